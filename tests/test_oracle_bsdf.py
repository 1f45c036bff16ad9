"""Closed-form pins of the oracle BSDFs (materials.h, BXDF.cpp)."""
import math

import numpy as np
import pytest

UP = (0.0, 0.0, 1.0)
DOWN = (0.0, 0.0, -1.0)


def _dir(theta, phi=0.3):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def fresnel_exact(cos_i, n):
    """Unpolarised dielectric Fresnel reflectance from Snell + Fresnel equations."""
    sin_t = math.sqrt(max(0.0, 1 - cos_i * cos_i)) / n
    if sin_t >= 1:
        return 1.0
    cos_t = math.sqrt(1 - sin_t * sin_t)
    rs = (cos_i - n * cos_t) / (cos_i + n * cos_t)
    rp = (n * cos_i - cos_t) / (n * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("ior", [1.33, 1.5, 1.7])
def test_glass_normal_incidence_fresnel(oracle, ior):
    out = oracle.bsdf(("G", ior, 0.0, 0.0, (1, 1, 1)), 1, UP, UP, DOWN, u3=(0.0, 0.5, 0.5))
    assert out[10] == 1 + 8 + 16  # reflection | glossy | specular
    assert out[6] == pytest.approx(((ior - 1) / (ior + 1)) ** 2, rel=1e-5)
    assert out[12] == 1  # delta


@pytest.mark.parametrize("theta", [0.2, 0.6, 1.0, 1.3])
def test_glass_fresnel_and_snell(oracle, theta):
    ior = 1.5
    wo = _dir(theta)
    d = -wo
    refl = oracle.bsdf(("G", ior, 0, 0, (1, 1, 1)), 0, UP, UP, d, u3=(0.0, 0.5, 0.5))
    F = refl[6]
    assert F == pytest.approx(fresnel_exact(math.cos(theta), ior), rel=2e-5)
    np.testing.assert_allclose(refl[3:6], [-wo[0], -wo[1], wo[2]], atol=2e-6)
    tr = oracle.bsdf(("G", ior, 0, 0, (1, 1, 1)), 0, UP, UP, d, u3=(0.999, 0.5, 0.5))
    assert tr[10] == 2 + 8 + 16  # transmission
    wi = tr[3:6]
    assert wi[2] < 0
    sin_t = math.hypot(wi[0], wi[1])
    assert sin_t == pytest.approx(math.sin(theta) / ior, rel=1e-5)
    assert tr[6] == pytest.approx(1 - F, rel=1e-5)


def test_glass_total_internal_reflection(oracle):
    ior = 1.5
    theta = math.asin(1 / ior) + 0.1  # beyond the critical angle, from inside
    wo_inside = _dir(theta)
    wo_inside[2] = -wo_inside[2]  # below the surface (inside the glass)
    d = -wo_inside
    out = oracle.bsdf(("G", ior, 0, 0, (1, 1, 1)), 1, UP, UP, d, u3=(0.9999, 0.5, 0.5))
    assert out[6] == 1.0 and out[10] == 1 + 8 + 16


def test_radiance_scaling_of_refraction(oracle):
    ior = 1.5
    d = -_dir(0.4)
    rad = oracle.bsdf(("G", ior, 0, 0, (1, 1, 1)), 0, UP, UP, d, u3=(0.999, 0.5, 0.5))
    imp = oracle.bsdf(("G", ior, 0, 0, (1, 1, 1)), 1, UP, UP, d, u3=(0.999, 0.5, 0.5))
    assert imp[0] == 1.0
    assert rad[0] == pytest.approx((1 / ior) ** 2, rel=1e-6)


def _theta_quadrature(n=20000):
    th = (np.arange(n) + 0.5) * (math.pi / 2) / n
    return th, (math.pi / 2) / n


@pytest.mark.parametrize("kind,rough,spec", [("M", 0.2, 0.0), ("M", 0.4, 0.0), ("M", 0.7, 0.0),
                                             ("P", 0.6, 0.5), ("P", 0.8, 1.0)])
def test_ggx_pdf_integrates_to_its_closed_form(oracle, kind, rough, spec):
    """At normal incidence a GGX half vector within 45 degrees of the normal
    reflects into the upper hemisphere, and P(theta_h < 45 deg) = 1 / (1 + a2)
    for the GGX distribution (a2 = roughness^2 for metal, Metal_Material ->
    Microfacet_Reflection_GGX, BXDF.cpp:274-367; roughness^4 for plastic,
    Plastic_Material squares it once more, materials.h:62).  So the quadrature
    of the oracle's forward pdf over the upper hemisphere is 1 / (1 + a2) for
    metal and (1 - F0) + F0 / (1 + a2) for the plastic's mixture of the
    cosine lobe and the GGX lobe (F0 = 0.08 specular, BXDF.cpp:373-548)."""
    a2 = rough ** 2 if kind == "M" else rough ** 4
    mat = (kind, 1.5, rough, spec, (1.0, 1.0, 1.0))
    th, dth = _theta_quadrature()
    pdf = np.asarray([oracle.bsdf(mat, 1, UP, UP, DOWN, wi=_dir(t, 0.7), op=2)[6] for t in th])
    integral = float(np.sum(pdf * np.sin(th)) * 2 * math.pi * dth)
    f0 = 0.08 * spec
    want = 1 / (1 + a2) if kind == "M" else (1 - f0) + f0 / (1 + a2)
    assert integral == pytest.approx(want, rel=2e-3)


def test_ggx_pdf_integral_equals_the_samplers_valid_share(oracle):
    """At oblique incidence the mass of the pdf over the upper hemisphere is
    the share of sampled directions that stay above it (the sampler returns
    zero for the rest): quadrature of op 2 against the sampler's own count."""
    mat = ("M", 1.0, 0.5, 0.0, (1.0, 1.0, 1.0))
    wo = _dir(1.0)
    nt, nph = 240, 240
    th = (np.arange(nt) + 0.5) * (math.pi / 2) / nt
    ph = (np.arange(nph) + 0.5) * (2 * math.pi) / nph
    total = 0.0
    for t in th:
        row = sum(oracle.bsdf(mat, 1, UP, UP, -wo, wi=_dir(t, p), op=2)[6] for p in ph)
        total += row * math.sin(t)
    integral = total * (math.pi / 2 / nt) * (2 * math.pi / nph)
    rng = np.random.default_rng(5)
    n = 20000
    valid = sum(oracle.bsdf(mat, 1, UP, UP, -wo, u3=rng.random(3))[0] != 0.0 for _ in range(n))
    share = valid / n
    assert integral == pytest.approx(share, abs=4 * math.sqrt(share * (1 - share) / n) + 3e-3)
    assert share < 0.97  # the case is one where directions do leave the hemisphere


def test_ggx_sample_weights_average_to_the_integral_of_evaluate(oracle):
    """E[f cos / pdf] over the sampler equals the integral of f cos over the
    hemisphere: the sampler's weights (BXDF.cpp:302-337) and evaluate
    (:339-353) describe the same lobe (white furnace of the metal, R = 1)."""
    mat = ("M", 1.0, 0.45, 0.0, (1.0, 1.0, 1.0))
    wo = _dir(0.6)
    nt, nph = 240, 240
    th = (np.arange(nt) + 0.5) * (math.pi / 2) / nt
    ph = (np.arange(nph) + 0.5) * (2 * math.pi) / nph
    total = 0.0
    for t in th:
        total += sum(oracle.bsdf(mat, 1, UP, UP, -wo, wi=_dir(t, p), op=1)[0] for p in ph) * math.sin(t)
    integral = total * (math.pi / 2 / nt) * (2 * math.pi / nph)
    rng = np.random.default_rng(6)
    w = np.asarray([oracle.bsdf(mat, 1, UP, UP, -wo, u3=rng.random(3))[0] for _ in range(20000)])
    assert integral <= 1.0 + 1e-3  # energy conserving
    assert w.mean() == pytest.approx(integral, abs=4 * w.std() / math.sqrt(len(w)) + 2e-3)


def test_metal_mirror_and_reciprocity(oracle):
    wo = _dir(0.5)
    mirror = oracle.bsdf(("M", 1.0, 0.0, 0.0, (0.9, 0.8, 0.7)), 1, UP, UP, -wo, u3=(0.1, 0.2, 0.3))
    np.testing.assert_allclose(mirror[3:6], [-wo[0], -wo[1], wo[2]], atol=2e-6)
    np.testing.assert_allclose(mirror[0:3], [0.9, 0.8, 0.7], rtol=1e-6)
    assert mirror[12] == 1
    # rough metal: f = eval / cos(wi) is symmetric in (wo, wi)
    mat = ("M", 1.0, 0.4, 0.0, (0.9, 0.9, 0.9))
    for a, b in [(0.3, 0.7), (0.1, 1.2), (0.9, 0.2)]:
        w1, w2 = _dir(a, 0.2), _dir(b, 2.0)
        e12 = oracle.bsdf(mat, 1, UP, UP, -w1, wi=w2, op=1)[0] / w2[2]
        e21 = oracle.bsdf(mat, 1, UP, UP, -w2, wi=w1, op=1)[0] / w1[2]
        assert e12 == pytest.approx(e21, rel=2e-5)


def test_rough_metal_sample_pdf_consistent(oracle):
    mat = ("M", 1.0, 0.5, 0.0, (1, 1, 1))
    wo = _dir(0.7)
    rng = np.random.default_rng(0)
    for _ in range(50):
        u = rng.random(3)
        s = oracle.bsdf(mat, 1, UP, UP, -wo, u3=u)
        if s[0] == 0:
            continue
        p = oracle.bsdf(mat, 1, UP, UP, -wo, wi=s[3:6], op=2)
        assert p[6] == pytest.approx(s[6], rel=1e-4)


def test_diffuse_plastic_is_cosine_sampled(oracle):
    # roughness 1, specular 0: Fresnel 0, so every sample is the diffuse lobe
    mat = ("P", 1.5, 1.0, 0.0, (0.8, 0.8, 0.8))
    wo = _dir(0.4)
    rng = np.random.default_rng(1)
    cz = []
    for _ in range(4000):
        s = oracle.bsdf(mat, 1, UP, UP, -wo, u3=rng.random(3))
        assert s[10] == 4
        np.testing.assert_allclose(s[0:3], 0.8, rtol=1e-6)
        assert s[6] == pytest.approx(s[5] / math.pi, rel=1e-5)  # pdf = cos / pi
        cz.append(s[5])
    cz = np.array(cz)
    # cos(theta) of a cosine-weighted hemisphere: P(cos <= c) = c^2
    for c in (0.25, 0.5, 0.75):
        assert (cz <= c).mean() == pytest.approx(c * c, abs=0.03)
    ev = oracle.bsdf(mat, 1, UP, UP, -wo, wi=_dir(0.9), op=1)
    assert ev[0] == pytest.approx(0.8 / math.pi, rel=1e-6)


def test_smooth_plastic_fresnel_reflectance(oracle):
    # specular s maps to F0 = 0.08 s (materials.h:62); at normal incidence the
    # mirror lobe is chosen with probability F0
    s = 0.5
    out = oracle.bsdf(("P", 1.5, 0.0, s, (0.5, 0.5, 0.5)), 1, UP, UP, DOWN, u3=(0.0, 0.5, 0.5))
    assert out[10] == 16 + 1 + 8
    assert out[6] == pytest.approx(0.08 * s, rel=1e-4)
    assert out[12] == 0  # not delta: photons are stored on it


def test_ggx_masking_from_evaluate_and_pdf(oracle):
    """The masking term recovered from the oracle's metal lobe (R = 1):
    evaluate / pdf = G2 |wo.h| / (|wo.z| |h.z|) (BXDF.cpp:339-367), with h the
    half vector of (wo, wi) computed here.  G2 lies in [0, 1], is symmetric in
    wo and wi, is 1 at normal incidence both ways and tends to 1 as the
    roughness vanishes."""
    def g2(mat, wo, wi):
        e = oracle.bsdf(mat, 1, UP, UP, -wo, wi=wi, op=1)[0]
        p = oracle.bsdf(mat, 1, UP, UP, -wo, wi=wi, op=2)[6]
        h = (wo + wi) / np.linalg.norm(wo + wi)
        return e / p * abs(wo[2]) * abs(h[2]) / abs(wo @ h)
    for rough in (0.1, 0.5, 1.0):
        mat = ("M", 1.0, rough, 0.0, (1.0, 1.0, 1.0))
        for a, b in [(0.2, 0.4), (0.7, 1.2), (1.3, 0.1)]:
            w1, w2 = _dir(a, 0.3), _dir(b, 2.1)
            x, y = g2(mat, w1, w2), g2(mat, w2, w1)
            assert 0.0 <= x <= 1.0 + 1e-5
            assert x == pytest.approx(y, rel=1e-4)
        assert g2(mat, _dir(0.0), _dir(0.0)) == pytest.approx(1.0, rel=1e-5)
    smooth = ("M", 1.0, 0.02, 0.0, (1.0, 1.0, 1.0))
    assert g2(smooth, _dir(0.5, 0.0), _dir(0.6, math.pi)) == pytest.approx(1.0, abs=2e-3)
    assert g2(("M", 1.0, 1.0, 0.0, (1.0, 1.0, 1.0)), _dir(1.4, 0.0), _dir(1.4, math.pi)) < 0.5
