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
    # radiance mode scales by eta^2 entering the denser medium
    trr = oracle.bsdf(("G", ior, 0, 0, (1, 1, 1)), 0, UP, UP, d, u3=(0.999, 0.5, 0.5))
    trr_rad = oracle.bsdf(("G", ior, 0, 0, (1, 1, 1)), 0, UP, UP, d, u3=(0.999, 0.5, 0.5))
    assert trr[0] == trr_rad[0]


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


def ggx_D(c, a2):
    den = c * c * (a2 - 1) + 1
    return a2 / (math.pi * den * den)


@pytest.mark.parametrize("a2", [0.0001, 0.01, 0.1, 0.5, 1.0])
def test_ggx_normalisation_closed_form(a2):
    # integral over the hemisphere of D(h) cos(h) = 1 (substitution u = cos^2)
    th = np.linspace(0, math.pi / 2, 400001)
    c = np.cos(th)
    den = c * c * (a2 - 1) + 1
    f = a2 / (math.pi * den * den) * c * np.sin(th) * 2 * math.pi
    assert np.trapezoid(f, th) == pytest.approx(1.0, rel=2e-3)


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


def test_ggx_g2_bounds_and_symmetry():
    def G2(nw, nl, a2):
        A = nw * math.sqrt(a2 + (1 - a2) * nl * nl)
        B = nl * math.sqrt(a2 + (1 - a2) * nw * nw)
        return 2 * nl * nw / (A + B)
    for a2 in (0.01, 0.3, 1.0):
        for x in (0.1, 0.5, 0.9):
            for y in (0.2, 0.7, 1.0):
                g = G2(x, y, a2)
                assert 0 <= g <= 1 + 1e-12
                assert g == pytest.approx(G2(y, x, a2))
    assert G2(1.0, 1.0, 0.5) == pytest.approx(1.0)
