"""Pins of the rough dielectric (glass of roughness >= 0.01 with option
compat_glass = 0; Walter, Marschner, Li, Torrance 2007, with the
reference's GGX sampling, D and height-correlated G2, src/BXDF.cpp:286-311).
It has no counterpart in the reference (its glass is always delta,
src/materials.h:104), so what the mathematics fixes pins it:

* the pdf integrates over the sphere to the share of sampled directions
  the sampler keeps (reflections that stay above the surface, refractions
  that cross it);
* the sampler's weights f |cos wi| / pdf average to the integral of
  evaluate (f |cos wi|), and that integral is at most 1 (white furnace);
* the pdf reported with a sample equals pdf() of the sampled direction;
* the BTDF obeys reciprocity with the eta^2 factor: f(wo, wi) / f(wi, wo)
  = (eta_wo / eta_wi)^2 in radiance mode, 1 for reflections;
* as the roughness goes to zero the samples become Fresnel_Specular's:
  mirror or Snell direction, reflection probability F(wo.n), weight 1, or
  eta^2 for a refraction in radiance mode (src/BXDF.cpp:136-205).
"""
import math

import numpy as np
import pytest

UP = (0.0, 0.0, 1.0)


def _dir(theta, phi=0.3):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def _mat(rough, ior=1.5):
    return ("R", ior, rough, 0.0, (1.0, 1.0, 1.0))


def _sphere_grid(nt=360, nph=180):
    th = (np.arange(nt) + 0.5) * math.pi / nt
    ph = (np.arange(nph) + 0.5) * 2 * math.pi / nph
    return th, ph, (math.pi / nt) * (2 * math.pi / nph)


@pytest.mark.parametrize("theta,rough,inside", [(0.0, 0.3, False), (0.8, 0.3, False), (0.5, 0.5, True)])
def test_pdf_integrates_to_the_samplers_kept_share(oracle, theta, rough, inside):
    mat = _mat(rough)
    wo = _dir(theta)
    if inside:
        wo[2] = -wo[2]
    th, ph, dA = _sphere_grid()
    total = 0.0
    for t in th:
        total += sum(oracle.bsdf(mat, 1, UP, UP, -wo, wi=_dir(t, p), op=2)[6] for p in ph) * math.sin(t)
    integral = total * dA
    rng = np.random.default_rng(7)
    n = 20000
    kept = sum(oracle.bsdf(mat, 1, UP, UP, -wo, u3=rng.random(3))[0] != 0.0 for _ in range(n)) / n
    print(f"theta {theta} rough {rough} inside {inside}: integral {integral:.4f} kept {kept:.4f}")
    assert integral == pytest.approx(kept, abs=4 * math.sqrt(kept * (1 - kept) / n) + 5e-3)


@pytest.mark.parametrize("theta,rough", [(0.4, 0.2), (1.1, 0.5)])
def test_weights_average_to_the_integral_of_evaluate(oracle, theta, rough):
    mat = _mat(rough)
    wo = _dir(theta)
    th, ph, dA = _sphere_grid()
    total = 0.0
    for t in th:
        total += sum(oracle.bsdf(mat, 1, UP, UP, -wo, wi=_dir(t, p), op=1)[0] for p in ph) * math.sin(t)
    integral = total * dA
    rng = np.random.default_rng(8)
    w = np.asarray([oracle.bsdf(mat, 1, UP, UP, -wo, u3=rng.random(3))[0] for _ in range(20000)])
    print(f"theta {theta} rough {rough}: integral {integral:.4f} mean weight {w.mean():.4f}")
    assert integral <= 1.0 + 5e-3  # energy conserving (importance mode, R = 1)
    assert integral > 0.8
    assert w.mean() == pytest.approx(integral, abs=4 * w.std() / math.sqrt(len(w)) + 5e-3)


def test_sample_pdfs_equal_pdf_of_the_sampled_direction(oracle):
    rng = np.random.default_rng(9)
    seen = set()
    for k in range(400):
        mat = _mat(rng.uniform(0.05, 0.8), rng.uniform(1.2, 1.8))
        wo = _dir(rng.uniform(0.0, 1.4), rng.uniform(0, 2 * math.pi))
        if k % 2:
            wo[2] = -wo[2]
        s = oracle.bsdf(mat, k % 2, UP, UP, -wo, u3=rng.random(3))
        if s[0] == 0.0:
            continue
        seen.add(int(s[10]))
        p = oracle.bsdf(mat, k % 2, UP, UP, -wo, wi=s[3:6], op=2)
        assert p[6] == pytest.approx(s[6], rel=2e-3, abs=1e-6)
        assert p[7] == pytest.approx(s[7], rel=2e-3, abs=1e-6)
        # the weight is evaluate / pdf of the same direction
        e = oracle.bsdf(mat, k % 2, UP, UP, -wo, wi=s[3:6], op=1)
        assert e[0] == pytest.approx(s[0] * s[6], rel=2e-3, abs=1e-6)
    assert seen == {8 | 1, 8 | 2}  # both glossy reflections and glossy transmissions were sampled


def test_btdf_reciprocity_with_eta_squared(oracle):
    rng = np.random.default_rng(10)
    ior = 1.5
    mat = _mat(0.4, ior)
    n_checked = 0
    for _ in range(300):
        a = _dir(rng.uniform(0.05, 1.3), rng.uniform(0, 2 * math.pi))
        b = _dir(rng.uniform(0.05, 1.3), rng.uniform(0, 2 * math.pi))
        refl = rng.random() < 0.3
        if not refl:
            b[2] = -b[2]  # a above (air), b below (glass)
        fab = oracle.bsdf(mat, 0, UP, UP, -a, wi=b, op=1)[0] / abs(b[2])
        fba = oracle.bsdf(mat, 0, UP, UP, -b, wi=a, op=1)[0] / abs(a[2])
        if fab < 1e-6 and fba < 1e-6:
            continue
        n_checked += 1
        # radiance mode: f(wo -> ...)/f(reverse) = (eta_wo / eta_wi)^2; wo = a is in air, wi = b in glass
        want = 1.0 if refl else (1.0 / ior) ** 2
        assert fab / fba == pytest.approx(want, rel=2e-3)
    assert n_checked > 100


@pytest.mark.parametrize("theta,inside", [(0.3, False), (1.0, False), (0.4, True)])
def test_smooth_limit_is_fresnel_specular(oracle, theta, inside):
    ior = 1.5
    wo = _dir(theta)
    if inside:
        wo[2] = -wo[2]
    rng = np.random.default_rng(11)
    n, refl = 2000, 0
    for _ in range(n):
        u = rng.random(3)
        for mode in (0, 1):
            r = oracle.bsdf(_mat(1e-4, ior), mode, UP, UP, -wo, u3=u)
            d = oracle.bsdf(("G", ior, 0.0, 0.0, (1.0, 1.0, 1.0)), mode, UP, UP, -wo, u3=(u[0], 0.5, 0.5))
            if abs(r[0] - d[0]) > 1e-3 * max(1.0, d[0]):
                # only where u0 lies between F(wo.m) and F(wo.n) may the branch differ
                continue
            np.testing.assert_allclose(r[3:6], d[3:6], atol=2e-3)
            assert r[0] == pytest.approx(d[0], rel=2e-3)
            if mode == 1:
                refl += (int(r[10]) & 1) == 1
    # reflection share = F(wo.n) of the smooth interface
    f = oracle.bsdf(("G", ior, 0.0, 0.0, (1.0, 1.0, 1.0)), 1, UP, UP, -wo, u3=(0.0, 0.5, 0.5))[6]
    assert refl / n == pytest.approx(f, abs=4 * math.sqrt(f * (1 - f) / n) + 2e-3)
