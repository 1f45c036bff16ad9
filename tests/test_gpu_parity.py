"""Parity of the gfx950 engine with the CPU oracle (run on the GPU box).

Bars: first-hit triangle ids equal except near-ties (|dt| <= 1e-5 t);
photon lists of the plain estimator bit-exact per path; gather sets
bit-exact; images within a per-pixel relative tolerance of the oracle
(summation order of the merge differs, 2e-4 relative + 1e-7 absolute)."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine, photons_from_oracle

pytestmark = pytest.mark.gpu


def _rays(args, n, seed):
    rng = np.random.default_rng(seed)
    cam = np.asarray(args.camera.position, dtype=np.float64)
    c = np.asarray(args.camera.corners, dtype=np.float64)
    # half camera rays, half rays from random points in random directions
    m = n // 2
    s = rng.random((m, 2))
    pts = c[2] + s[:, :1] * (c[1] - c[2]) + s[:, 1:] * (c[3] - c[2])
    d1 = pts - cam
    o2 = rng.uniform(-3, 3, (n - m, 3)) + np.asarray([0, 0, 1.2])
    d2 = rng.normal(size=(n - m, 3))
    o = np.concatenate([np.tile(cam, (m, 1)), o2])
    d = np.concatenate([d1, d2])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def test_first_hit_matches_brute_force(oracle):
    args = scenes.c0_scene(64, 64, subdiv=3)
    eng = Engine(args)
    orc = oracle.OracleScene(args)
    rays = _rays(args, 20000, 1)
    tri_g, t_g = eng.first_hit(rays)
    tri_o, t_o = orc.first_hit(rays, brute=True)
    hit = tri_o >= 0
    assert np.array_equal(tri_g >= 0, hit)
    same = tri_g == tri_o
    # identical triangle -> identical distance bits (same Moller-Trumbore arithmetic)
    assert np.array_equal(t_g[same & hit], t_o[same & hit].astype(np.float32))
    tie = ~same & hit
    assert np.all(np.abs(t_g[tie] - t_o[tie]) <= 1e-5 * t_o[tie])
    assert tie.sum() <= 0.001 * hit.sum()
    eng.release()


def test_photon_list_bit_exact(oracle):
    args = scenes.c0_scene(64, 64, subdiv=3)
    eng = Engine(args, photons=32768, seed=7)
    orc = oracle.OracleScene(args, photons=32768, seed=7)
    ph_g, uid_g = eng.trace_photons(0)
    ph_o = orc.trace_light(0)
    uid_o = ph_o[:, 10].astype(np.uint32)
    assert len(ph_o) > 1000
    common, ig, io = np.intersect1d(uid_g, uid_o, return_indices=True)
    # paths that diverge (closest-hit ties) are allowed only rarely
    assert len(common) >= 0.999 * max(len(uid_g), len(uid_o))
    a = ph_g[ig]
    b = photons_from_oracle(ph_o[io])
    exact = np.all(a.view(np.uint32) == b.view(np.uint32), axis=1)
    assert exact.mean() >= 0.999, exact.mean()
    eng.release()


def test_gather_sets_bit_exact(oracle):
    rng = np.random.default_rng(3)
    args = scenes.c0_scene(16, 16, subdiv=1)
    eng = Engine(args)
    P, Q = 60000, 3000
    pos = rng.uniform(-1, 1, (P, 3)).astype(np.float32)
    pos[: P // 2, 2] = 0.0  # a planar layer, like photons on a floor
    q = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    q[: Q // 2, 2] = 0.0
    r2 = np.float32(0.05) ** 2
    off_g, idx_g = eng.gather_sets(pos, q, r2)
    off_o, idx_o = oracle.gather_sets(pos, q, r2)
    assert np.array_equal(off_g, off_o)
    for i in range(Q):
        a = np.sort(idx_g[off_g[i]:off_g[i + 1]])
        b = np.sort(idx_o[off_o[i]:off_o[i + 1]])
        assert np.array_equal(a, b)
    eng.release()


@pytest.mark.parametrize("exact,group", [(0, 8), (1, 8), (0, 1), (1, 1)])
def test_camera_pass_given_photons(oracle, exact, group):
    # exact = 1: correctly rounded merge arithmetic; 0: hardware rcp/sqrt (<= 2 ulp);
    # group 8: the grouped merge (the default), 1: one query per lane
    args = scenes.c0_scene(48, 48, subdiv=3)
    eng = Engine(args, photons=20000, seed=3, exact_merge=exact, group_queries=group)
    orc = oracle.OracleScene(args, photons=20000, seed=3)
    ph = orc.trace_light(0)
    img_o = orc.camera_pass(0, ph)
    img_g = eng.camera_pass(0, photons_from_oracle(ph))
    tol = 2e-4 * np.abs(img_o) + 1e-7
    assert np.all(np.abs(img_g - img_o) <= tol)
    assert img_o.sum() > 0
    eng.release()


def test_render_matches_oracle(oracle):
    args = scenes.c0_scene(32, 32, subdiv=3, samples=2)
    eng = Engine(args, photons=16384, seed=11)
    orc = oracle.OracleScene(args, photons=16384, seed=11)
    eng.render_iteration()
    eng.render_iteration()
    img_g = eng.image(scaled=True)
    img_o = orc.render(2)
    tol = 2e-4 * np.abs(img_o) + 1e-7
    bad = np.abs(img_g - img_o) > tol
    assert bad.mean() <= 0.002, bad.mean()
    assert img_o.sum() > 0
    eng.release()
