"""The autotuner (rc_autotune, rc_batch_autotune) on the GPU.

It must leave the film and the iteration count alone, pick values from its
candidate lists, and the tuned launch shapes must render the same image: the
photon list and the merge sets do not depend on them; only the summation
order inside a merge changes with the cell size and the merge group size
(relative 1e-5)."""
import numpy as np
import pytest

from realcausticsengine_amd import convergence, scenes
from realcausticsengine_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu

CANDIDATES = {"path_chunk": {64, 128, 256, 512, 1024}, "light_blocks_per_cu": {4, 6, 8, 12, 16},
              "camera_blocks_per_cu": {4, 6, 8, 12, 16}, "gather_blocks_per_cu": {5, 8, 12, 16, 24},
              "light_refill": {1, 8, 16, 32}, "light_order": {0, 1}, "camera_refill": {1, 8, 16, 32},
              "query_order": {0, 1},
              "group_queries": {1, 4, 8, 16}, "gather_tile": {0, 256, 1024},
              "gather_block": {0, 1, 2, 4}, "overlap": {0, 1, 2}}


def _check_report(r):
    print("autotune", r, flush=True)
    for k, c in CANDIDATES.items():
        assert int(r[k]) in c, k
    assert min(abs(r["cell_scale"] - v) for v in (0.67, 0.8, 1.0, 1.25)) < 1e-6
    assert 0 < r["ms_after"] <= 1.5 * r["ms_before"]


def test_autotune_keeps_film_and_image():
    args = scenes.water_scene(64, 96, 64, samples=3)
    ref = Engine(args, photons=262144, seed=2, write_exr=0)
    ref.prepare()
    for _ in range(2):
        ref.render_iteration()
    want = ref.image(scaled=False)
    ref.release()

    eng = Engine(args, photons=262144, seed=2, write_exr=0)
    eng.prepare()
    eng.render_iteration()
    before = eng.image(scaled=False)
    r = eng.autotune(repeats=1)
    _check_report(r)
    assert eng.get("iterations") == 1
    assert np.array_equal(eng.image(scaled=False), before)
    for k in CANDIDATES:
        assert eng.get(k) == r[k]
    eng.render_iteration()
    got = eng.image(scaled=False)
    eng.release()
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7 * want.max())


def test_online_tuner_renders_the_same_iterations():
    """convergence.OnlineTuner: every candidate renders a real iteration of the run, so after the tuning
    iterations the image is the image of as many iterations with fixed launch shapes (to the merge's
    summation order)."""
    args = scenes.water_scene(64, 96, 64, samples=3)
    n = sum(len(c) for _, c, _ in convergence.ONLINE_KNOBS) + 2
    ref = Engine(args, photons=65536, seed=2, write_exr=0)
    ref.prepare()
    for _ in range(n):
        ref.render_iteration()
    want = ref.image(scaled=False)
    ref.release()

    eng = Engine(args, photons=65536, seed=2, write_exr=0)
    eng.prepare()
    tuner = convergence.OnlineTuner(eng)
    for _ in range(n):
        tuner.before_iteration()
        eng.render_iteration()
        if not tuner.done:
            tuner.after_iteration(eng.stats())
    r = tuner.report()
    print("online", r, flush=True)
    assert tuner.done and r["tuning_iterations"] == n - 2
    assert eng.get("iterations") == n
    for k, c in CANDIDATES.items():
        assert int(r[k]) in c and eng.get(k) == r[k], k
    got = eng.image(scaled=False)
    eng.release()
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7 * want.max())


def test_batch_autotune_keeps_films():
    cases = [(scenes.water_scene(32, 24, 16, samples=2), dict(photons=20000, seed=3)),
             (scenes.c0_scene(32, 32, subdiv=2, samples=2), dict(photons=50000, seed=5))]
    b = Batch([a for a, _ in cases], [dict(o, write_exr=0) for _, o in cases]).prepare()
    b.render_iteration()
    before = [b.image(k, scaled=False) for k in range(2)]
    r = b.autotune(repeats=1)
    _check_report(r)
    for k in range(2):
        assert np.array_equal(b.image(k, scaled=False), before[k])
    assert b.render_iteration() == 1
    b.release()


def test_tuned_render_matches_the_oracle(oracle):
    # the tuned launch shapes against the oracle's render of the same two iterations
    args = scenes.c0_scene(32, 32, subdiv=3, samples=2)
    eng = Engine(args, photons=16384, seed=11, write_exr=0)
    eng.prepare()
    eng.render_iteration()
    _check_report(eng.autotune(repeats=1))
    eng.render_iteration()
    img_g = eng.image(scaled=True)
    eng.release()
    img_o = oracle.OracleScene(args, photons=16384, seed=11).render(2)
    bad = np.abs(img_g - img_o) > 2e-4 * np.abs(img_o) + 1e-7
    assert img_o.sum() > 0 and bad.mean() <= 0.002, bad.mean()
