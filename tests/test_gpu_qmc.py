"""Stratified camera samples (option camera_qmc) on the GPU.

The camera pass with camera_qmc = 1 against the oracle's restatement on the
same photon list (the merge tolerance of the plain parity test), at two
iterations (the sample index is the iteration); the light pass is untouched;
and an A/B check of what the option is for: over many iterations of a fixed
merge radius the running image approaches the converged image faster than
with independent draws, while the two long-run means agree."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine, photons_from_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("iteration", [0, 5])
def test_qmc_camera_pass_matches_oracle(oracle, iteration):
    args = scenes.water_scene(64, 48, 48, samples=1)
    eng = Engine(args, photons=40000, seed=3, camera_qmc=1)
    orc = oracle.OracleScene(args, photons=40000, seed=3, camera_qmc=1)
    ph = orc.trace_light(iteration)
    img_o = orc.camera_pass(iteration, ph)
    img_g = eng.camera_pass(iteration, photons_from_oracle(ph))
    tol = 2e-4 * np.abs(img_o) + 1e-7
    assert img_o.sum() > 0
    assert np.all(np.abs(img_g - img_o) <= tol)
    eng.release()


def _running(qmc, iters, seed):
    args = scenes.water_scene(128, 64, 64, samples=iters, alpha=1.0)
    eng = Engine(args, photons=262144, seed=seed, write_exr=0, compat_rows=0, camera_qmc=qmc)
    eng.prepare()
    out = []
    for i in range(iters):
        eng.render_iteration()
        if i + 1 in (16, 64, iters):
            out.append(eng.image(scaled=True).astype(np.float64))
    eng.release()
    return out


def test_qmc_converges_faster_with_the_same_mean():
    iters = 256
    ref = 0.5 * (_running(0, iters, 101)[-1] + _running(1, iters, 102)[-1])  # a long-run estimate of the mean
    m = ref.mean()
    errs = {}
    for qmc in (0, 1):
        run = _running(qmc, 64, 7)
        errs[qmc] = [float(np.sqrt(((img - ref) ** 2).mean()) / m) for img in run]
    print("relative RMSE at 16 / 64 iterations: independent", np.round(errs[0], 4), "stratified",
          np.round(errs[1], 4), flush=True)
    assert errs[1][0] < errs[0][0] and errs[1][1] < errs[0][1]
