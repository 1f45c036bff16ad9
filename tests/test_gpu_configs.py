"""The measured configurations at their stated sizes, against the oracle.

* C0 as stated: glass icosphere (subdivision 4, 5,120 triangles) above a
  10 x 10 plane, point light at height 5, 256 x 256 camera, 1,048,576 light
  paths per iteration.  The reference's point light emits nothing
  (src/Lights.h:78-81): with compat_point_light = 1 the film stays black, as
  the reference renders it; with 0 the light emits isotropically and the
  photon lists (bit for bit) and the camera pass (2e-4) match the oracle.
* C1 at full size: the 1024 x 1024 water caustic with 16,777,216 light paths
  per iteration and a fixed merge radius, whose running image must converge
  to the FP64 oracle's image (tests/golden/water1024_fp64_s61.npz, made with
  another seed): after K iterations its RMSE^2 must be the noise the
  iterations' own variance predicts, V / K + the fixture's noise^2 -- a
  bias of about 1 % per pixel would add 1e-4 and fail.
* C4: one at-scale scene (2.2 M-triangle water heightfield, GPU-built BVH,
  four mixed lights -- point, quad, sphere and one more -- with the corrected
  light pdf, light paths traced in emission order) with its learned proposal
  trained on the GPU: guided photon lists bit-exact against the oracle given
  the GPU's proposal table, and the camera pass within 2e-4 (camera reduced
  to 128 x 128).
"""
import os

import numpy as np
import pytest

from realcausticsengine_amd import convergence, scenes
from realcausticsengine_amd.engine import Engine, photons_from_oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _compare_photons(ph_g, uid_g, ph_o):
    uid_o = ph_o[:, 10].astype(np.uint32)
    common, ig, io = np.intersect1d(uid_g, uid_o, return_indices=True)
    assert len(common) >= 0.999 * max(len(uid_g), len(uid_o)), (len(common), len(uid_g), len(uid_o))
    exact = np.all(ph_g[ig].view(np.uint32) == photons_from_oracle(ph_o[io]).view(np.uint32), axis=1)
    assert exact.mean() >= 0.999, exact.mean()


def test_c0_as_stated_reference_point_light_renders_black():
    args = scenes.c0_scene(256, 256, subdiv=4, parity_light=False, samples=16)
    eng = Engine(args, photons=1 << 20, seed=0, write_exr=0)  # compat_point_light = 1, the reference
    ph, _ = eng.trace_photons(0)
    eng.render_iteration()
    img = eng.image(scaled=True)
    eng.release()
    assert len(ph) == 0 and not img.any()


def test_c0_as_stated_isotropic_point_light(oracle):
    args = scenes.c0_scene(256, 256, subdiv=4, parity_light=False, samples=16)
    opts = dict(photons=1 << 20, seed=0, compat_point_light=0)
    eng = Engine(args, write_exr=0, **opts)
    ph_g, uid_g = eng.trace_photons(0)
    orc = oracle.OracleScene(args, threads=16, **opts)
    ph_o = orc.trace_light(0)
    assert len(ph_o) > 10000
    _compare_photons(ph_g, uid_g, ph_o)
    img_g = eng.camera_pass(0, photons_from_oracle(ph_o))
    eng.release()
    img_o = orc.camera_pass(0, ph_o)
    assert img_o.sum() > 0
    bad = np.abs(img_g - img_o) > 2e-4 * np.abs(img_o) + 1e-7
    assert bad.mean() <= 1e-4, bad.mean()


def test_c1_full_size_converges_to_the_fp64_reference():
    ref = convergence.SubsetReference.load(os.path.join(ROOT, "tests", "golden", "water1024_fp64_s61.npz"))
    args = scenes.water_scene(512, 1024, 1024, alpha=1.0, samples=1, base_radius=0.01)
    eng = Engine(args, photons=1 << 24, seed=0, write_exr=0, compat_rows=0)
    eng.prepare()
    K = 256
    prev = np.zeros((len(ref.rows), 3))
    deltas = []
    for _ in range(K):
        eng.render_iteration()
        acc = eng.image(scaled=False)[ref.rows, ref.cols].astype(np.float64)
        deltas.append(acc - prev)
        prev = acc
    full = eng.image(scaled=False)
    eng.release()
    rmse = ref.rmse(full, K)
    V = convergence.pixel_variance(deltas)
    expected = V / K + ref.noise ** 2
    print(f"C1 full size: RMSE {rmse:.5f} after {K} iterations, predicted {np.sqrt(expected):.5f} "
          f"(V {V:.4f}, fixture noise {ref.noise:.4f})", flush=True)
    assert 0.8 * expected <= rmse ** 2 <= 1.25 * expected


def test_c4_scene_with_mixed_lights_and_proposal(oracle):
    args, _ = scenes.c4_scene(1, res=128, photons=1 << 17)
    n = 1 << 17
    opts = dict(compat_point_light=0, compat_light_pdf=0)
    eng = Engine(args, photons=n, seed=4, write_exr=0, bvh_build=1, light_order=1, guide=1, guide_batch=65536,
                 guide_warmup=1, **opts)
    eng.prepare()
    for _ in range(3):  # the proposal trains on the GPU
        eng.render_iteration()
    assert eng.stats()["guide_rows"] > 0
    ph_g, uid_g = eng.trace_photons(3)
    cdf, _ = eng.guide_table()
    lights = [args.lights[k] for k in range(args.n_lights)]
    guidable = [k for k, L in enumerate(lights) if L.type == b"A" or (L.type == b"P" and L.radius > 0)]
    slot_of_light = [guidable.index(k) if k in guidable else -1 for k in range(len(lights))]
    assert len(cdf) == len(guidable) * 16 * 16
    orc = oracle.OracleScene(args, photons=n, seed=4, threads=16, **opts)
    orc.set_guide(cdf, slots=len(guidable), cells=16, slot_of_light=slot_of_light, alpha=0.2)
    ph_o = orc.trace_light(3)
    assert len(ph_o) > 5000
    _compare_photons(ph_g, uid_g, ph_o)
    orc.set_guide(None, 0, 0, [], 0.0)
    plain = orc.trace_light(3)
    img_o = orc.camera_pass(3, plain)
    img_g = eng.camera_pass(3, photons_from_oracle(plain))
    eng.release()
    assert img_o.sum() > 0
    bad = np.abs(img_g - img_o) > 2e-4 * np.abs(img_o) + 1e-7
    assert bad.mean() <= 1e-3, bad.mean()
