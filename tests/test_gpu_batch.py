"""Scenes rendered together (rc_batch_*) against the same scenes rendered alone.

Each scene of a batch keeps its own Philox key, iteration index, photon-map
grid and film, its photons sort into its own bucket range in the same
(bucket, light path) order as alone (a path's photons in bounce order), and
every pixel receives one addition per iteration: the batched film equals the
film of the scene rendered alone, bit for bit.  A launch holds up to 2^31
light paths (photon ids are 32-bit light path indices)."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu


def _cases():
    return [
        (scenes.water_scene(32, 24, 16, samples=2), dict(photons=20000, seed=3)),
        (scenes.c0_scene(32, 32, subdiv=2, samples=2), dict(photons=50000, seed=5, max_path_length=4)),
        (scenes.random_mesh_scene(seed=1, width=16, height=24), dict(photons=30011, seed=7, compat_rows=0)),
        (scenes.slab_scene(20, 20, samples=2), dict(photons=4096, seed=11)),
    ]


def _alone(args, opts, iters):
    eng = Engine(args, **dict(opts, write_exr=0))
    eng.prepare()
    for _ in range(iters):
        eng.render_iteration()
    img = eng.image(scaled=False)
    eng.release()
    return img


def test_batch_matches_scenes_rendered_alone():
    cases = _cases()
    iters = 2
    alone = [_alone(a, o, iters) for a, o in cases]
    b = Batch([a for a, _ in cases], [dict(o, write_exr=0) for _, o in cases]).prepare()
    for i in range(iters):
        assert b.render_iteration() == i
    st = b.stats()
    assert st["light_paths"] == sum(o["photons"] for _, o in cases)
    assert st["camera_paths"] == sum(a.camera.width * a.camera.height for a, _ in cases)
    for k, (a, _) in enumerate(cases):
        img = b.image(k, scaled=False)
        print(f"scene {k}: sum {img.sum():.6g} alone {alone[k].sum():.6g}", flush=True)
        assert img.sum() > 0
        assert np.array_equal(img, alone[k]), k
    b.release()


def test_batch_beyond_2_29_light_paths_matches_scenes_alone():
    # 2^28 + k light paths per scene: the second scene's photon ids (global light path indices) pass 2^29
    cases = [(scenes.c0_scene(16, 16, subdiv=2, samples=1), dict(photons=(1 << 28) + 12345, seed=21)),
             (scenes.c0_scene(24, 16, subdiv=3, samples=1), dict(photons=(1 << 28) + 777, seed=22))]
    alone = [_alone(a, o, 1) for a, o in cases]
    b = Batch([a for a, _ in cases], [dict(o, write_exr=0) for _, o in cases]).prepare()
    b.render_iteration()
    st = b.stats()
    assert st["light_paths"] == sum(o["photons"] for _, o in cases) > (1 << 29)
    for k in range(len(cases)):
        img = b.image(k, scaled=False)
        print(f"scene {k}: sum {img.sum():.6g} alone {alone[k].sum():.6g}", flush=True)
        assert img.sum() > 0
        assert np.array_equal(img, alone[k]), k
    b.release()


def test_batch_scene_order_does_not_matter():
    cases = _cases()[:3]
    b1 = Batch([a for a, _ in cases], [dict(o, write_exr=0) for _, o in cases]).prepare()
    b1.render_iteration()
    first = [b1.image(k, scaled=False) for k in range(3)]
    b1.release()
    rev = cases[::-1]
    b2 = Batch([a for a, _ in rev], [dict(o, write_exr=0) for _, o in rev]).prepare()
    b2.render_iteration()
    for k in range(3):
        assert np.array_equal(b2.image(2 - k, scaled=False), first[k])
    b2.release()


def test_batch_rejects_mixed_merge_arithmetic():
    cases = _cases()[:2]
    opts = [dict(cases[0][1], write_exr=0), dict(cases[1][1], write_exr=0, exact_merge=1)]
    b = Batch([a for a, _ in cases], opts)
    with pytest.raises(Exception, match="exact_merge"):
        b.prepare()
    b.release()


def _guided_cases():
    return [
        (scenes.water_scene(32, 24, 16, samples=3), dict(photons=40000, seed=3)),
        (scenes.c0_scene(32, 32, subdiv=2, samples=3), dict(photons=50000, seed=5, max_path_length=4)),
        (scenes.slab_scene(20, 20, samples=3), dict(photons=4096, seed=11)),  # no proposal: plain emission
    ]


def test_batch_with_fixed_proposals_matches_scenes_rendered_alone():
    """Per-scene proposals (fixed networks: same initial parameters, no
    training) sample the same directions in the batch as alone."""
    guide = dict(guide=1, guide_train=0, guide_warmup=0, guide_batch=4096)
    cases = _guided_cases()
    opts = [dict(o, write_exr=0, **(guide if k < 2 else {})) for k, (_, o) in enumerate(cases)]
    iters = 2
    alone = [_alone(a, o, iters) for (a, _), o in zip(cases, opts)]
    b = Batch([a for a, _ in cases], opts).prepare()
    for _ in range(iters):
        b.render_iteration()
    for k in range(len(cases)):
        img = b.image(k, scaled=False)
        assert img.sum() > 0
        assert np.array_equal(img, alone[k]), k
    b.release()


def test_batch_trains_each_scene_proposal():
    """Training inside the batch: every guided scene takes its own step on its
    own first guide_batch paths each iteration; the films stay close to the
    films of the scenes trained alone (same photons and targets; float atomics
    in the targets and gradients reorder sums, so not bit for bit)."""
    guide = dict(guide=1, guide_warmup=1, guide_batch=8192)
    cases = _guided_cases()
    opts = [dict(o, write_exr=0, **(guide if k < 2 else {})) for k, (_, o) in enumerate(cases)]
    iters = 3
    alone = [_alone(a, o, iters) for (a, _), o in zip(cases, opts)]
    b = Batch([a for a, _ in cases], opts).prepare()
    for _ in range(iters):
        b.render_iteration()
        st = b.stats()
        assert st["guide_rows"] == 2 * 8192
        assert st["ms_guide"] > 0
    for k in range(len(cases)):
        img = b.image(k, scaled=False)
        assert np.all(np.isfinite(img)) and img.sum() > 0
        rel = abs(img.sum() - alone[k].sum()) / alone[k].sum()
        print(f"scene {k}: batch {img.sum():.6g} alone {alone[k].sum():.6g} rel {rel:.2e}", flush=True)
        assert rel < 0.05, k
        if k == 2:
            assert np.array_equal(img, alone[k])  # the unguided scene is untouched by its neighbours' training
    b.release()


def test_batched_scenes_match_the_oracle(oracle):
    # each scene of a batch against the oracle's own render of it (not only against the scene alone)
    cases = [(scenes.c0_scene(32, 32, subdiv=3, samples=2), dict(photons=16384, seed=11)),
             (scenes.materials_scene(40, 24), dict(photons=16384, seed=12)),
             (scenes.water_scene(32, 24, 16, samples=2), dict(photons=8192, seed=13))]
    b = Batch([a for a, _ in cases], [dict(o, write_exr=0) for _, o in cases]).prepare()
    b.render_iteration()
    b.render_iteration()
    for k, (args, opts) in enumerate(cases):
        img_g = b.image(k, scaled=True)
        img_o = oracle.OracleScene(args, **opts).render(2)
        bad = np.abs(img_g - img_o) > 2e-4 * np.abs(img_o) + 1e-7
        print(f"scene {k}: sum {img_g.sum():.6g} vs oracle {img_o.sum():.6g}, outside {bad.mean():.4f}", flush=True)
        assert img_o.sum() > 0 and bad.mean() <= 0.002, (k, bad.mean())
    b.release()


def test_frames_render_consecutive_iterations():
    """K copies of one scene with iteration_offset k and iteration_stride K
    (convergence.FrameBatch) render iterations 0 .. K B - 1 of the scene: each
    frame's film equals the scene alone with the same offset and stride (bit
    for bit), and the mean of the K films equals the scene alone rendering
    K B consecutive iterations (the same terms, summed per frame first)."""
    from realcausticsengine_amd import convergence

    args = scenes.water_scene(32, 24, 16, samples=8, alpha=1.0)
    opts = dict(photons=20000, seed=9, write_exr=0, camera_qmc=1)
    K, B = 4, 2
    fb = convergence.FrameBatch(args, K, opts)
    for _ in range(B):
        fb.render_iteration()
    frames = [fb.batch.image(k, scaled=False) for k in range(K)]
    mean = fb.image(scaled=False)
    fb.release()
    for k in range(K):
        alone = _alone(args, dict(opts, iteration_offset=k, iteration_stride=K), B)
        assert frames[k].sum() > 0
        assert np.array_equal(frames[k], alone), k
    consecutive = _alone(args, opts, K * B)
    rel = np.abs(mean * K - consecutive) / (np.abs(consecutive) + 1e-6)
    print(f"frames mean vs consecutive: max rel {rel.max():.2e}", flush=True)
    assert np.allclose(mean * K, consecutive, rtol=1e-5, atol=1e-6)


@pytest.mark.skip(reason="the batch's photon-buffer regrow faulted the GPU before its fix (the scenes' photon-map "
                         "views kept the freed sorted-photon arrays); kept off the GPU until the fix is reviewed")
def test_batch_photon_buffer_regrow():
    """A batch whose light pass stores more photons than the shared buffer's
    first size (a quarter of the light paths, at least 2^20) grows the buffer
    and runs the light pass again; the merge must then read the new sorted
    photon arrays.  Films equal each scene rendered alone."""
    cases = [(scenes.slab_scene(32, 32, light_size=4.0, samples=2), dict(photons=1 << 21, seed=3)),
             (scenes.slab_scene(24, 24, light_size=4.0, samples=2), dict(photons=1 << 21, seed=4))]
    iters = 2
    alone = [_alone(a, o, iters) for a, o in cases]
    b = Batch([a for a, _ in cases], [dict(o, write_exr=0) for _, o in cases]).prepare()
    for _ in range(iters):
        b.render_iteration()
        st = b.stats()
        assert st["photons"] > max(st["light_paths"] / 4, 1 << 20), st["photons"]  # the first buffer overflowed
    for k in range(len(cases)):
        img = b.image(k, scaled=False)
        assert img.sum() > 0
        assert np.array_equal(img, alone[k]), k
    b.release()
