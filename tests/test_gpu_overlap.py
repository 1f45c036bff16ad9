"""The overlapped camera trace (option overlap) when the merge query buffer
overflows.

Between two facing diffuse planes of albedo 0.9 a camera path makes a merge
query at every bounce, so the first iteration's trace makes more queries
than the buffer's first size (two per camera path); a glass sphere under the
light gives the caustic photons the queries merge.  Serially the trace runs
again into a larger buffer before the merge; overlapped, the host finds the
overflow in the count the trace copied to pinned memory and traces again on
the main stream.  Both must give the same film and query count."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _two_planes(width=48, height=32, samples=3):
    fv, fn, ft = scenes.plane(10.0, 10.0, 0.0)
    cv, cn, ct = scenes.plane(10.0, 10.0, 1.0)
    cn = -cn  # the ceiling faces down
    ct = np.ascontiguousarray(ct[:, ::-1])
    sv, sn, st = scenes.icosphere(2)
    m = np.eye(4, dtype=np.float32) * 0.25
    m[3, 3] = 1.0
    m[:3, 3] = (0.0, 1.0, 0.55)
    meshes = [scenes.MeshData(fv, fn, ft, material=0), scenes.MeshData(cv, cn, ct, material=0),
              scenes.MeshData(sv, sn, st, matrix=m, material=1)]
    mats = [scenes.MaterialData("P", roughness=1.0, specular=0.0, color=(0.9, 0.9, 0.9)),
            scenes.MaterialData("G", ior=1.5, color=(1.0, 1.0, 1.0))]
    lights = [scenes.LightData("A", position=(0.0, 1.0, 0.95), width=0.3, height=0.3, power=1.0)]
    cam = scenes.look_at_camera(width, height, (0.0, -1.0, 0.5), (0.0, 3.0, 0.3), 50.0)
    return scenes.SceneArgs(meshes, lights, mats, cam, samples=samples, base_radius=0.02, radius_reduction=0.75)


def _render(overlap, iters=3):
    args = _two_planes(samples=iters)
    eng = Engine(args, photons=65536, seed=9, write_exr=0, overlap=overlap)
    eng.prepare()
    films, gathers = [], []
    for _ in range(iters):
        eng.render_iteration()
        films.append(eng.image(scaled=False))
        gathers.append(eng.stats()["gathers"])
    paths = eng.stats()["camera_paths"]
    eng.release()
    return films, gathers, paths


@pytest.mark.parametrize("overlap", [1, 2])
def test_overlapped_trace_retraces_on_query_overflow(overlap):
    serial, g0, paths = _render(0)
    overlapped, g1, _ = _render(overlap)
    assert g0[0] > 2 * paths, (g0[0], paths)  # the first trace overflowed the first buffer
    assert g0 == g1
    for a, b in zip(serial, overlapped):
        assert a.sum() > 0
        assert np.array_equal(a, b)


def _render_batch(overlap, iters=3):
    """The two-plane scene twice in one batch (the second with another seed
    and a smaller film) beside a glass-sphere scene: the batch's first query
    buffer (2 queries per camera path + 1024) overflows in the first iteration."""
    from realcausticsengine_amd.engine import Batch

    cases = [_two_planes(samples=iters), _two_planes(40, 24, samples=iters), scenes.c0_scene(32, 32, subdiv=2,
                                                                                            samples=iters)]
    opts = [dict(photons=65536, seed=9, write_exr=0, overlap=overlap),
            dict(photons=32768, seed=4, write_exr=0, overlap=overlap),
            dict(photons=16384, seed=2, write_exr=0, overlap=overlap)]
    b = Batch(cases, opts)
    b.prepare()
    films, gathers = [], []
    for _ in range(iters):
        b.render_iteration()
        films.append([b.image(i, scaled=False) for i in range(len(cases))])
        gathers.append(b.stats()["gathers"])
    paths = b.stats()["camera_paths"]
    b.release()
    return films, gathers, paths


@pytest.mark.parametrize("overlap", [1, 2])
def test_overlapped_batch_trace_retraces_on_query_overflow(overlap):
    serial, g0, paths = _render_batch(0)
    overlapped, g1, _ = _render_batch(overlap)
    assert g0[0] > 2 * paths, (g0[0], paths)  # the batch's first trace overflowed its first buffer
    assert g0 == g1
    for fs, fo in zip(serial, overlapped):
        for a, b in zip(fs, fo):
            assert a.sum() > 0
            assert np.array_equal(a, b)
