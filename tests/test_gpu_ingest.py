"""Scene ingest on the GPU (option device_ingest, with bvh_build = 1) against
the host ingest.

init hands over Blender's arrays (MVert, MPoly, MLoop per mesh, an object
matrix and a material slot table; src/main.cpp:89-143, src/mesh.cpp:12-26).
The host ingest transforms them on the CPU (Transform.cpp:3-31) and uploads
triangles; the device ingest copies the raw bytes and converts them in two
kernels with the same arithmetic.  Both must give the same triangles, so the
same GPU-built tree: first hits (triangle and distance bits) and rendered
films equal bit for bit, on scenes of several transformed meshes whose
polygons have more than three loops, loops out of order and negative or
several material slots.  Invalid meshes are rejected with the host ingest's
messages."""
import ctypes

import numpy as np
import pytest

from realcausticsengine_amd import blender, scenes
from realcausticsengine_amd.engine import Engine, EngineError

pytestmark = pytest.mark.gpu


def _rays(args, n, seed):
    rng = np.random.default_rng(seed)
    cam = np.asarray(args.camera.position, dtype=np.float64)
    c = np.asarray(args.camera.corners, dtype=np.float64)
    s = rng.random((n // 2, 2))
    d1 = c[2] + s[:, :1] * (c[1] - c[2]) + s[:, 1:] * (c[3] - c[2]) - cam
    o2 = rng.uniform(-3.0, 3.0, (n - n // 2, 3)) + np.asarray([0.0, 0.0, 1.0])
    d2 = rng.normal(size=(n - n // 2, 3))
    o = np.concatenate([np.tile(cam, (n // 2, 1)), o2])
    d = np.concatenate([d1, d2])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def _scramble_polygons(args, mesh, seed):
    """Rewrites one mesh's MPoly/MLoop arrays: polygons in shuffled loop order,
    every third polygon with a fourth (unused) loop, mat_nr -1 for slot 0; the
    same triangles (the first three loops of each polygon) as before."""
    mv, ml, mp, blk = args._keep[4 * mesh: 4 * mesh + 4]
    tris = ml["v"].reshape(-1, 3)
    nt = len(tris)
    rng = np.random.default_rng(seed)
    extra = (np.arange(nt) % 3) == 0
    sizes = 3 + extra.astype(np.int64)
    order = rng.permutation(nt)  # polygon p's loops sit at the order[p]-th place
    starts = np.zeros(nt, dtype=np.int64)
    starts[order] = np.concatenate([[0], np.cumsum(sizes[order])[:-1]])
    loops = np.zeros(int(sizes.sum()), dtype=blender.MLOOP_DTYPE)
    for k in range(3):
        loops["v"][starts + k] = tris[:, k]
    loops["v"][starts[extra] + 3] = tris[extra, 0]
    poly = mp.copy()
    poly["loopstart"] = starts.astype(np.int32)
    poly["totloop"] = sizes.astype(np.int32)
    poly["mat_nr"] = np.where(poly["mat_nr"] == 0, -1, poly["mat_nr"])
    blk[blender.MESH_OFF_MPOLY:blender.MESH_OFF_MPOLY + 8] = np.frombuffer(np.uint64(poly.ctypes.data).tobytes(), np.uint8)
    blk[blender.MESH_OFF_MLOOP:blender.MESH_OFF_MLOOP + 8] = np.frombuffer(np.uint64(loops.ctypes.data).tobytes(),
                                                                           np.uint8)
    args._keep += [poly, loops]


def _multi_slot_scene():
    """random_mesh_scene with its first mesh's triangles spread over three
    material slots and its polygons scrambled."""
    args = scenes.random_mesh_scene(seed=3, width=48, height=32)
    mp = args._keep[2]
    mp["mat_nr"] = (np.arange(len(mp)) % 3).astype(np.int16)
    slots = (ctypes.c_int * 3)(0, 1, 2)
    args.mat_index_arrays[0] = slots
    args.materials_indices[0] = ctypes.cast(slots, ctypes.POINTER(ctypes.c_int))
    _scramble_polygons(args, 0, seed=5)
    return args


CASES = {
    "random_meshes": lambda: scenes.random_mesh_scene(seed=1, width=48, height=32),
    "materials": lambda: scenes.materials_scene(),
    "slots_scrambled": _multi_slot_scene,
}


def _engine(args, ingest, **opts):
    return Engine(args, bvh_build=1, device_ingest=ingest, seed=4, write_exr=0, **opts)


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_ingest_equals_host_ingest(name):
    args = CASES[name]()
    rays = _rays(args, 8192, seed=11)
    out = []
    for ingest in (0, 1):
        eng = _engine(args, ingest, photons=65536)
        eng.prepare()
        tri, t = eng.first_hit(rays)
        eng.render_iteration()
        eng.render_iteration()
        out.append((tri, t, eng.image(scaled=False), eng.stats()["tris"], eng.get("device_ingest")))
        eng.release()
    (tri0, t0, img0, n0, _), (tri1, t1, img1, n1, _) = out
    assert n0 == n1 > 0
    assert (tri0 >= 0).mean() > 0.3
    assert np.array_equal(tri0, tri1)
    assert np.array_equal(t0.view(np.uint32), t1.view(np.uint32))
    assert img0.sum() > 0
    assert np.array_equal(img0, img1)


def test_device_ingest_at_scale_scene():
    args, _ = scenes.c4_scene(1, res=64, photons=1 << 16)
    rays = _rays(args, 65536, seed=12)
    hits = []
    for ingest in (0, 1):
        eng = _engine(args, ingest, photons=1 << 16, compat_point_light=0)
        eng.prepare()
        hits.append(eng.first_hit(rays))
        eng.release()
    assert (hits[0][0] >= 0).mean() > 0.3
    assert np.array_equal(hits[0][0], hits[1][0])
    assert np.array_equal(hits[0][1].view(np.uint32), hits[1][1].view(np.uint32))


def _rejected(args):
    try:
        eng = _engine(args, 1)
    except EngineError as e:
        return str(e)
    eng.release()
    return ""


def test_device_ingest_rejects_invalid_meshes():
    args = scenes.c0_scene(8, 8, subdiv=1)
    args._keep[0]["co"][3, 1] = np.nan
    assert "non-finite" in _rejected(args)
    args = scenes.c0_scene(8, 8, subdiv=1)
    args._keep[5]["v"][4] = 10 ** 6  # mesh 1's loops
    assert "loop vertex index" in _rejected(args)
    args = scenes.c0_scene(8, 8, subdiv=1)
    args._keep[2]["totloop"][4] = 2
    assert "fewer than 3 loops" in _rejected(args)
    args = scenes.c0_scene(8, 8, subdiv=1)
    args.mat_index_arrays[1][0] = 7
    assert "material index" in _rejected(args)
    assert _rejected(scenes.c0_scene(8, 8, subdiv=1)) == ""
