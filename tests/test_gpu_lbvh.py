"""The GPU-built linear BVH (option bvh_build = 1), traversed as binary nodes
(bvh_width = 2) and collapsed on the GPU to 4-wide nodes (bvh_width = 4, the
default), against brute force, the oracle and the host-built SAH tree.

The tree changes only which nodes a ray visits: the closest hit is the same
triangle with the same Moller-Trumbore distance bits, except where two
triangles tie in distance (shared edges), so first hits, photon lists and
films follow the bars of the host-tree parity tests."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine, photons_from_oracle
from realcausticsengine_amd.scenes import LightData, MaterialData, MeshData, SceneArgs, look_at_camera

pytestmark = pytest.mark.gpu


def _rays(args, n, seed, spread=3.0, z=1.2):
    rng = np.random.default_rng(seed)
    cam = np.asarray(args.camera.position, dtype=np.float64)
    c = np.asarray(args.camera.corners, dtype=np.float64)
    m = n // 2
    s = rng.random((m, 2))
    pts = c[2] + s[:, :1] * (c[1] - c[2]) + s[:, 1:] * (c[3] - c[2])
    d1 = pts - cam
    o2 = rng.uniform(-spread, spread, (n - m, 3)) + np.asarray([0, 0, z])
    d2 = rng.normal(size=(n - m, 3))
    o = np.concatenate([np.tile(cam, (m, 1)), o2])
    d = np.concatenate([d1, d2])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def _check_hits(tri_g, t_g, tri_o, t_o):
    hit = tri_o >= 0
    assert np.array_equal(tri_g >= 0, hit)
    same = tri_g == tri_o
    assert np.array_equal(t_g[same & hit].view(np.uint32), t_o[same & hit].astype(np.float32).view(np.uint32))
    tie = ~same & hit
    assert np.all(np.abs(t_g[tie] - t_o[tie]) <= 1e-5 * t_o[tie])
    assert tie.sum() <= 0.001 * max(1, hit.sum())
    return hit.sum()


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("scene", ["c0", "random"])
def test_lbvh_first_hit_matches_brute_force(oracle, scene, width):
    args = scenes.c0_scene(64, 64, subdiv=3) if scene == "c0" else scenes.random_mesh_scene(seed=2)
    eng = Engine(args, bvh_build=1, bvh_width=width)
    eng.prepare()
    ntri = len(np.concatenate([m.tris for m in args.meshes]))
    assert eng.get("bvh_build") == 1 and eng.get("node_bytes") == 32 * width
    assert eng.get("nodes") == ntri - 1
    if width == 4:
        # the internal nodes of even depth
        assert (ntri - 1) / 4 <= eng.get("nodes4") < ntri - 1
    else:
        assert eng.get("nodes4") == 0
    orc = oracle.OracleScene(args)
    rays = _rays(args, 20000, 1)
    tri_g, t_g = eng.first_hit(rays)
    tri_o, t_o = orc.first_hit(rays, brute=True)
    assert _check_hits(tri_g, t_g, tri_o, t_o) > 5000
    eng.release()


@pytest.mark.parametrize("width", [2, 4])
def test_lbvh_two_triangles(oracle, width):
    """The smallest tree the GPU build makes: one internal node over two leaves."""
    pv, pn, pt = scenes.plane(4.0, 4.0, 0.0)
    args = SceneArgs([MeshData(pv, pn, pt, material=0)], [LightData("A", position=(0.0, 0.0, 3.0), width=1.0,
                                                                    height=1.0, power=1.0)],
                     [MaterialData("P", roughness=1.0, specular=0.0, color=(0.8, 0.8, 0.8))],
                     look_at_camera(16, 16, (0.0, -3.0, 3.0), (0.0, 0.0, 0.0), 40.0), samples=1)
    eng = Engine(args, bvh_build=1, bvh_width=width)
    eng.prepare()
    assert eng.get("nodes") == 1 and eng.get("bvh_depth") == 1 and eng.get("nodes4") == (1 if width == 4 else 0)
    orc = oracle.OracleScene(args)
    rays = _rays(args, 4000, 5, spread=1.5, z=1.0)
    tri_g, t_g = eng.first_hit(rays)
    tri_o, t_o = orc.first_hit(rays, brute=True)
    assert _check_hits(tri_g, t_g, tri_o, t_o) > 1000
    eng.release()


@pytest.mark.parametrize("width", [2, 4])
def test_lbvh_photon_list_matches_oracle(oracle, width):
    args = scenes.c0_scene(64, 64, subdiv=3)
    eng = Engine(args, photons=32768, seed=7, bvh_build=1, bvh_width=width)
    orc = oracle.OracleScene(args, photons=32768, seed=7)
    ph_g, uid_g = eng.trace_photons(0)
    ph_o = orc.trace_light(0)
    uid_o = ph_o[:, 10].astype(np.uint32)
    common, ig, io = np.intersect1d(uid_g, uid_o, return_indices=True)
    assert len(common) >= 0.999 * max(len(uid_g), len(uid_o))
    exact = np.all(ph_g[ig].view(np.uint32) == photons_from_oracle(ph_o[io]).view(np.uint32), axis=1)
    assert exact.mean() >= 0.999, exact.mean()
    eng.release()


def test_lbvh_two_million_triangles_match_host_tree():
    """A 1024 x 1024-vertex water surface (2.09 M triangles): GPU-built trees
    (binary and collapsed 4-wide) and the host-built tree give the same first
    hits and the same photons."""
    args = scenes.water_scene(1024, 64, 64, samples=1)
    ntri = sum(len(m.tris) for m in args.meshes)
    assert ntri > 2_000_000
    rays = _rays(args, 100000, 9, spread=1.5, z=0.2)
    out = {}
    for build, width in ((0, 2), (1, 2), (1, 4)):
        eng = Engine(args, photons=1 << 18, seed=3, bvh_build=build, bvh_width=width)
        eng.prepare()
        depth, ms = eng.get("bvh_depth"), eng.get("bvh_build_ms")
        print(f"bvh_build={build} width={width}: depth {depth:.0f}, build {ms:.1f} ms, nodes {eng.get('nodes'):.0f}, "
              f"4-wide nodes {eng.get('nodes4'):.0f}", flush=True)
        assert 0 < depth < 64
        tri, t = eng.first_hit(rays)
        ph, uid = eng.trace_photons(0)
        out[(build, width)] = (tri, t, ph, uid)
        eng.release()
    tri_h, t_h, ph_h, uid_h = out[(0, 2)]
    for key in ((1, 2), (1, 4)):
        tri_l, t_l, ph_l, uid_l = out[key]
        assert _check_hits(tri_l, t_l, tri_h, t_h) > 50000
        common, ia, ib = np.intersect1d(uid_l, uid_h, return_indices=True)
        assert len(common) >= 0.999 * max(len(uid_l), len(uid_h))
        exact = np.all(ph_l[ia].view(np.uint32) == ph_h[ib].view(np.uint32), axis=1)
        assert exact.mean() >= 0.999, (key, exact.mean())


@pytest.mark.parametrize("build", [0, 1])
def test_deep_tree_traverses_within_its_stack(oracle, build):
    """A tree whose 4-wide traversal could need more than the 64 stack
    entries the kernels hold (tests/test_bvh_stack_cpu.py builds it): the
    host build keeps the binary traversal, and first hits still equal brute
    force on both builds."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_bvh_stack_cpu import _deep_spine
    from realcausticsengine_amd.engine import bvh_shape

    tris = _deep_spine()
    shape = bvh_shape(tris)
    assert shape["width"] == 2 and shape["stack_need"] <= 64
    verts = tris.reshape(-1, 3)
    idx = np.arange(len(verts), dtype=np.uint32).reshape(-1, 3)
    normals = np.tile(np.asarray([0.0, 0.0, 1.0]), (len(verts), 1))
    args = SceneArgs([MeshData(verts, normals, idx, material=0)],
                     [LightData("A", position=(0.0, 0.0, 5.0), width=1.0, height=1.0, power=1.0)],
                     [MaterialData("P", roughness=1.0, specular=0.0)],
                     look_at_camera(8, 8, (0.0, -6.0, 4.0), (0.0, 0.0, 0.0), 40.0), samples=1)
    eng = Engine(args, bvh_build=build)
    eng.prepare()
    print(f"build {build}: depth {eng.get('bvh_depth'):.0f}, depth4 {eng.get('bvh_depth4'):.0f}, "
          f"width {eng.get('bvh_traversal_width'):.0f}", flush=True)
    if build == 0:
        assert eng.get("bvh_traversal_width") == 2 and eng.get("bvh_depth") == shape["depth"]
    w = eng.get("bvh_traversal_width")
    need = 3 * (eng.get("bvh_depth4") + 1) if w == 4 else eng.get("bvh_depth")
    assert need <= 64
    # vertical rays onto the centroids of the triangles from 17^-1 to 17^11
    # (Triangle::hit rejects the smaller ones by its determinant epsilon)
    c = tris.astype(np.float64).mean(axis=1)[32 * 11:32 * 24]
    o = c + np.asarray([0.0, 0.0, 1.0]) * np.maximum(1e-3, 0.1 * np.abs(c[:, :1]))
    d = np.tile(np.asarray([0.0, 0.0, -1.0]), (len(c), 1))
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    orc = oracle.OracleScene(args)
    tri_g, t_g = eng.first_hit(rays, tmin=0.0)
    tri_o, t_o = orc.first_hit(rays, tmin=0.0, brute=True)
    assert _check_hits(tri_g, t_g, tri_o, t_o) >= 0.9 * len(rays)
    eng.release()
