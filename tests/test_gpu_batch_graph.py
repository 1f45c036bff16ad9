"""Graph-captured batched iterations on a subset of the uneven-scene
workload (scenes.uneven_batch: displaced glass blobs of 1,200-499,392
triangles, log-uniform, GGX roughness U(0, 0.3) rendered as the rough
dielectric with compat_glass = 0; films and light paths scaled down for the
test).  With scene 0's option graph the batch's photon map and camera pass
replay one captured graph, the per-scene values coming from the device
tables refilled before each replay: every film must equal the film of the
same batch with direct launches, bit for bit, and each scene's film the
film of that scene rendered alone."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu

ITERS = 3


def _cases():
    return scenes.uneven_batch(8, samples=ITERS, res_scale=8, photon_scale=32)


def _batch(graph):
    cases = _cases()
    opts = [dict(photons=p, seed=100 + i, write_exr=0, compat_glass=0, graph=graph if i == 0 else 0)
            for i, (_, p) in enumerate(cases)]
    b = Batch([sc for sc, _ in cases], opts).prepare()
    films, stats = [], []
    for it in range(ITERS):
        assert b.render_iteration() == it
        films.append([b.image(k, scaled=False) for k in range(len(cases))])
        stats.append(b.stats())
    b.release()
    return films, stats


def test_graph_batch_matches_direct_batch():
    cases = _cases()
    tris = [sc.total_tris - 2 for sc, _ in cases]
    assert min(tris) >= 1000 and max(tris) <= 500000
    direct, st_d = _batch(0)
    graphed, st_g = _batch(1)
    assert st_d[-1]["graph_captures"] == 0
    assert 1 <= st_g[-1]["graph_captures"] <= ITERS
    for it in range(ITERS):
        assert st_g[it]["gathers"] == st_d[it]["gathers"]
        for k in range(len(cases)):
            assert direct[it][k].sum() > 0
            assert np.array_equal(graphed[it][k], direct[it][k]), (it, k)


def test_rough_blob_batch_matches_scenes_alone():
    cases = _cases()[:4]
    alone = []
    for i, (sc, p) in enumerate(cases):
        eng = Engine(sc, photons=p, seed=100 + i, write_exr=0, compat_glass=0)
        eng.prepare()
        for _ in range(2):
            eng.render_iteration()
        alone.append(eng.image(scaled=False))
        eng.release()
    b = Batch([sc for sc, _ in cases], [dict(photons=p, seed=100 + i, write_exr=0, compat_glass=0, graph=1 if i == 0
                                             else 0) for i, (_, p) in enumerate(cases)]).prepare()
    for _ in range(2):
        b.render_iteration()
    for k in range(len(cases)):
        assert np.array_equal(b.image(k, scaled=False), alone[k]), k
    b.release()
