"""The merge in spatial query order (option query_order = 1, the default)
against the merge in trace order, and the grouped merge (group_queries = 4,
8 -- the default -- or 16 queries per wave) and the merge with one query per
lane (group_queries = 1) against the per-query merge.

In either merge a query's sum is formed by one wave in an order fixed by the
query's cell window alone (whichever queries share its wave), and the film
adds every query's contribution through its path slot, so the films of the
two query orders must be bit-identical; only the L2 reuse between
neighbouring queries changes.  The grouped and per-query merges add the same
terms in different orders: films within float rounding of each other."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu


def _films(order: int, graph: int = 0, iters: int = 2, group: int = 8, tile: int = 1024, block: int = 0,
           exact: int = 0, overlap: int = 1):
    args = scenes.water_scene(128, 96, 80, samples=iters)
    eng = Engine(args, photons=200000, seed=4, write_exr=0, query_order=order, graph=graph, group_queries=group,
                 gather_tile=tile, gather_block=block, exact_merge=exact, overlap=overlap)
    eng.prepare()
    films = []
    for _ in range(iters):
        eng.render_iteration()
        films.append(eng.image(scaled=False))
    st = eng.stats()
    eng.release()
    return films, st


@pytest.mark.parametrize("group", [0, 8, 1])
def test_sorted_merge_matches_trace_order(group):
    plain, st0 = _films(0, group=group)
    ordered, st1 = _films(1, group=group)
    assert st1["ms_query_sort"] > st0["ms_query_sort"]  # order 0: two events back to back
    for a, b in zip(plain, ordered):
        assert a.sum() > 0
        assert np.array_equal(a, b)


def test_sorted_merge_in_graph_matches_trace_order():
    plain, _ = _films(0)
    graphed, _ = _films(1, graph=1)
    for a, b in zip(plain, graphed):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("group", [4, 8, 16, 1])
def test_grouped_merge_matches_per_query_merge(group):
    per_query, s0 = _films(1, group=0)
    grouped, s1 = _films(1, group=group)
    for a, b in zip(per_query, grouped):
        assert a.sum() > 0
        np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-6 * float(a.max()))
    assert s0["gathers"] == s1["gathers"]


def test_grouped_merge_sizes_agree_bitwise():
    """A query's sum depends only on its window's photon stream, not on the group size."""
    films = [_films(1, group=g)[0] for g in (4, 8, 16)]
    for f in films[1:]:
        for a, b in zip(films[0], f):
            assert np.array_equal(a, b)


def test_gather_tiles_keep_the_film():
    """XCD tiles and the dealing of groups to waves change which wave merges a
    group, not a query's sum: one contiguous segment per XCD in a fixed order
    (tile 0) and tiles of 64 or 1024 queries taken as waves free up give the
    same film, in a graph too."""
    base = _films(1, group=16, tile=0)[0]
    for tile, graph in ((64, 0), (1024, 0), (1024, 1)):
        for a, b in zip(base, _films(1, group=16, tile=tile, graph=graph)[0]):
            assert a.sum() > 0
            assert np.array_equal(a, b)


@pytest.mark.parametrize("exact", [0, 1])
def test_rolled_query_loop_keeps_the_film(exact):
    """The rolled query loop over 2 or 4 photon steps at a time (sums in LDS)
    adds each query's terms in the unrolled loop's order: the same film, for
    every group size, with fast and exact merge arithmetic."""
    for group in (4, 8, 16):
        base = _films(1, group=group, exact=exact)[0]
        for block in (2, 4):
            for a, b in zip(base, _films(1, group=group, block=block, exact=exact)[0]):
                assert a.sum() > 0
                assert np.array_equal(a, b), (group, block)


@pytest.mark.parametrize("order,graph,overlap", [(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 0, 2), (1, 0, 2), (1, 1, 2)])
def test_overlapped_camera_trace_keeps_the_film(order, graph, overlap):
    """The camera trace on a second stream, beside the light pass and the
    photon map (option overlap = 1, the default; 2: queued after the light
    pass), traces the same queries from the same keys as the serial trace:
    the same film and query count, with direct launches and with the graph
    replay."""
    serial, s0 = _films(order, group=8, overlap=0, iters=3)
    overlapped, s1 = _films(order, group=8, overlap=overlap, iters=3, graph=graph)
    for a, b in zip(serial, overlapped):
        assert a.sum() > 0
        assert np.array_equal(a, b)
    assert s0["gathers"] == s1["gathers"] > 0


def test_grouped_merge_counts_match_per_query_merge():
    args = scenes.water_scene(128, 96, 80, samples=1)
    st = []
    for group, block in ((0, 0), (8, 0), (8, 4), (1, 0)):
        eng = Engine(args, photons=200000, seed=4, write_exr=0, count_events=1, group_queries=group,
                     gather_block=block)
        eng.prepare()
        eng.render_iteration()
        st.append(eng.stats())
        eng.release()
    for k in ("gathers", "photons_tested", "photons_accepted"):
        assert st[0][k] == st[1][k] == st[2][k] == st[3][k], k
    assert st[3]["cell_visits"] == st[0]["cell_visits"]  # one lookup per cell of every query's window
    assert st[1]["cell_visits"] <= st[0]["cell_visits"]  # one window lookup per sub-run of a group
    assert st[2]["cell_visits"] == st[1]["cell_visits"] and st[2]["photons_streamed"] == st[1]["photons_streamed"]


def test_sorted_merge_in_batch_matches_trace_order():
    cases = [(scenes.water_scene(32, 24, 16, samples=2), dict(photons=20000, seed=3)),
             (scenes.c0_scene(32, 32, subdiv=2, samples=2), dict(photons=50000, seed=5, max_path_length=4)),
             (scenes.slab_scene(20, 20, samples=2), dict(photons=4096, seed=11))]
    out = []
    for order in (0, 1):
        b = Batch([a for a, _ in cases], [dict(o, write_exr=0, query_order=order) for _, o in cases]).prepare()
        b.render_iteration()
        b.render_iteration()
        out.append([b.image(k, scaled=False) for k in range(len(cases))])
        b.release()
    for a, b in zip(*out):
        assert a.sum() > 0
        assert np.array_equal(a, b)


def test_refill_thresholds_keep_the_film():
    """Waves that wait for idle lanes before taking new paths trace the same
    paths with the same keys; photons and queries land in other buffer slots,
    which the photon-map sort and the merge order undo."""
    args = scenes.water_scene(128, 96, 80, samples=2)
    films = []
    for r in (1, 32):
        eng = Engine(args, photons=200000, seed=4, write_exr=0, light_refill=r, camera_refill=r)
        eng.prepare()
        eng.render_iteration()
        eng.render_iteration()
        films.append(eng.image(scaled=False))
        eng.release()
    assert films[0].sum() > 0
    assert np.array_equal(films[0], films[1])


def test_light_path_order_keeps_the_film():
    """Light paths traced in the order of their emission samples: each path is
    computed from its own index and key, so the photons are the same and the
    photon-map sort restores their order."""
    args = scenes.water_scene(128, 96, 80, samples=2)
    films = []
    for order in (0, 1):
        eng = Engine(args, photons=200000, seed=4, write_exr=0, light_order=order)
        eng.prepare()
        eng.render_iteration()
        eng.render_iteration()
        films.append(eng.image(scaled=False))
        eng.release()
    assert films[0].sum() > 0
    assert np.array_equal(films[0], films[1])


def test_light_path_order_in_batch_keeps_the_films():
    cases = [(scenes.water_scene(32, 24, 16, samples=2), dict(photons=20000, seed=3)),
             (scenes.c0_scene(32, 32, subdiv=2, samples=2), dict(photons=50000, seed=5, max_path_length=4))]
    out = []
    for order in (0, 1):
        b = Batch([a for a, _ in cases], [dict(o, write_exr=0, light_order=order) for _, o in cases]).prepare()
        b.render_iteration()
        out.append([b.image(k, scaled=False) for k in range(len(cases))])
        b.release()
    for a, b in zip(*out):
        assert a.sum() > 0
        assert np.array_equal(a, b)


def test_light_emission_kernel_keeps_photons_and_films():
    """Option light_emit (the light samples of every path made by a coherent
    kernel before the persistent light pass): the same photons, bit for bit,
    and the same film -- single scene and batch."""
    from realcausticsengine_amd.engine import Batch

    args = scenes.c0_scene(48, 48, subdiv=3, samples=2)
    out = []
    for emit in (0, 1):
        eng = Engine(args, photons=65536, seed=5, write_exr=0, light_emit=emit)
        ph, uid = eng.trace_photons(0)
        eng.render_iteration()
        eng.render_iteration()
        out.append((ph, uid, eng.image(scaled=False)))
        eng.release()
    (p0, u0, f0), (p1, u1, f1) = out
    o0, o1 = np.argsort(u0), np.argsort(u1)
    assert np.array_equal(u0[o0], u1[o1]) and np.array_equal(p0[o0].view(np.uint32), p1[o1].view(np.uint32))
    assert np.array_equal(f0, f1) and f0.sum() > 0
    films = []
    for emit in (0, 1):
        cases = [scenes.c0_scene(32, 32, subdiv=2, samples=2), scenes.water_scene(32, 24, 16, samples=2)]
        b = Batch(cases, [dict(photons=20000, seed=3, write_exr=0, light_emit=emit),
                          dict(photons=30000, seed=4, write_exr=0, light_emit=emit)]).prepare()
        b.render_iteration()
        films.append([b.image(k, scaled=False) for k in range(2)])
        b.release()
    for a, c in zip(*films):
        assert np.array_equal(a, c)
