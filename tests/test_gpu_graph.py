"""The captured iteration graph (option "graph") against direct launches.

With the plain light pass the whole iteration is one graph (film copy, light
pass, photon map, camera pass, resolve) with every count read from device
memory and the per-iteration values from a device struct; with the learned
proposal the light pass stays a direct launch before the graph.  The kernels
and their data are the same as in direct launches, so each replay must leave
the film bit-identical to the direct run's film.  Every replay is checked,
with a readback between replays; a photon buffer that overflows inside the
graph undoes the iteration's film and runs it again with larger buffers."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _run(graph: int, iters: int, schedule=None, **opts):
    args = scenes.water_scene(64, 64, 48, samples=iters)
    eng = Engine(args, seed=1, write_exr=0, graph=graph, **opts)
    eng.prepare()
    films = []
    for i in range(iters):
        if schedule:
            eng.set("photons", schedule[i])
        eng.render_iteration()
        films.append(eng.image(scaled=False))
    caps = eng.get("graph_captures")
    eng.release()
    return films, caps


@pytest.mark.parametrize("light_order", [0, 1])
def test_graph_replays_match_direct_launches(light_order):
    direct, _ = _run(0, 6, photons=65536, light_order=light_order)
    graphed, caps = _run(1, 6, photons=65536, light_order=light_order)
    assert caps == 1
    for i in range(6):
        assert direct[i].sum() > 0
        assert np.array_equal(graphed[i], direct[i]), i


def test_graph_is_captured_again_when_buffers_change():
    sched = [40000, 40000, 90000, 90000, 20000]  # the photon count changes the key layout and the buffers
    direct, _ = _run(0, 5, schedule=sched, photon_capacity=1000)
    graphed, caps = _run(1, 5, schedule=sched, photon_capacity=1000)
    assert caps >= 3
    for i in range(5):
        assert np.array_equal(graphed[i], direct[i]), i


def test_graph_with_learned_proposal_runs():
    films, caps = _run(1, 5, photons=65536, guide=1, guide_warmup=1)
    assert caps == 1
    assert np.all(np.isfinite(films[-1])) and films[-1].sum() > 0


@pytest.mark.parametrize("overlap", [0, 1])
def test_graph_render_matches_the_oracle(oracle, overlap):
    # the replayed iteration against the oracle's render (whole-iteration graph without overlap)
    args = scenes.c0_scene(32, 32, subdiv=3, samples=2)
    eng = Engine(args, photons=16384, seed=11, write_exr=0, graph=1, overlap=overlap)
    eng.prepare()
    eng.render_iteration()
    eng.render_iteration()
    img_g = eng.image(scaled=True)
    assert eng.get("graph_captures") >= 1
    eng.release()
    img_o = oracle.OracleScene(args, photons=16384, seed=11).render(2)
    bad = np.abs(img_g - img_o) > 2e-4 * np.abs(img_o) + 1e-7
    assert img_o.sum() > 0 and bad.mean() <= 0.002, bad.mean()
