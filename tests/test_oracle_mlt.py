"""The oracle's restatement of the engine's chain-parallel MLT light pass
(rng_serial = 2) against its restatement of the reference's serial loop
(rng_serial = 1, Vertex_Merging.cpp:220-282, MLT_Sampler.cpp:7-67).

The two draw their samples from different streams (Philox words against the
reference's xorshift), so they agree only in distribution.  A single chain
runs the reference's algorithm; its mutation acceptance and the brightness
of its photon set relative to the plain estimator's, (visible / N) x MLT flux
/ plain flux, are compared pass by pass over 16 light passes against the
serial loop's: the means agree within three standard errors (the adaptive
chain's acceptance varies from pass to pass between about 0.05 and 0.23 in
both).  Short chains (many chains of 64 indices, the engine's default) adapt
little; the test records that their acceptance is lower, so their estimate
is darker than the reference's (its expectation is plain x (visible +
accepted) / N)."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.blender import LightData, MaterialData, MeshData, SceneArgs, look_at_camera

N = 65536
PASSES = 16


def _stats(oracle, **opts):
    args = scenes.c0_scene(64, 64, subdiv=3, samples=4)
    o = oracle.OracleScene(args, photons=N, **opts)
    plain = oracle.OracleScene(args, photons=N)
    acc, ratio = [], []
    for it in range(PASSES):
        ph, pl = o.trace_light(it), plain.trace_light(it)
        vis, mut, a = o.get("visible_paths"), o.get("mlt_mutations"), o.get("mlt_accepted")
        assert vis + mut == N and 0 <= a <= mut
        acc.append(a / mut)
        flux = ph[:, 6:9].sum() / pl[:, 6:9].sum()
        # the serial loop's photons already carry visible / N (Vertex_Merging.cpp:278-282)
        ratio.append(flux if opts["rng_serial"] == 1 else flux * vis / N)
    return np.array(acc), np.array(ratio)


def test_single_chain_agrees_with_the_serial_loop_in_distribution(oracle):
    a_s, r_s = _stats(oracle, rng_serial=1)
    a_c, r_c = _stats(oracle, rng_serial=2, mlt_chains=1)
    print("serial acceptance", a_s.round(3), "ratio", r_s.round(3))
    print("chain  acceptance", a_c.round(3), "ratio", r_c.round(3))
    for x, y in ((a_s, a_c), (r_s, r_c)):
        se = np.sqrt(x.var(ddof=1) / PASSES + y.var(ddof=1) / PASSES)
        assert abs(x.mean() - y.mean()) < 3 * se + 0.01
    # the adaptation drives the acceptance towards the reference's target of 0.234
    assert 0.1 < a_c.mean() < 0.25


def test_short_chains_accept_less(oracle):
    a_s, r_s = _stats(oracle, rng_serial=1)
    a_m, r_m = _stats(oracle, rng_serial=2)  # N / 64 chains of 64 indices
    print("1024 chains acceptance", a_m.round(3), "ratio", r_m.round(3))
    assert a_m.mean() < a_s.mean() and r_m.mean() < r_s.mean()
    assert a_m.std() < 0.01  # many short chains average out the adaptation noise


# ---------------------------------------------------------------------------
# What the algorithm itself fixes.  The chain's target is the indicator of
# "the path stores a photon" (times the probability of the Russian-roulette
# draws that keep it, which the serial loop takes from its stream), its
# mutations are symmetric moves on the torus of primary samples, and a
# visible uniform sample replaces the state outright; each of these keeps
# the distribution of the plain estimator's visible paths invariant.  So the
# sequence of chain states -- one per light-path index, a rejected mutation
# repeating the state before it -- is distributed as the visible uniform
# paths, and where every uniform path is visible the loop never mutates.


def _first_photon_r(ph):
    """Path index and distance from the scene axis of each path's first photon."""
    uid = ph[:, 10].astype(np.int64)
    path, ln = uid >> 3, uid & 7
    o = np.lexsort((ln, path))
    path, ph = path[o], ph[o]
    first = np.r_[True, path[1:] != path[:-1]]
    return path[first], ph[first, :3]


def _bins(xyz, edges):
    """4 rings about the light axis x 2 half planes: 8 bins."""
    ring = np.searchsorted(edges, np.hypot(xyz[:, 0], xyz[:, 1]))
    return ring * 2 + (xyz[:, 0] > 0)


def _chain_state_bins(ph, n, chains, edges):
    """Bin of each chain's state after each of its light-path indices
    (forward-filled over rejected mutations; -1 before its first visible
    uniform path, which ends the start-up transient)."""
    path, xyz = _first_photon_r(ph)
    b = np.full(n, -1)
    b[path] = _bins(xyz, edges)
    out = []
    for c in range(chains):
        seq = b[c::chains]
        have = seq >= 0
        fill = np.maximum.accumulate(np.where(have, np.arange(len(seq)), -1))
        st = np.where(fill >= 0, seq[np.maximum(fill, 0)], -1)
        # the chain's states count from its first visible uniform path on
        plain_vis = np.isin(np.arange(c, n, chains), _PLAIN_VISIBLE[0])
        start = np.argmax(plain_vis) if plain_vis.any() else len(seq)
        out.append(st[start:])
    return out


_PLAIN_VISIBLE = [np.zeros(0, dtype=np.int64)]


@pytest.mark.parametrize("form", ["serial", "chains", "one_chain"])
def test_chain_states_are_distributed_as_visible_uniform_paths(oracle, form):
    args = scenes.c0_scene(64, 64, subdiv=3, samples=4)
    plain = oracle.OracleScene(args, photons=N)
    ref = []
    for it in range(8):
        ref.append(_first_photon_r(plain.trace_light(it))[1])
    ref = np.concatenate(ref)
    edges = np.quantile(np.hypot(ref[:, 0], ref[:, 1]), [0.25, 0.5, 0.75])
    p_ref = np.bincount(_bins(ref, edges), minlength=8) / len(ref)
    se_ref = np.sqrt(p_ref * (1 - p_ref) / len(ref))
    opts = {"serial": dict(rng_serial=1), "chains": dict(rng_serial=2), "one_chain": dict(rng_serial=2, mlt_chains=1)}
    chains = N // 64 if form == "chains" else 1
    o = oracle.OracleScene(args, photons=N, **opts[form])
    for it in range(2):
        ph = o.trace_light(it)
        # the visible uniform paths of the pass: those the plain estimator of the same
        # words finds visible (chain forms); the serial loop's own stream decides its own
        if form == "serial":
            _PLAIN_VISIBLE[0] = np.arange(N)  # its transient ends with the first visible path
        else:
            _PLAIN_VISIBLE[0] = _first_photon_r(plain.trace_light(it))[0]
        seqs = _chain_state_bins(ph, N, chains, edges)
        if chains == 1:
            seq = seqs[0]
            batches = np.array_split(seq[seq >= 0], 32)  # batch means absorb the chain's autocorrelation
        else:
            batches = [s[s >= 0] for s in seqs if (s >= 0).any()]
        f = np.asarray([np.bincount(bb, minlength=8) / len(bb) for bb in batches])
        w = np.asarray([len(bb) for bb in batches], dtype=float)
        mean = (f * w[:, None]).sum(0) / w.sum()
        se = np.sqrt(((f - mean) ** 2 * w[:, None] ** 2).sum(0)) / w.sum()
        z = (mean - p_ref) / np.sqrt(se ** 2 + se_ref ** 2)
        print(form, it, mean.round(4), p_ref.round(4), z.round(2))
        assert np.all(np.abs(z) < 4.5), z
        assert o.get("mlt_accepted") > 0.01 * o.get("mlt_mutations")  # mutations do contribute


def _enclosed_light_scene():
    """A small area light inside a closed sphere of nearly smooth GGX metal
    (roughness 0.011, a2 = 1.2e-4): each path meets the sphere near normal
    incidence, reflects and stores its photon at the second hit."""
    sv, sn, st = scenes.icosphere(3)
    m = np.eye(4, dtype=np.float32) * 2.0
    m[3, 3] = 1.0
    return SceneArgs([MeshData(sv, sn, st, matrix=m, material=0)],
                     [LightData("A", position=(0.0, 0.0, 0.0), width=0.01, height=0.01, power=1.0)],
                     [MaterialData("M", roughness=0.011, color=(0.9, 0.9, 0.9))],
                     look_at_camera(16, 16, (0.0, -1.0, 0.0), (0.0, 0.0, 0.0), 40.0), samples=1)


def test_every_path_visible_means_no_mutation(oracle):
    args = _enclosed_light_scene()
    n = 8192
    plain = oracle.OracleScene(args, photons=n, seed=2)
    ph_plain = plain.trace_light(0)
    assert plain.get("visible_paths") == n  # the special case holds for these words
    chains = oracle.OracleScene(args, photons=n, seed=2, rng_serial=2)
    ph = chains.trace_light(0)
    assert chains.get("mlt_mutations") == 0 and chains.get("visible_paths") == n
    # the same photons (the chains emit them in chain order), weights unscaled (visible / N = 1)
    assert np.array_equal(ph[np.argsort(ph[:, 10], kind="stable")], ph_plain[np.argsort(ph_plain[:, 10], kind="stable")])
    # the reference loop on its own serial stream: every uniform path visible on
    # this pass, so it never mutates and the visibility factor is 1
    serial = oracle.OracleScene(args, photons=4096, rng_serial=1)
    ph_s = serial.trace_light(0)
    assert serial.get("visible_paths") == 4096 and serial.get("mlt_mutations") == 0
    assert len(np.unique(ph_s[:, 10].astype(np.int64) >> 3)) == 4096
    # its photons are unscaled: first photons weigh as the plain estimator's (another stream)
    first = ph_s[(ph_s[:, 10].astype(np.int64) & 7) == 1][:, 6:9]
    first_plain = ph_plain[(ph_plain[:, 10].astype(np.int64) & 7) == 1][:, 6:9]
    assert first.mean() == pytest.approx(first_plain.mean(), rel=2e-3)
