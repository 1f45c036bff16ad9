"""The production merge accepts exactly the brute-force radius sets.

rc_merge_sets runs the camera pass against a given photon list with the merge
kernel the options select (the grouped merge by default: queries sharing a
photon-map cell window walk its photons together) in its counting form, and
returns per merge query its position, the number of photons the merge
accepted and the sum of SplitMix64(input index) over them.  The sets of
find_closest_in_radius (src/KD_Tree.cpp:217-280: every photon with
fl(fl(dx^2 + dy^2) + dz^2) < r^2) are computed by brute force on the CPU from
the same query positions and r^2 bits; count and hash sum must be equal for
every query.  The scene carries every BSDF class, so both the diffuse-only
form and the general form of the merge run."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine, photons_from_oracle

pytestmark = pytest.mark.gpu


def mix64(x):
    with np.errstate(over="ignore"):
        z = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def test_mix64_known_answer():
    # first output of SplitMix64 seeded with 0 (Steele, Lea and Flood 2014)
    assert int(mix64(0)) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("group_queries,gather_block,exact", [(8, 2, 0), (16, 0, 1), (4, 4, 0), (0, 2, 0), (1, 2, 0),
                                                            (1, 2, 1)])
def test_merge_sets_equal_brute_force(oracle, group_queries, gather_block, exact):
    args = scenes.materials_scene(96, 64)
    args.base_radius = 0.05  # sets of tens of photons
    orc = oracle.OracleScene(args, photons=131072, seed=3)
    ph = orc.trace_light(0)
    eng = Engine(args, photons=131072, seed=3, group_queries=group_queries, gather_block=gather_block,
                 exact_merge=exact)
    qp, slot, cnt, hs, r2 = eng.merge_sets(0, photons_from_oracle(ph))
    eng.release()
    assert len(qp) > 2000 and len(np.unique(slot)) == len(slot)
    off, idx = oracle.gather_sets(ph[:, :3].astype(np.float32), qp, np.float32(r2))
    want_cnt = np.diff(off).astype(np.uint64)
    assert want_cnt.sum() > 10 * len(qp)  # the sets are not trivially empty
    h = mix64(idx)
    csum = np.concatenate([[np.uint64(0)], np.cumsum(h, dtype=np.uint64)])
    with np.errstate(over="ignore"):
        want_hash = csum[off[1:]] - csum[off[:-1]]
    print(f"queries {len(qp)}, accepted {int(want_cnt.sum())}", flush=True)
    assert np.array_equal(cnt, want_cnt)
    assert np.array_equal(hs, want_hash)
