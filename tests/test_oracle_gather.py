"""The restated KD tree (KD_Tree.cpp) returns exactly the brute-force radius sets."""
import numpy as np
import pytest


@pytest.mark.parametrize("case", ["uniform", "planar", "duplicates", "tiny"])
def test_kd_sets_equal_brute_force(oracle, case):
    rng = np.random.default_rng(hash(case) % 1000)
    if case == "tiny":
        pos = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
    else:
        pos = rng.uniform(-1, 1, (20000, 3)).astype(np.float32)
    if case == "planar":
        pos[:, 2] = 0.0
    if case == "duplicates":
        pos[::3] = pos[0]
    q = rng.uniform(-1, 1, (1500, 3)).astype(np.float32)
    if case == "planar":
        q[:, 2] = 0.0
    q[0] = pos[0]
    r2 = np.float32(0.07) ** 2
    ob, ib = oracle.gather_sets(pos, q, r2, kd=False)
    ok, ik = oracle.gather_sets(pos, q, r2, kd=True)
    assert np.array_equal(ob, ok)
    for i in range(len(q)):
        assert np.array_equal(np.sort(ib[ob[i]:ob[i + 1]]), np.sort(ik[ok[i]:ok[i + 1]]))


def test_gather_strict_radius(oracle):
    # a photon at exactly distance r is excluded (strict <, KD_Tree.cpp:316)
    pos = np.array([[0.5, 0, 0], [0.25, 0, 0]], dtype=np.float32)
    q = np.zeros((1, 3), dtype=np.float32)
    off, idx = oracle.gather_sets(pos, q, 0.25)
    assert idx.tolist() == [1]


def test_gather_empty(oracle):
    off, idx = oracle.gather_sets(np.zeros((0, 3), np.float32), np.zeros((4, 3), np.float32), 1.0, kd=True)
    assert off.tolist() == [0] * 5 and len(idx) == 0
