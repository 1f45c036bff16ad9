"""Pins of the oracle's counter RNG and its own sin/cos/acos."""
import math

import numpy as np


def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32-10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert oracle.philox(ctr, key).tolist() == want


def test_own_sin_cos_accuracy(oracle):
    xs = np.concatenate([np.linspace(-math.pi / 4, 3 * math.pi / 4, 2001), np.linspace(0, 2 * math.pi, 4001)])
    xs = xs.astype(np.float32).astype(np.float64)
    s = np.array([oracle.own_math(0, x) for x in xs])
    c = np.array([oracle.own_math(1, x) for x in xs])
    assert np.max(np.abs(s - np.sin(xs))) < 4e-7
    assert np.max(np.abs(c - np.cos(xs))) < 4e-7
    # exact special points
    assert oracle.own_math(0, 0.0) == 0.0 and oracle.own_math(1, 0.0) == 1.0


def test_own_acos_accuracy(oracle):
    xs = np.linspace(-1, 1, 8001).astype(np.float32).astype(np.float64)
    a = np.array([oracle.own_math(2, x) for x in xs])
    assert np.max(np.abs(a - np.arccos(xs))) < 5e-7
    assert oracle.own_math(2, 1.0) == 0.0
    assert abs(oracle.own_math(2, 0.0) - math.pi / 2) < 1e-7
