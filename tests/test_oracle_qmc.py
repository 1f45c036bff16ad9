"""Stratified camera samples (option camera_qmc) in the oracle's restatement.

Sample i of a pixel is point i of the (0,2)-sequence (van der Corput, Sobol's
second dimension) for the jitter and the base-3 radical inverse for the first
bounce's first dimension, shifted modulo 1 by a per-pixel Philox offset.
Pins: with the shift removed, the first 2^m jitter points form a (0,m,2)-net
(every 2^a x 2^b box with a + b = m holds exactly one point) and the first
3^k third coordinates hit every 1/3^k interval once; the shift is the same for
every iteration of a pixel and differs between pixels; the unshifted third
coordinate matches the radical inverse computed in exact fractions."""
from fractions import Fraction

import numpy as np

from oracle import binding

SEED = 12345


def _unshifted(pixel, n):
    key = np.array([SEED & 0xFFFFFFFF, SEED >> 32], dtype=np.uint32)
    sh = binding.philox(np.array([0, pixel, 0xFFFFFFFF, 4], dtype=np.uint32), key) >> 8
    w = np.array([binding.qmc_words(SEED, pixel, i) for i in range(n)], dtype=np.uint64) >> 8
    return (w - sh[:3].astype(np.uint64)) & 0xFFFFFF, sh


def test_jitter_is_a_02_sequence(oracle):
    for pixel in (0, 777):
        v, _ = _unshifted(pixel, 1 << 10)
        for m in range(1, 11):
            pts = v[: 1 << m, :2]
            for a in range(m + 1):
                b = m - a
                cx = (pts[:, 0] >> (24 - a)) if a else np.zeros(len(pts), dtype=np.uint64)
                cy = (pts[:, 1] >> (24 - b)) if b else np.zeros(len(pts), dtype=np.uint64)
                cells = cx * (1 << b) + cy
                assert len(np.unique(cells)) == 1 << m, (pixel, m, a)


def test_third_dimension_is_the_base3_radical_inverse(oracle):
    v, _ = _unshifted(5, 243)
    for i in range(243):
        x, f, d = i, Fraction(0), Fraction(1, 3)
        while x:
            f += (x % 3) * d
            x //= 3
            d /= 3
        assert v[i, 2] == int(f * (1 << 24)), i
    for k in range(1, 6):  # the first 3^k points are the multiples j / 3^k, each once (24-bit floors)
        j = np.round(v[: 3 ** k, 2].astype(np.float64) * 3 ** k / (1 << 24)).astype(np.int64)
        assert np.array_equal(np.sort(j), np.arange(3 ** k))


def test_shift_is_per_pixel(oracle):
    _, s0 = _unshifted(3, 1)
    _, s1 = _unshifted(4, 1)
    assert not np.array_equal(s0[:3], s1[:3])
    a = binding.qmc_words(SEED, 3, 0)
    b = binding.qmc_words(SEED, 3, 0)
    assert np.array_equal(a, b)
