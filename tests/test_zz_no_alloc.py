"""rc_gather_sets_noalloc against rc_gather_sets (run on the GPU box).

The entry point takes device buffers and allocates nothing: its photon map
and per-query counts live in the library's fixed scratch array.  Two calls
in a row (the second smaller, so stale scratch from the first would show)
must give the same CSR radius sets as rc_gather_sets, bit for bit, and the
device's free memory must not move across the calls.  The test's own device
buffers come from the HIP runtime the library itself loaded."""
import ctypes

import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine, EngineError, load

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        load()  # libRealCaustics.so pulls in the runtime; this is the same library object
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        self.held = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), max(1, nbytes)) == 0
        self.held.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free_bytes(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert self.lib.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    def release(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def _inputs(seed, P, Q):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (P, 3)).astype(np.float32)
    pos[: P // 2, 2] = 0.0  # a planar layer, like photons on a floor
    q = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    q[: Q // 2, 2] = 0.0
    return pos, q


def test_noalloc_gather_sets_match_rc_gather_sets():
    args = scenes.c0_scene(16, 16, subdiv=1)
    eng = Engine(args)
    eng.prepare()
    hip = _Hip()
    r2 = np.float32(0.05) ** 2
    try:
        for call, (P, Q, seed) in enumerate([(200000, 20000, 3), (60000, 3000, 4)]):
            print(f"call {call}: {P} photons, {Q} queries", flush=True)
            pos, q = _inputs(seed, P, Q)
            off_ref, idx_ref = eng.gather_sets(pos, q, r2)
            total = int(off_ref[-1])
            d_pos, d_q = hip.upload(pos), hip.upload(q)
            d_off = hip.alloc(8 * (Q + 1))
            cap = total + 16
            d_idx = hip.upload(np.full(cap, -7, dtype=np.int32))
            free0 = hip.free_bytes()
            n = eng.gather_sets_noalloc(d_pos, P, d_q, Q, r2, d_off, d_idx, cap)
            free1 = hip.free_bytes()
            assert free1 == free0, (free0, free1)
            assert n == total
            off = hip.download(d_off, Q + 1, np.int64)
            idx = hip.download(d_idx, cap, np.int32)
            assert np.array_equal(off, off_ref)
            assert np.array_equal(idx[:total], idx_ref)  # the same grid: the same order within each set
            assert np.all(idx[total:] == -7)
            # a cap below the total: the total comes back and nothing is written
            d_small = hip.upload(np.full(8, -7, dtype=np.int32))
            assert eng.gather_sets_noalloc(d_pos, P, d_q, Q, r2, d_off, d_small, 8) == total
            assert np.all(hip.download(d_small, 8, np.int32) == -7)
        # sizes beyond the scratch array are refused with a clear error
        with pytest.raises(EngineError, match="at most 262144"):
            eng.gather_sets_noalloc(d_pos, (1 << 18) + 1, d_q, 10, r2, d_off, d_idx, cap)
    finally:
        hip.release()
        eng.release()


def test_noalloc_needs_a_prepared_scene():
    eng = Engine(scenes.c0_scene(8, 8, subdiv=1))
    with pytest.raises(EngineError, match="not prepared"):
        eng.gather_sets_noalloc(0, 0, 0, 0, 0.01, 0, 0, 0)
    eng.release()
