"""ctypes binding of libRealCaustics.so, shaped like the Blender plugin's calls.

The plugin loads ``Real Caustics.dll`` and calls ``init(...)`` then ``main()``
(``src/DLL.h:14-15``).  ``Engine`` does the same against the gfx950 library and
adds the non-breaking ``rc_*`` entry points.  There is no fallback: if the
library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .blender import DebugStuff, PythonLight, PythonMaterial, PythonMatrix4x4, SceneArgs, Vec3

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libRealCaustics.so")

# Every symbol include/real_caustics.h declares.
EXPORTS = [
    "init", "main", "rc_set_option", "rc_set_option_str", "rc_get_option", "rc_prepare", "rc_render_iteration",
    "rc_get_image", "rc_write_exr", "rc_release", "rc_last_error", "rc_cancel", "rc_stats", "rc_first_hit",
    "rc_trace_photons", "rc_gather_sets", "rc_camera_pass_photons",
    "rc_guide_get_params", "rc_guide_set_params", "rc_guide_forward", "rc_guide_train_step", "rc_guide_table",
    "rc_guide_records", "rc_batch_add", "rc_batch_size", "rc_batch_prepare", "rc_batch_render_iteration",
    "rc_batch_render_scene", "rc_batch_get_image", "rc_batch_stats", "rc_batch_release", "rc_autotune",
    "rc_batch_autotune", "rc_batch_set_option", "rc_render_debug", "rc_bvh_shape",
    "rc_gather_sets_noalloc", "rc_merge_sets", "rc_write_exr_image",
]
TUNE_KEYS = ["ms_before", "ms_after", "path_chunk", "light_blocks_per_cu", "camera_blocks_per_cu", "cell_scale",
             "light_refill", "light_order", "camera_refill", "group_queries", "gather_tile", "gather_block",
             "gather_blocks_per_cu", "query_order", "overlap"]
GUIDE_PARAMS = 87232
GUIDE_BINS = 1024
STAT_NAMES = ["light_paths", "photons", "camera_paths", "gathers", "photons_tested", "photons_accepted",
              "ms_light", "ms_map", "ms_camera", "ms_total", "radius", "nodes", "tris",
              "light_node_visits", "light_tri_tests", "light_hits", "camera_node_visits", "camera_tri_tests",
              "camera_hits", "cell_visits", "sort_passes", "ms_camera_trace", "ms_gather", "ms_resolve",
              "ms_guide", "guide_visible", "guide_hit_fraction", "guide_loss", "ms_guide_step", "guide_rows",
              "ms_query_sort", "photons_streamed", "mlt_visible", "mlt_mutations", "mlt_accepted",
              "graph_captures"]
# Matrix-core FLOPs of one training step per row: forward 2*(192*64 + 2*64*64 + 64*1024), data
# gradients of the three upper layers 2*(1024*64 + 2*64*64), weight gradients 2*(64*1024 + 2*64*64 + 192*64).
GUIDE_STEP_FLOPS_PER_ROW = 2 * (192 * 64 + 2 * 64 * 64 + 64 * 1024) + 2 * (1024 * 64 + 2 * 64 * 64) + \
    2 * (64 * 1024 + 2 * 64 * 64 + 192 * 64)

_lib = None


class EngineError(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineError(f"{path} is not built; run realcausticsengine_amd/build.py (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    lib.init.restype = None
    lib.init.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, P(ctypes.c_longlong), ctypes.c_uint,
                         P(PythonMatrix4x4), P(ctypes.c_uint), P(ctypes.c_uint), ctypes.c_uint, ctypes.c_uint,
                         Vec3, Vec3, Vec3, Vec3, Vec3, P(PythonLight), ctypes.c_uint, P(PythonMaterial),
                         ctypes.c_uint, P(P(ctypes.c_int)), ctypes.c_int, DebugStuff]
    lib.main.restype = ctypes.c_int
    lib.main.argtypes = []
    lib.rc_set_option.argtypes = [ctypes.c_char_p, ctypes.c_double]
    lib.rc_set_option.restype = ctypes.c_int
    lib.rc_set_option_str.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.rc_set_option_str.restype = ctypes.c_int
    lib.rc_get_option.argtypes = [ctypes.c_char_p]
    lib.rc_get_option.restype = ctypes.c_double
    lib.rc_prepare.restype = ctypes.c_int
    lib.rc_render_iteration.restype = ctypes.c_int
    lib.rc_get_image.argtypes = [P(ctypes.c_float), ctypes.c_int]
    lib.rc_get_image.restype = ctypes.c_int
    lib.rc_write_exr.argtypes = [ctypes.c_char_p]
    lib.rc_write_exr.restype = ctypes.c_int
    lib.rc_release.restype = None
    lib.rc_last_error.restype = ctypes.c_char_p
    lib.rc_cancel.restype = None
    lib.rc_stats.argtypes = [P(ctypes.c_double), ctypes.c_int]
    lib.rc_stats.restype = ctypes.c_int
    lib.rc_first_hit.argtypes = [P(ctypes.c_float), ctypes.c_longlong, ctypes.c_float, P(ctypes.c_int),
                                 P(ctypes.c_float)]
    lib.rc_first_hit.restype = ctypes.c_int
    lib.rc_trace_photons.argtypes = [ctypes.c_int, P(ctypes.c_float), P(ctypes.c_uint), ctypes.c_longlong]
    lib.rc_trace_photons.restype = ctypes.c_longlong
    lib.rc_gather_sets.argtypes = [P(ctypes.c_float), ctypes.c_longlong, P(ctypes.c_float), ctypes.c_longlong,
                                   ctypes.c_float, P(ctypes.c_longlong), P(ctypes.c_int), ctypes.c_longlong]
    lib.rc_gather_sets.restype = ctypes.c_longlong
    lib.rc_camera_pass_photons.argtypes = [ctypes.c_int, P(ctypes.c_float), ctypes.c_longlong, P(ctypes.c_float)]
    lib.rc_camera_pass_photons.restype = ctypes.c_int
    FP, IP = P(ctypes.c_float), P(ctypes.c_int)
    lib.rc_guide_get_params.argtypes = [FP, ctypes.c_longlong]
    lib.rc_guide_set_params.argtypes = [FP, ctypes.c_longlong]
    lib.rc_guide_forward.argtypes = [FP, ctypes.c_longlong, FP]
    lib.rc_guide_train_step.argtypes = [FP, IP, FP, ctypes.c_longlong, FP, FP, ctypes.c_int]
    for fn in (lib.rc_guide_get_params, lib.rc_guide_set_params, lib.rc_guide_forward, lib.rc_guide_train_step):
        fn.restype = ctypes.c_int
    lib.rc_guide_table.argtypes = [FP, FP, ctypes.c_longlong]
    lib.rc_guide_table.restype = ctypes.c_longlong
    lib.rc_guide_records.argtypes = [FP, IP, ctypes.c_longlong]
    lib.rc_guide_records.restype = ctypes.c_longlong
    for fn in (lib.rc_batch_add, lib.rc_batch_size, lib.rc_batch_prepare, lib.rc_batch_render_iteration):
        fn.argtypes = []
        fn.restype = ctypes.c_int
    lib.rc_batch_set_option.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_double]
    lib.rc_batch_set_option.restype = ctypes.c_int
    lib.rc_batch_render_scene.argtypes = [ctypes.c_int]
    lib.rc_batch_render_scene.restype = ctypes.c_int
    lib.rc_batch_get_image.argtypes = [ctypes.c_int, FP, ctypes.c_int]
    lib.rc_batch_get_image.restype = ctypes.c_int
    lib.rc_batch_stats.argtypes = [P(ctypes.c_double), ctypes.c_int]
    lib.rc_batch_stats.restype = ctypes.c_int
    lib.rc_batch_release.restype = None
    lib.rc_render_debug.argtypes = [ctypes.c_int]
    lib.rc_render_debug.restype = ctypes.c_int
    lib.rc_gather_sets_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong,
                                           ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong]
    lib.rc_gather_sets_noalloc.restype = ctypes.c_longlong
    lib.rc_merge_sets.argtypes = [ctypes.c_int, P(ctypes.c_float), ctypes.c_longlong, P(ctypes.c_float),
                                  P(ctypes.c_longlong), P(ctypes.c_ulonglong), P(ctypes.c_ulonglong),
                                  ctypes.c_longlong, P(ctypes.c_float)]
    lib.rc_merge_sets.restype = ctypes.c_longlong
    lib.rc_write_exr_image.argtypes = [ctypes.c_char_p, P(ctypes.c_float), ctypes.c_int, ctypes.c_int]
    lib.rc_write_exr_image.restype = ctypes.c_int
    lib.rc_bvh_shape.argtypes = [P(ctypes.c_float), ctypes.c_longlong, ctypes.c_int, P(ctypes.c_int)]
    lib.rc_bvh_shape.restype = ctypes.c_int
    for fn in (lib.rc_autotune, lib.rc_batch_autotune):
        fn.argtypes = [ctypes.c_int, P(ctypes.c_double), ctypes.c_int]
        fn.restype = ctypes.c_int
    _lib = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Engine:
    """One scene in libRealCaustics (the library holds a single global scene)."""

    def __init__(self, args: SceneArgs, **options):
        self.lib = load()
        self.args = args
        for k, v in options.items():
            self.set(k, v)
        self.lib.init(*args.init_args())
        self._check(0)

    # -- errors ---------------------------------------------------------
    def error(self) -> str:
        return self.lib.rc_last_error().decode()

    def _check(self, rc, what="call"):
        err = self.error()
        if (isinstance(rc, int) and rc < 0) or err:
            raise EngineError(f"{what}: {err or rc}")
        return rc

    # -- options --------------------------------------------------------
    def set(self, key: str, value) -> None:
        if isinstance(value, str):
            r = self.lib.rc_set_option_str(key.encode(), value.encode())
        else:
            r = self.lib.rc_set_option(key.encode(), float(value))
        if r != 0:
            raise KeyError(key)

    def get(self, key: str) -> float:
        return self.lib.rc_get_option(key.encode())

    # -- rendering ------------------------------------------------------
    def prepare(self):
        self._check(self.lib.rc_prepare(), "rc_prepare")
        return self

    def render_iteration(self) -> int:
        return self._check(self.lib.rc_render_iteration(), "rc_render_iteration")

    def image(self, scaled: bool = True) -> np.ndarray:
        W, H = self.args.camera.width, self.args.camera.height
        out = np.zeros(W * H * 3, dtype=np.float32)
        self._check(self.lib.rc_get_image(_fp(out), int(scaled)), "rc_get_image")
        return out.reshape(H, W, 3)

    def main(self) -> int:
        rc = self.lib.main()
        if rc != 0:
            raise EngineError(f"main returned {rc}: {self.error()}")
        return rc

    def stats(self) -> dict:
        buf = (ctypes.c_double * len(STAT_NAMES))()
        n = self.lib.rc_stats(buf, len(STAT_NAMES))
        return {STAT_NAMES[i]: buf[i] for i in range(n)}

    # -- diagnostics ------------------------------------------------------
    def first_hit(self, rays: np.ndarray, tmin: float = 0.001):
        r = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(r)
        tri = np.zeros(n, dtype=np.int32)
        t = np.zeros(n, dtype=np.float32)
        self._check(self.lib.rc_first_hit(_fp(r), n, tmin, tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _fp(t)),
                    "rc_first_hit")
        return tri, t

    def trace_photons(self, iteration: int, cap: int = 1 << 22):
        buf = np.zeros((cap, 10), dtype=np.float32)
        uid = np.zeros(cap, dtype=np.uint32)
        n = self._check(self.lib.rc_trace_photons(iteration, _fp(buf), uid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                                  cap), "rc_trace_photons")
        if n > cap:
            return self.trace_photons(iteration, int(n))
        return buf[:n].copy(), uid[:n].copy()

    def gather_sets(self, positions: np.ndarray, queries: np.ndarray, r2: float):
        p = np.ascontiguousarray(positions, dtype=np.float32)
        q = np.ascontiguousarray(queries, dtype=np.float32)
        off = np.zeros(len(q) + 1, dtype=np.int64)
        cap = max(1, 4 * len(q))
        while True:
            idx = np.zeros(cap, dtype=np.int32)
            n = self._check(self.lib.rc_gather_sets(_fp(p), len(p), _fp(q), len(q), float(r2),
                                                    off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                                    idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cap),
                            "rc_gather_sets")
            if n <= cap:
                return off, idx[:n].copy()
            cap = int(n)

    def merge_sets(self, iteration: int, photons: np.ndarray, cap: int = 1 << 20):
        """Set fingerprints of the production merge (rc_merge_sets): per query
        its position, slot, accepted count and SplitMix64 hash sum, and r2."""
        ph = np.ascontiguousarray(photons, dtype=np.float32)
        qp = np.zeros((cap, 3), dtype=np.float32)
        sl = np.zeros(cap, dtype=np.int64)
        cnt = np.zeros(cap, dtype=np.uint64)
        hs = np.zeros(cap, dtype=np.uint64)
        r2 = np.zeros(1, dtype=np.float32)
        P = ctypes.POINTER
        n = self._check(self.lib.rc_merge_sets(iteration, _fp(ph), len(ph), _fp(qp), sl.ctypes.data_as(P(ctypes.c_longlong)),
                                               cnt.ctypes.data_as(P(ctypes.c_ulonglong)),
                                               hs.ctypes.data_as(P(ctypes.c_ulonglong)), cap, _fp(r2)),
                        "rc_merge_sets")
        if n > cap:
            return self.merge_sets(iteration, photons, int(n))
        return qp[:n], sl[:n], cnt[:n], hs[:n], float(r2[0])

    def gather_sets_noalloc(self, d_positions: int, count: int, d_queries: int, nq: int, r2: float, d_offsets: int,
                            d_indices: int, cap: int) -> int:
        """rc_gather_sets_noalloc on device addresses (ints); returns the total set size."""
        return self._check(self.lib.rc_gather_sets_noalloc(d_positions, count, d_queries, nq, float(r2), d_offsets,
                                                           d_indices, cap), "rc_gather_sets_noalloc")

    def camera_pass(self, iteration: int, photons: np.ndarray) -> np.ndarray:
        W, H = self.args.camera.width, self.args.camera.height
        img = np.zeros(W * H * 3, dtype=np.float32)
        ph = np.ascontiguousarray(photons, dtype=np.float32)
        self._check(self.lib.rc_camera_pass_photons(iteration, _fp(ph), len(ph), _fp(img)), "rc_camera_pass_photons")
        return img.reshape(H, W, 3)

    def render_debug(self, mode: int) -> np.ndarray:
        """Debug render mode 2, 3 or 4 (src/Debug_scene.cpp) into the film; returns it."""
        self._check(self.lib.rc_render_debug(int(mode)), "rc_render_debug")
        return self.image(scaled=True)

    def autotune(self, repeats: int = 2) -> dict:
        """Chooses launch shapes and queue / tile sizes for this scene (rc_autotune)."""
        buf = (ctypes.c_double * len(TUNE_KEYS))()
        self._check(self.lib.rc_autotune(repeats, buf, len(TUNE_KEYS)), "rc_autotune")
        return dict(zip(TUNE_KEYS, list(buf)))

    # -- learned emission proposal -----------------------------------------
    def guide_params(self) -> np.ndarray:
        out = np.zeros(GUIDE_PARAMS, dtype=np.float32)
        self._check(self.lib.rc_guide_get_params(_fp(out), GUIDE_PARAMS), "rc_guide_get_params")
        return out

    def guide_set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != GUIDE_PARAMS:
            raise ValueError("guide parameter vector must hold 87,232 floats")
        self._check(self.lib.rc_guide_set_params(_fp(p), GUIDE_PARAMS), "rc_guide_set_params")

    def guide_forward(self, x: np.ndarray) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        out = np.zeros((len(xx), GUIDE_BINS), dtype=np.float32)
        self._check(self.lib.rc_guide_forward(_fp(xx), len(xx), _fp(out)), "rc_guide_forward")
        return out

    def guide_train_step(self, x: np.ndarray, bins: np.ndarray, w: np.ndarray, apply: bool = False):
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        b = np.ascontiguousarray(bins, dtype=np.int32)
        ww = np.ascontiguousarray(w, dtype=np.float32)
        grad = np.zeros(GUIDE_PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.lib.rc_guide_train_step(_fp(xx), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _fp(ww),
                                                 len(xx), _fp(grad), _fp(loss), int(apply)), "rc_guide_train_step")
        return float(loss[0]), grad

    def guide_table(self):
        n = self._check(self.lib.rc_guide_table(None, None, 0), "rc_guide_table")
        cdf = np.zeros((n, GUIDE_BINS + 1), dtype=np.float32)
        inp = np.zeros((n, 6), dtype=np.float32)
        self._check(self.lib.rc_guide_table(_fp(cdf), _fp(inp), n), "rc_guide_table")
        return cdf, inp

    def guide_records(self, n: int):
        x = np.zeros((n, 6), dtype=np.float32)
        b = np.zeros(n, dtype=np.int32)
        m = self._check(self.lib.rc_guide_records(_fp(x), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n),
                        "rc_guide_records")
        return x[:m], b[:m]

    def release(self):
        self.lib.rc_release()


class Batch:
    """Several scenes rendered by the same launches (rc_batch_*)."""

    def __init__(self, scenes, options=None):
        """scenes: SceneArgs list; options: one dict of rc_set_option values per scene (or None)."""
        self.lib = load()
        self.lib.rc_batch_release()
        self.sizes = []
        for i, args in enumerate(scenes):
            for k, v in (options[i] if options else {}).items():
                if self.lib.rc_set_option(k.encode(), float(v)) != 0:
                    raise KeyError(k)
            self.lib.init(*args.init_args())
            if self.lib.rc_batch_add() < 0:
                raise EngineError(f"rc_batch_add: {self.lib.rc_last_error().decode()}")
            self.sizes.append((args.camera.width, args.camera.height))
        self._keep = list(scenes)

    def _check(self, rc, what):
        if rc < 0:
            raise EngineError(f"{what}: {self.lib.rc_last_error().decode()}")
        return rc

    def prepare(self):
        self._check(self.lib.rc_batch_prepare(), "rc_batch_prepare")
        return self

    def render_iteration(self) -> int:
        return self._check(self.lib.rc_batch_render_iteration(), "rc_batch_render_iteration")

    def set(self, scene: int, key: str, value) -> None:
        if self.lib.rc_batch_set_option(scene, key.encode(), float(value)) != 0:
            raise EngineError(f"rc_batch_set_option({scene}, {key}): {self.lib.rc_last_error().decode() or 'unknown key'}")

    def render_scene(self, scene: int) -> int:
        """One iteration of one scene with its own launches (the one-at-a-time baseline)."""
        return self._check(self.lib.rc_batch_render_scene(scene), "rc_batch_render_scene")

    def image(self, scene: int, scaled: bool = True) -> np.ndarray:
        W, H = self.sizes[scene]
        out = np.zeros(W * H * 3, dtype=np.float32)
        self._check(self.lib.rc_batch_get_image(scene, _fp(out), int(scaled)), "rc_batch_get_image")
        return out.reshape(H, W, 3)

    def stats(self) -> dict:
        buf = (ctypes.c_double * len(STAT_NAMES))()
        n = self.lib.rc_batch_stats(buf, len(STAT_NAMES))
        return {STAT_NAMES[i]: buf[i] for i in range(n)}

    def autotune(self, repeats: int = 2) -> dict:
        buf = (ctypes.c_double * len(TUNE_KEYS))()
        self._check(self.lib.rc_batch_autotune(repeats, buf, len(TUNE_KEYS)), "rc_batch_autotune")
        return dict(zip(TUNE_KEYS, list(buf)))

    def release(self):
        self.lib.rc_batch_release()


def write_exr_image(path: str, rgb: np.ndarray) -> None:
    """Writes an (H, W, 3) float image with the library's EXR writer (rc_write_exr_image; no GPU needed)."""
    img = np.ascontiguousarray(rgb, dtype=np.float32)
    H, W = img.shape[:2]
    lib = load()
    if lib.rc_write_exr_image(path.encode(), _fp(img), W, H) != 0:
        raise EngineError(f"rc_write_exr_image: {lib.rc_last_error().decode()}")


def bvh_shape(tris: np.ndarray, width: int = 4) -> dict:
    """Shape of the host-built BVH over (n, 3, 3) world-space triangles
    (rc_bvh_shape; no GPU needed): binary depth, 4-wide depth, traversal width
    kept and the stack entries it can need (ok False: the tree cannot be traversed)."""
    t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    out = np.zeros(4, dtype=np.int32)
    rc = load().rc_bvh_shape(_fp(t), len(t), int(width), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return {"depth": int(out[0]), "depth4": int(out[1]), "width": int(out[2]), "stack_need": int(out[3]),
            "ok": rc == 0}


def photons_from_oracle(records: np.ndarray) -> np.ndarray:
    """Oracle photon records (11 doubles) -> engine layout (10 floats)."""
    out = np.zeros((len(records), 10), dtype=np.float32)
    out[:, :9] = records[:, :9]
    out[:, 9] = records[:, 9]
    return out
