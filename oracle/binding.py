"""ctypes binding of the CPU oracle (test infrastructure only).

Loads ``oracle/_build/liboracle_f32.so`` (reference-mode float estimator) or
``liboracle_f64.so`` (FP64 oracle).  Only tests/, ``__graft_entry__.smoke`` and
``bench.py``'s cpu_baseline leg may import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
PHOTON_STRIDE = 11

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)
_up = ctypes.POINTER(ctypes.c_uint)

_libs = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(precision: str = "f32") -> ctypes.CDLL:
    if precision in _libs:
        return _libs[precision]
    path = os.path.join(BUILD, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_uint,
                               _fp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, _fp,
                               ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p]
    lib.orc_free.argtypes = [ctypes.c_void_p]
    lib.orc_set_option.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
    lib.orc_set_option.restype = ctypes.c_int
    lib.orc_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.orc_get.restype = ctypes.c_double
    lib.orc_prepare.argtypes = [ctypes.c_void_p]
    lib.orc_iteration_params.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
    lib.orc_trace_light.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_longlong]
    lib.orc_trace_light.restype = ctypes.c_longlong
    lib.orc_camera_pass.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_longlong, _dp]
    lib.orc_render.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    lib.orc_first_hit.argtypes = [ctypes.c_void_p, _fp, ctypes.c_longlong, ctypes.c_float, ctypes.c_int, _ip, _dp]
    for fn in (lib.orc_gather_sets, lib.orc_gather_sets_kd):
        fn.argtypes = [_fp, ctypes.c_longlong, _fp, ctypes.c_longlong, ctypes.c_float, _llp, _ip, ctypes.c_longlong]
        fn.restype = ctypes.c_longlong
    lib.orc_philox.argtypes = [_up, _up, _up]
    lib.orc_math.argtypes = [ctypes.c_int, ctypes.c_double]
    lib.orc_math.restype = ctypes.c_double
    lib.orc_bsdf.argtypes = [_fp, ctypes.c_int, _fp, _fp, _fp, _fp, _fp, ctypes.c_int, _dp]
    lib.orc_light_sample.argtypes = [ctypes.c_void_p, _fp, _dp]
    lib.orc_counters.argtypes = [ctypes.c_void_p, _llp]
    lib.orc_debug_aov.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
    lib.orc_set_guide.argtypes = [ctypes.c_void_p, _fp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_double]
    lib.orc_set_guide.restype = ctypes.c_int
    lib.orc_qmc_words.argtypes = [ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)]
    lib.orc_qmc_words.restype = None
    _libs[precision] = lib
    return lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


class OracleScene:
    """A scene in the CPU oracle, created from the same ``SceneArgs`` as the engine."""

    def __init__(self, args, precision: str = "f32", **options):
        self.lib = load(precision)
        self.args = args
        self.mats = args.matrices_flat()
        self.cam = args.camera_flat()
        self.h = self.lib.orc_create(args.samples, args.base_radius, args.radius_reduction,
                                     ctypes.cast(args.meshes_pointers, ctypes.c_void_p), args.n_meshes,
                                     _ptr(self.mats, _fp), ctypes.cast(args.n_verts, ctypes.c_void_p),
                                     ctypes.cast(args.n_tris, ctypes.c_void_p), args.camera.width,
                                     args.camera.height, _ptr(self.cam, _fp),
                                     ctypes.cast(args.lights, ctypes.c_void_p), args.n_lights,
                                     ctypes.cast(args.materials, ctypes.c_void_p), args.n_materials,
                                     ctypes.cast(args.materials_indices, ctypes.c_void_p))
        for k, v in options.items():
            self.set(k, v)
        self.prepared = False

    def set(self, key: str, value: float) -> None:
        if self.lib.orc_set_option(self.h, key.encode(), float(value)) != 0:
            raise KeyError(key)

    def get(self, key: str) -> float:
        return self.lib.orc_get(self.h, key.encode())

    def prepare(self):
        if not self.prepared:
            self.lib.orc_prepare(self.h)
            self.prepared = True
        return self

    def iteration_params(self, it: int):
        r, r2, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.lib.orc_iteration_params(self.h, it, ctypes.byref(r), ctypes.byref(r2), ctypes.byref(n))
        return r.value, r2.value, n.value

    def trace_light(self, it: int, cap: int = 1 << 22) -> np.ndarray:
        self.prepare()
        buf = np.zeros((cap, PHOTON_STRIDE), dtype=np.float64)
        n = self.lib.orc_trace_light(self.h, it, _ptr(buf, _dp), cap)
        if n > cap:
            return self.trace_light(it, int(n))
        return buf[:n].copy()

    def camera_pass(self, it: int, photons: np.ndarray, image: np.ndarray | None = None) -> np.ndarray:
        self.prepare()
        W, H = self.args.camera.width, self.args.camera.height
        img = np.zeros(W * H * 3, dtype=np.float64) if image is None else image.reshape(-1).astype(np.float64)
        ph = np.ascontiguousarray(photons, dtype=np.float64)
        self.lib.orc_camera_pass(self.h, it, _ptr(ph, _dp), len(ph), _ptr(img, _dp))
        return img.reshape(H, W, 3)

    def render(self, iterations: int) -> np.ndarray:
        self.prepare()
        W, H = self.args.camera.width, self.args.camera.height
        img = np.zeros(W * H * 3, dtype=np.float64)
        self.lib.orc_render(self.h, iterations, _ptr(img, _dp))
        return img.reshape(H, W, 3)

    def first_hit(self, rays: np.ndarray, tmin: float = 0.001, brute: bool = False):
        self.prepare()
        r = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(r)
        tri = np.zeros(n, dtype=np.int32)
        t = np.zeros(n, dtype=np.float64)
        self.lib.orc_first_hit(self.h, _ptr(r, _fp), n, tmin, int(brute), _ptr(tri, _ip), _ptr(t, _dp))
        return tri, t

    def counters(self):
        c = np.zeros(3, dtype=np.int64)
        self.lib.orc_counters(self.h, _ptr(c, _llp))
        return c

    def light_sample(self, u5) -> np.ndarray:
        self.prepare()
        u = np.ascontiguousarray(u5, dtype=np.float32)
        out = np.zeros(10, dtype=np.float64)
        self.lib.orc_light_sample(self.h, _ptr(u, _fp), _ptr(out, _dp))
        return out

    def debug_aov(self, mode: int, ior: float, brute: bool = False) -> np.ndarray:
        """Debug render mode 2, 3 or 4 of the reference (Debug_scene.cpp)."""
        self.prepare()
        W, H = self.args.camera.width, self.args.camera.height
        img = np.zeros(W * H * 3, dtype=np.float64)
        self.lib.orc_debug_aov(self.h, mode, float(ior), int(brute), _ptr(img, _dp))
        return img.reshape(H, W, 3)

    def set_guide(self, cdf, slots: int, cells: int, slot_of_light, alpha: float) -> None:
        """Emission from the engine's learned proposal table (rows of 1025-float CDFs); cdf=None: plain."""
        if cdf is None:
            self.lib.orc_set_guide(self.h, None, 0, 0, None, 0.0)
            return
        self._guide_cdf = np.ascontiguousarray(cdf, dtype=np.float32)
        sl = np.full(8, -1, dtype=np.int32)
        sl[:len(slot_of_light)] = slot_of_light
        if self.lib.orc_set_guide(self.h, _ptr(self._guide_cdf, _fp), slots, cells, _ptr(sl, _ip), alpha) != 0:
            raise MemoryError("orc_set_guide")

    def __del__(self):
        try:
            self.lib.orc_free(self.h)
        except Exception:
            pass


def gather_sets(pos: np.ndarray, queries: np.ndarray, r2: float, kd: bool = False, precision: str = "f32"):
    """CSR radius sets: (offsets[Q+1], indices)."""
    lib = load(precision)
    p = np.ascontiguousarray(pos, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    off = np.zeros(len(q) + 1, dtype=np.int64)
    cap = 1 << 20
    while True:
        idx = np.zeros(cap, dtype=np.int32)
        fn = lib.orc_gather_sets_kd if kd else lib.orc_gather_sets
        n = fn(_ptr(p, _fp), len(p), _ptr(q, _fp), len(q), float(r2), _ptr(off, _llp), _ptr(idx, _ip), cap)
        if n <= cap:
            return off, idx[:n].copy()
        cap = int(n)


def philox(ctr, key) -> np.ndarray:
    lib = load("f32")
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib.orc_philox(_ptr(c, _up), _ptr(k, _up), _ptr(out, _up))
    return out


def own_math(which: int, x: float) -> float:
    return load("f32").orc_math(which, float(x))


MAT_CODES = {"P": 80, "G": 71, "M": 77, "R": 82}  # "R": glass as the rough dielectric (compat_glass = 0)


def bsdf(mat, mode: int, shade_n, geo_n, direction, u3=(0.5, 0.5, 0.5), wi=(0.0, 0.0, 1.0), op: int = 0,
         precision: str = "f32") -> np.ndarray:
    """mat = (type, ior, roughness, specular, (r, g, b)); mode 0 radiance, 1 importance."""
    lib = load(precision)
    t, ior, rough, spec, col = mat
    m = np.asarray([MAT_CODES.get(t, ord(t)), ior, rough, spec, *col], dtype=np.float32)
    arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in (shade_n, geo_n, direction, u3, wi)]
    out = np.zeros(13, dtype=np.float64)
    lib.orc_bsdf(_ptr(m, _fp), mode, *[_ptr(a, _fp) for a in arrs], op, _ptr(out, _dp))
    return out


def qmc_words(seed: int, pixel: int, iteration: int) -> np.ndarray:
    """The three stratified camera words (option camera_qmc) of one pixel sample."""
    out = np.zeros(3, dtype=np.uint32)
    load("f32").orc_qmc_words(seed, pixel, iteration, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    return out
