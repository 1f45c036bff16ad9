"""Byte layouts of the structures the Blender plugin hands to ``init``.

These mirror, field for field, the types the reference's C-ABI takes
(``src/DLL.h:14-15``, ``src/main.cpp:89-108``):

* ``vec3`` / ``colorf``: three floats (``src/vec3.h:13-19``, ``src/Color.h:5-10``);
* ``PythonMatrix4x4``: 16 floats, row-major ``t00..t33`` (``src/matrix.h:151-171``);
* ``Python_Light``: 64 bytes (``src/Lights.h:15-36``);
* ``Python_Material``: 28 bytes (``src/materials.h:21-32``);
* ``Debug_stuff``: 24 bytes (``src/Vertex_Merging.h:13-19``);
* Blender ``Mesh``: 1776 bytes of which only ``mpoly@200``, ``mloop@208`` and
  ``mvert@256`` are read (``src/Blender_definitions.h:214-311``, ``src/mesh.cpp:16-26``);
  ``MVert`` 20 B, ``MPoly`` 12 B, ``MLoop`` 8 B (``src/Blender_definitions.h:183-212``).

``SceneArgs`` owns every buffer so that the borrowed pointers (mesh vertices and
the light array, ``src/mesh.cpp:16`` and ``src/main.cpp:134``) stay valid until
the render returns.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

MESH_STRUCT_SIZE = 1776
MESH_OFF_MPOLY = 200
MESH_OFF_MLOOP = 208
MESH_OFF_MVERT = 256

MVERT_DTYPE = np.dtype([("co", "<f4", 3), ("no", "<i2", 3), ("flag", "i1"), ("bweight", "i1")])
MPOLY_DTYPE = np.dtype([("loopstart", "<i4"), ("totloop", "<i4"), ("mat_nr", "<i2"), ("flag", "i1"), ("pad", "i1")])
MLOOP_DTYPE = np.dtype([("v", "<u4"), ("e", "<u4")])
assert MVERT_DTYPE.itemsize == 20 and MPOLY_DTYPE.itemsize == 12 and MLOOP_DTYPE.itemsize == 8


class Vec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class PythonMatrix4x4(ctypes.Structure):
    _fields_ = [(f"t{r}{c}", ctypes.c_float) for r in range(4) for c in range(4)]


class PythonLight(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_char),
        ("color", ctypes.c_float * 3),
        ("position", ctypes.c_float * 3),
        ("width", ctypes.c_float),
        ("height", ctypes.c_float),
        ("spread", ctypes.c_float),
        ("rotation", ctypes.c_float * 3),
        ("power", ctypes.c_float),
        ("angle", ctypes.c_float),
        ("radius", ctypes.c_float),
    ]


class PythonMaterial(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_char),
        ("ior", ctypes.c_float),
        ("roughness", ctypes.c_float),
        ("specular", ctypes.c_float),
        ("color", ctypes.c_float * 3),
    ]


class DebugStuff(ctypes.Structure):
    _fields_ = [
        ("number_of_points", ctypes.c_int),
        ("direction", Vec3),
        ("roughness", ctypes.c_float),
        ("specular", ctypes.c_float),
    ]


assert ctypes.sizeof(PythonLight) == 64
assert ctypes.sizeof(PythonMaterial) == 28
assert ctypes.sizeof(DebugStuff) == 24
assert ctypes.sizeof(PythonMatrix4x4) == 64


@dataclass
class MeshData:
    """One triangle mesh in object space with its world matrix."""

    verts: np.ndarray  # (V, 3) float32
    normals: np.ndarray  # (V, 3) float, unit
    tris: np.ndarray  # (T, 3) uint32
    matrix: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))
    material: int = 0  # global material index of slot 0
    mat_nr: np.ndarray | None = None  # per-triangle slot (default 0)
    slots: Sequence[int] | None = None  # slot -> global material (default [material])


@dataclass
class LightData:
    type: str = "A"
    color: Sequence[float] = (1.0, 1.0, 1.0)
    position: Sequence[float] = (0.0, 0.0, 5.0)
    width: float = 1.0
    height: float = 1.0
    spread: float = 0.0
    rotation: Sequence[float] = (0.0, 0.0, 0.0)
    power: float = 1.0
    angle: float = 0.0
    radius: float = 0.0


@dataclass
class MaterialData:
    type: str = "P"
    ior: float = 1.5
    roughness: float = 1.0
    specular: float = 0.0
    color: Sequence[float] = (0.8, 0.8, 0.8)


@dataclass
class CameraData:
    width: int
    height: int
    position: Sequence[float]
    corners: Sequence[Sequence[float]]  # top right, bottom right, bottom left, top left


def look_at_camera(width: int, height: int, origin, target, fov_deg: float, up=(0.0, 0.0, 1.0)) -> CameraData:
    """Pinhole camera corners at unit distance (the four corners Blender passes)."""
    o = np.asarray(origin, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - o
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, dtype=np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    hh = np.tan(np.radians(fov_deg) * 0.5)
    hw = hh * width / height
    c = o + f
    tr = c + r * hw + u * hh
    br = c + r * hw - u * hh
    bl = c - r * hw - u * hh
    tl = c - r * hw + u * hh
    return CameraData(width, height, tuple(o), [tuple(tr), tuple(br), tuple(bl), tuple(tl)])


class SceneArgs:
    """Every argument of ``init`` as ctypes objects, with the buffers kept alive."""

    def __init__(self, meshes: List[MeshData], lights: List[LightData], materials: List[MaterialData],
                 camera: CameraData, samples: int = 1, base_radius: float = 0.01,
                 radius_reduction: float = 0.75, debug_test: int = 1):
        self.meshes = meshes
        self.samples = int(samples)
        self.base_radius = float(base_radius)
        self.radius_reduction = float(radius_reduction)
        self.debug_test = int(debug_test)
        self.camera = camera
        self._keep = []
        n = len(meshes)
        self.mesh_structs = []
        ptrs = (ctypes.c_longlong * max(n, 1))()
        self.mesh_matrices = (PythonMatrix4x4 * max(n, 1))()
        self.n_verts = (ctypes.c_uint * max(n, 1))()
        self.n_tris = (ctypes.c_uint * max(n, 1))()
        self.mat_index_arrays = []
        mat_ptrs = (ctypes.POINTER(ctypes.c_int) * max(n, 1))()
        for i, m in enumerate(meshes):
            verts = np.ascontiguousarray(m.verts, dtype=np.float32)
            nrm = np.asarray(m.normals, dtype=np.float64)
            mv = np.zeros(len(verts), dtype=MVERT_DTYPE)
            mv["co"] = verts
            mv["no"] = np.clip(np.rint(nrm * 32767.0), -32767, 32767).astype(np.int16)
            tris = np.ascontiguousarray(m.tris, dtype=np.uint32)
            nt = len(tris)
            ml = np.zeros(3 * nt, dtype=MLOOP_DTYPE)
            ml["v"] = tris.reshape(-1)
            mp = np.zeros(nt, dtype=MPOLY_DTYPE)
            mp["loopstart"] = np.arange(nt, dtype=np.int32) * 3
            mp["totloop"] = 3
            if m.mat_nr is not None:
                mp["mat_nr"] = np.asarray(m.mat_nr, dtype=np.int16)
            blk = np.zeros(MESH_STRUCT_SIZE, dtype=np.uint8)
            blk[MESH_OFF_MPOLY:MESH_OFF_MPOLY + 8] = np.frombuffer(np.uint64(mp.ctypes.data).tobytes(), np.uint8)
            blk[MESH_OFF_MLOOP:MESH_OFF_MLOOP + 8] = np.frombuffer(np.uint64(ml.ctypes.data).tobytes(), np.uint8)
            blk[MESH_OFF_MVERT:MESH_OFF_MVERT + 8] = np.frombuffer(np.uint64(mv.ctypes.data).tobytes(), np.uint8)
            self._keep += [mv, ml, mp, blk]
            ptrs[i] = blk.ctypes.data
            mat = np.asarray(m.matrix, dtype=np.float32).reshape(4, 4)
            for r in range(4):
                for c in range(4):
                    setattr(self.mesh_matrices[i], f"t{r}{c}", float(mat[r, c]))
            self.n_verts[i] = len(verts)
            self.n_tris[i] = nt
            slots = list(m.slots) if m.slots is not None else [m.material]
            arr = (ctypes.c_int * len(slots))(*slots)
            self.mat_index_arrays.append(arr)
            mat_ptrs[i] = ctypes.cast(arr, ctypes.POINTER(ctypes.c_int))
        self.meshes_pointers = ptrs
        self.materials_indices = mat_ptrs
        self.lights = (PythonLight * max(len(lights), 1))()
        for i, l in enumerate(lights):
            L = self.lights[i]
            L.type = l.type.encode()
            L.color[:] = list(l.color)
            L.position[:] = list(l.position)
            L.width, L.height, L.spread = l.width, l.height, l.spread
            L.rotation[:] = list(l.rotation)
            L.power, L.angle, L.radius = l.power, l.angle, l.radius
        self.n_lights = len(lights)
        self.materials = (PythonMaterial * max(len(materials), 1))()
        for i, m in enumerate(materials):
            M = self.materials[i]
            M.type = m.type.encode()
            M.ior, M.roughness, M.specular = m.ior, m.roughness, m.specular
            M.color[:] = list(m.color)
        self.n_materials = len(materials)
        self.debug_stuff = DebugStuff()
        self.cam_position = Vec3(*camera.position)
        self.cam_corners = [Vec3(*c) for c in camera.corners]

    @property
    def n_meshes(self) -> int:
        return len(self.meshes)

    @property
    def total_tris(self) -> int:
        return int(sum(len(m.tris) for m in self.meshes))

    def matrices_flat(self) -> np.ndarray:
        out = np.zeros((max(self.n_meshes, 1), 16), dtype=np.float32)
        for i in range(self.n_meshes):
            out[i] = [getattr(self.mesh_matrices[i], f"t{r}{c}") for r in range(4) for c in range(4)]
        return out

    def camera_flat(self) -> np.ndarray:
        vals = [self.camera.position] + list(self.camera.corners)
        return np.asarray(vals, dtype=np.float32).reshape(-1)

    def init_args(self):
        """Positional arguments of ``init`` in the reference's order (main.cpp:89-108)."""
        return (
            self.samples, self.base_radius, self.radius_reduction,
            self.meshes_pointers, self.n_meshes, self.mesh_matrices, self.n_verts, self.n_tris,
            self.camera.width, self.camera.height, self.cam_position,
            self.cam_corners[0], self.cam_corners[1], self.cam_corners[2], self.cam_corners[3],
            self.lights, self.n_lights, self.materials, self.n_materials, self.materials_indices,
            self.debug_test, self.debug_stuff,
        )
