"""Seeded synthetic scenes for the configurations the engine is measured on.

* ``c0_scene``: glass icosphere (subdivision 4, 5120 triangles, radius 1) above a
  10 x 10 diffuse plane; light at height 5 (a small area light in the parity
  variant, since the reference point light emits nothing: ``src/Lights.h:78-81``);
  256 x 256 pinhole camera.
* ``water_scene``: 512 x 512-vertex heightfield (522,242 triangles) made of 32
  seeded sine waves, IOR 1.33, over a diffuse floor 0.5 below, 2 x 2 area light,
  1024 x 1024 camera.
* small scenes used by the tests (``slab_scene``, ``two_sphere_scene``).
"""
from __future__ import annotations

import numpy as np

from .blender import CameraData, LightData, MaterialData, MeshData, SceneArgs, look_at_camera


def icosphere(subdiv: int):
    """Unit icosphere: 20 * 4**subdiv triangles, normals equal to positions."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.asarray(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.asarray(verts, dtype=np.float64)
    return V.astype(np.float32), V, np.asarray(faces, dtype=np.uint32)


def plane(size_x: float, size_y: float, z: float = 0.0, cx: float = 0.0, cy: float = 0.0):
    hx, hy = size_x / 2.0, size_y / 2.0
    verts = np.asarray([(cx - hx, cy - hy, z), (cx + hx, cy - hy, z), (cx + hx, cy + hy, z), (cx - hx, cy + hy, z)],
                       dtype=np.float32)
    normals = np.tile(np.asarray([0.0, 0.0, 1.0]), (4, 1))
    tris = np.asarray([(0, 1, 2), (0, 2, 3)], dtype=np.uint32)
    return verts, normals, tris


def box(lo, hi):
    """Closed axis-aligned box with flat per-face vertices (24 verts, 12 tris)."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    verts, normals, tris = [], [], []
    for axis in range(3):
        for side in (0, 1):
            n = np.zeros(3)
            n[axis] = 1.0 if side else -1.0
            u, w = [a for a in range(3) if a != axis]
            quad = []
            for a, b in ((0, 0), (1, 0), (1, 1), (0, 1)):
                p = np.where(np.arange(3) == axis, hi if side else lo, 0.0).astype(np.float64)
                p[u] = hi[u] if a else lo[u]
                p[w] = hi[w] if b else lo[w]
                quad.append(p)
            base = len(verts)
            verts += quad
            normals += [n] * 4
            if side:
                tris += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
            else:
                tris += [(base, base + 2, base + 1), (base, base + 3, base + 2)]
    return (np.asarray(verts, dtype=np.float32), np.asarray(normals), np.asarray(tris, dtype=np.uint32))


def heightfield(n: int = 512, waves: int = 32, seed: int = 1, total_amp: float = 0.05, extent: float = 1.0):
    """Sum of seeded sine waves on [-extent, extent]^2 with analytic normals."""
    rng = np.random.default_rng(seed)
    lam = rng.uniform(0.05, 1.0, waves)
    amp = lam / lam.sum() * total_amp
    theta = rng.uniform(0.0, 2 * np.pi, waves)
    phase = rng.uniform(0.0, 2 * np.pi, waves)
    xs = np.linspace(-extent, extent, n)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    Z = np.zeros_like(X)
    dZx = np.zeros_like(X)
    dZy = np.zeros_like(X)
    for k in range(waves):
        kx = 2 * np.pi / lam[k] * np.cos(theta[k])
        ky = 2 * np.pi / lam[k] * np.sin(theta[k])
        arg = kx * X + ky * Y + phase[k]
        Z += amp[k] * np.sin(arg)
        c = amp[k] * np.cos(arg)
        dZx += c * kx
        dZy += c * ky
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    nrm = np.stack([-dZx, -dZy, np.ones_like(Z)], axis=-1).reshape(-1, 3)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    i = np.arange(n - 1)
    I, J = np.meshgrid(i, i, indexing="xy")
    a = (J * n + I).reshape(-1)
    b = a + 1
    c = a + n
    d = c + 1
    tris = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)]).astype(np.uint32)
    return verts, nrm, tris


def translation(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = (x, y, z)
    return m


def c0_scene(width: int = 256, height: int = 256, subdiv: int = 4, parity_light: bool = True,
             samples: int = 16, base_radius: float = 0.01, alpha: float = 0.75) -> SceneArgs:
    sv, sn, st = icosphere(subdiv)
    pv, pn, pt = plane(10.0, 10.0, 0.0)
    meshes = [MeshData(sv, sn, st, matrix=translation(0.0, 0.0, 1.25), material=0),
              MeshData(pv, pn, pt, material=1)]
    mats = [MaterialData("G", ior=1.5, roughness=0.0, color=(1.0, 1.0, 1.0)),
            MaterialData("P", roughness=1.0, specular=0.0, color=(0.8, 0.8, 0.8))]
    if parity_light:
        lights = [LightData("A", position=(0.0, 0.0, 5.0), width=0.05, height=0.05, power=1.0)]
    else:
        lights = [LightData("P", position=(0.0, 0.0, 5.0), power=1.0)]
    cam = look_at_camera(width, height, (0.0, -6.0, 4.0), (0.0, 0.0, 0.5), 40.0)
    return SceneArgs(meshes, lights, mats, cam, samples=samples, base_radius=base_radius, radius_reduction=alpha)


def water_scene(n: int = 512, width: int = 1024, height: int = 1024, roughness: float = 0.0, samples: int = 64,
                base_radius: float = 0.01, alpha: float = 0.75, seed: int = 1) -> SceneArgs:
    hv, hn, ht = heightfield(n, 32, seed)
    fv, fn, ft = plane(4.0, 4.0, -0.5)
    meshes = [MeshData(hv, hn, ht, material=0), MeshData(fv, fn, ft, material=1)]
    mats = [MaterialData("G", ior=1.33, roughness=roughness, color=(1.0, 1.0, 1.0)),
            MaterialData("P", roughness=1.0, specular=0.0, color=(0.8, 0.8, 0.8))]
    lights = [LightData("A", position=(0.0, 0.0, 2.0), width=2.0, height=2.0, power=1.0)]
    cam = look_at_camera(width, height, (0.0, -1.6, 1.6), (0.0, 0.0, 0.0), 40.0)
    return SceneArgs(meshes, lights, mats, cam, samples=samples, base_radius=base_radius, radius_reduction=alpha)


def slab_scene(width: int = 64, height: int = 64, albedo: float = 0.5, light_size: float = 1.0,
               light_height: float = 5.0, samples: int = 1) -> SceneArgs:
    """Glass slab over a diffuse plane, camera between them looking straight down.

    Used for the closed-form photon-density and camera-estimate checks."""
    gv, gn, gt = box((-3.0, -3.0, 1.0), (3.0, 3.0, 1.2))
    pv, pn, pt = plane(10.0, 10.0, 0.0)
    meshes = [MeshData(gv, gn, gt, material=0), MeshData(pv, pn, pt, material=1)]
    mats = [MaterialData("G", ior=1.5, color=(1.0, 1.0, 1.0)),
            MaterialData("P", roughness=1.0, specular=0.0, color=(albedo, albedo, albedo))]
    lights = [LightData("A", position=(0.0, 0.0, light_height), width=light_size, height=light_size, power=1.0)]
    cam = look_at_camera(width, height, (0.0, 0.0, 0.5), (0.0, 0.0, 0.0), 30.0, up=(0.0, 1.0, 0.0))
    return SceneArgs(meshes, lights, mats, cam, samples=samples, base_radius=0.01, radius_reduction=0.75)


def random_mesh_scene(seed: int = 0, n_spheres: int = 3, subdiv: int = 2, width: int = 32, height: int = 32):
    """Several separated spheres of every material kind over a plane (for traversal tests)."""
    rng = np.random.default_rng(seed)
    meshes = []
    mats = [MaterialData("G", ior=1.45, color=(0.9, 0.95, 1.0)),
            MaterialData("P", roughness=1.0, specular=0.0, color=(0.8, 0.7, 0.6)),
            MaterialData("M", roughness=0.0, color=(0.9, 0.9, 0.9)),
            MaterialData("M", roughness=0.4, color=(0.9, 0.8, 0.7)),
            MaterialData("P", roughness=0.05, specular=0.5, color=(0.5, 0.6, 0.7)),
            MaterialData("P", roughness=0.5, specular=0.5, color=(0.3, 0.6, 0.3))]
    for k in range(n_spheres):
        sv, sn, st = icosphere(subdiv)
        m = np.eye(4, dtype=np.float32) * 0.6
        m[3, 3] = 1.0
        m[:3, 3] = (3.0 * (k - (n_spheres - 1) / 2.0), rng.uniform(-0.5, 0.5), 0.8)
        meshes.append(MeshData(sv, sn, st, matrix=m, material=int(k % len(mats))))
    pv, pn, pt = plane(12.0, 12.0, 0.0)
    meshes.append(MeshData(pv, pn, pt, material=1))
    lights = [LightData("A", position=(0.0, 0.0, 6.0), width=2.0, height=1.0, power=2.0, rotation=(10.0, 0.0, 20.0))]
    cam = look_at_camera(width, height, (0.0, -8.0, 5.0), (0.0, 0.0, 0.5), 40.0)
    return SceneArgs(meshes, lights, mats, cam, samples=1)


def materials_scene(width: int = 48, height: int = 32, subdiv: int = 2, lights=None,
                    glass_roughness: float = 0.0) -> SceneArgs:
    """One sphere of each BSDF class of the reference in a row over a glossy floor.

    Spheres, left to right: glass (``Fresnel_Specular``), mirror metal
    (``Specular_Reflection``, roughness < 0.01), rough metal
    (``Microfacet_Reflection_GGX``), glossy plastic with specular > 0
    (``Fresnel_Microfacet_GGX``), smooth plastic (``Fresnel_Specular_Microfacet_GGX``,
    roughness^2 < 0.01) and a diffuse plastic (specular 0); the floor is a
    glossy plastic with specular > 0, so merge queries on it take the general
    BSDF evaluation (``src/materials.h:59-146``).  A 2 x 1 area light rotated
    by (10, 0, 20) degrees sits above, unless `lights` replaces it.  The
    glass sphere takes `glass_roughness` (a rough dielectric with option
    compat_glass = 0; the reference ignores it)."""
    mats = [MaterialData("G", ior=1.45, roughness=glass_roughness, color=(0.9, 0.95, 1.0)),
            MaterialData("M", roughness=0.0, color=(0.9, 0.9, 0.9)),
            MaterialData("M", roughness=0.35, color=(0.9, 0.8, 0.7)),
            MaterialData("P", roughness=0.5, specular=0.6, color=(0.3, 0.6, 0.3)),
            MaterialData("P", roughness=0.05, specular=0.5, color=(0.5, 0.6, 0.7)),
            MaterialData("P", roughness=1.0, specular=0.0, color=(0.8, 0.7, 0.6)),
            MaterialData("P", roughness=0.6, specular=0.8, color=(0.7, 0.7, 0.7))]
    sv, sn, st = icosphere(subdiv)
    meshes = []
    for k in range(6):
        m = np.eye(4, dtype=np.float32) * 0.7
        m[3, 3] = 1.0
        m[:3, 3] = (1.8 * (k - 2.5), 0.3 * (k % 2), 0.75)
        meshes.append(MeshData(sv, sn, st, matrix=m, material=k))
    pv, pn, pt = plane(14.0, 8.0, 0.0)
    meshes.append(MeshData(pv, pn, pt, material=6))
    if lights is None:
        lights = [LightData("A", position=(0.0, 0.0, 5.0), width=2.0, height=1.0, power=2.0,
                            rotation=(10.0, 0.0, 20.0))]
    cam = look_at_camera(width, height, (0.0, -8.0, 4.5), (0.0, 0.0, 0.5), 50.0)
    return SceneArgs(meshes, lights, mats, cam, samples=1, base_radius=0.01, radius_reduction=0.75)


def mixed_lights() -> list:
    """A rotated area light, a sun, a point light and a sphere light of
    different powers and colours (``src/Lights.cpp``, ``src/Lights.h:78-81``)."""
    return [LightData("A", position=(-1.0, 0.5, 5.0), width=1.5, height=0.8, power=1.5, color=(1.0, 0.9, 0.8),
                      rotation=(15.0, -10.0, 30.0)),
            LightData("S", power=0.4, color=(1.0, 1.0, 0.9), rotation=(0.35, -0.25, 0.6)),
            LightData("P", position=(1.5, -0.5, 4.0), power=3.0, color=(0.8, 0.9, 1.0)),
            LightData("P", position=(0.5, 1.0, 3.5), power=2.0, color=(0.9, 0.7, 1.0), radius=0.3)]


_ICO_CACHE: dict = {}


def _icosphere_cached(subdiv: int):
    if subdiv not in _ICO_CACHE:
        _ICO_CACHE[subdiv] = icosphere(subdiv)
    return _ICO_CACHE[subdiv]


def displaced_sphere(subdiv: int, seed: int, amplitude: float = 0.1, waves: int = 8):
    """Icosphere with a radial displacement: amplitude * mean of `waves` seeded
    spherical sines sin(f (n . d) + phase); normals from the displaced faces."""
    v0, _, tris = _icosphere_cached(subdiv)
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(waves, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f = rng.uniform(2.0, 8.0, waves)
    ph = rng.uniform(0.0, 2.0 * np.pi, waves)
    n = v0.astype(np.float64)
    r = 1.0 + amplitude * np.mean(np.sin((n @ d.T) * f + ph), axis=1)
    verts = n * r[:, None]
    a, b, c = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    fn = np.cross(b - a, c - a)  # area-weighted face normals
    vn = np.zeros_like(verts)
    for k in range(3):
        np.add.at(vn, tris[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    return verts.astype(np.float32), vn, tris


def blob_scene(seed: int, samples: int = 1, res_scale: int = 1, photon_scale: int = 1):
    """One scene of the uneven batch: a seeded displaced glass blob over a
    diffuse plane.  The triangle count is drawn log-uniform in [1e3, 5e5] and
    met by a displaced cube sphere of m x m quads per face (12 m^2 triangles,
    the nearest such count); IOR U(1.3, 1.7); GGX roughness U(0, 0.3) (a rough
    dielectric with option compat_glass = 0; the reference renders glass as
    delta); resolution from {256, 512, 1024}^2 / res_scale; 1..8 x 2^20 /
    photon_scale light paths per iteration.  Returns (SceneArgs, light paths)."""
    rng = np.random.default_rng(1000 + seed)
    want = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
    m = min(204, max(10, int(round(np.sqrt(want / 12.0)))))  # 1,200 .. 499,392 triangles
    ior = float(rng.uniform(1.3, 1.7))
    rough = float(rng.uniform(0.0, 0.3))
    res = int(rng.choice([256, 512, 1024])) // res_scale
    photons = (int(rng.integers(1, 9)) << 20) // photon_scale
    bv, bn, bt = displaced_cube_sphere(m, seed)
    pv, pn, pt = plane(10.0, 10.0, 0.0)
    meshes = [MeshData(bv, bn, bt, matrix=translation(0.0, 0.0, 1.3), material=0), MeshData(pv, pn, pt, material=1)]
    mats = [MaterialData("G", ior=ior, roughness=rough, color=(1.0, 1.0, 1.0)),
            MaterialData("P", roughness=1.0, specular=0.0, color=(0.8, 0.8, 0.8))]
    lights = [LightData("A", position=(0.0, 0.0, 5.0), width=1.0, height=1.0, power=1.0)]
    cam = look_at_camera(res, res, (0.0, -6.0, 4.0), (0.0, 0.0, 0.8), 40.0)
    return SceneArgs(meshes, lights, mats, cam, samples=samples, base_radius=0.01, radius_reduction=0.75), photons


def uneven_batch(n: int = 64, samples: int = 1, res_scale: int = 1, photon_scale: int = 1):
    """The uneven-scene batch: blob_scene(seed) for seeds 0 .. n-1."""
    return [blob_scene(s, samples, res_scale, photon_scale) for s in range(n)]


def displaced_cube_sphere(m: int, seed: int, amplitude: float = 0.1, waves: int = 8):
    """Unit sphere made of the six faces of a cube, m x m quads per face
    (12 m^2 triangles, any count, unlike icosphere levels), radially displaced
    by amplitude * mean of `waves` seeded spherical sines sin(f (n . d) + phase).
    Face-edge vertices come from the same grid values on both faces, so the
    seams close exactly; normals are the analytic normals of the displaced
    surface, n - grad_t(r) / r (continuous across seams)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(waves, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f = rng.uniform(2.0, 8.0, waves)
    ph = rng.uniform(0.0, 2.0 * np.pi, waves)
    g = np.linspace(-1.0, 1.0, m + 1)
    U, V = np.meshgrid(g, g, indexing="xy")
    i = np.arange(m)
    I, J = np.meshgrid(i, i, indexing="xy")
    a = (J * (m + 1) + I).reshape(-1)
    b, c = a + 1, a + (m + 1)
    dd = c + 1
    verts, norms, tris = [], [], []
    base = 0
    for axis in range(3):
        for sign in (-1.0, 1.0):
            P = np.empty((m + 1, m + 1, 3))
            P[..., axis] = sign
            P[..., (axis + 1) % 3] = U
            P[..., (axis + 2) % 3] = V
            n = (P / np.sqrt((P * P).sum(-1, keepdims=True))).reshape(-1, 3)
            arg = (n @ d.T) * f + ph
            r = 1.0 + amplitude * np.mean(np.sin(arg), axis=1)
            grad = amplitude * (np.cos(arg) * f) @ d / waves
            grad_t = grad - (grad * n).sum(1, keepdims=True) * n
            nn = n - grad_t / r[:, None]
            nn /= np.linalg.norm(nn, axis=1, keepdims=True)
            verts.append((n * r[:, None]).astype(np.float32))
            norms.append(nn)
            if sign > 0:  # e_(axis+1) x e_(axis+2) = e_axis: outward winding
                t = np.concatenate([np.stack([a, b, dd], 1), np.stack([a, dd, c], 1)])
            else:
                t = np.concatenate([np.stack([a, dd, b], 1), np.stack([a, c, dd], 1)])
            tris.append((t + base).astype(np.uint32))
            base += (m + 1) * (m + 1)
    return np.concatenate(verts), np.concatenate(norms), np.concatenate(tris)


def _loguniform(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def c4_scene(seed: int, res: int = 2048, photons: int = 64 << 20, samples: int = 1):
    """One scene of the at-scale batch: a seeded displaced glass mesh (even
    seeds) or water heightfield (odd seeds) of 2M-10M triangles (log-uniform),
    IOR U(1.3, 1.7) for glass and 1.33 for water, roughness U(0, 0.2) (glass is
    a delta interface in the reference: recorded, not used), over a diffuse
    floor; four lights -- a point light, a quad, a sphere and one more of a
    random kind among those three -- with log-uniform powers in [0.25, 4];
    res^2 camera; `photons` light paths per iteration.  Point lights emit only
    with compat_point_light = 0 (isotropic emission).  Returns (SceneArgs, photons)."""
    rng = np.random.default_rng(4000 + seed)
    want = _loguniform(rng, 2e6, 1e7)
    rough = float(rng.uniform(0.0, 0.2))
    if seed % 2 == 0:
        m = int(round(np.sqrt(want / 12.0)))
        bv, bn, bt = displaced_cube_sphere(m, seed)
        ior = float(rng.uniform(1.3, 1.7))
        obj = MeshData(bv, bn, bt, matrix=translation(0.0, 0.0, 1.3), material=0)
        floor = plane(10.0, 10.0, 0.0)
        look, cam_pos = (0.0, 0.0, 0.8), (0.0, -6.0, 4.0)
    else:
        n = int(round(np.sqrt(want / 2.0))) + 1
        hv, hn, ht = heightfield(n, 32, seed)
        ior = 1.33
        obj = MeshData(hv, hn, ht, material=0)
        floor = plane(4.0, 4.0, -0.5)
        look, cam_pos = (0.0, 0.0, 0.0), (0.0, -1.6, 1.6)
    fv, fn, ft = floor
    meshes = [obj, MeshData(fv, fn, ft, material=1)]
    mats = [MaterialData("G", ior=ior, roughness=rough, color=(1.0, 1.0, 1.0)),
            MaterialData("P", roughness=1.0, specular=0.0, color=(0.8, 0.8, 0.8))]
    z = 5.0 if seed % 2 == 0 else 2.0
    kinds = ["point", "quad", "sphere", ["point", "quad", "sphere"][int(rng.integers(0, 3))]]
    lights = []
    for kind in kinds:
        pos = (float(rng.uniform(-1.0, 1.0)), float(rng.uniform(-1.0, 1.0)), z + float(rng.uniform(-0.3, 0.3)))
        power = _loguniform(rng, 0.25, 4.0)
        if kind == "point":
            lights.append(LightData("P", position=pos, power=power))
        elif kind == "sphere":
            lights.append(LightData("P", position=pos, power=power, radius=float(rng.uniform(0.1, 0.4))))
        else:
            s = float(rng.uniform(0.5, 2.0))
            tilt = (float(rng.uniform(-15.0, 15.0)), float(rng.uniform(-15.0, 15.0)), 0.0)
            lights.append(LightData("A", position=pos, width=s, height=s, power=power, rotation=tilt))
    cam = look_at_camera(res, res, cam_pos, look, 40.0)
    return SceneArgs(meshes, lights, mats, cam, samples=samples, base_radius=0.01, radius_reduction=0.75), photons


def c4_batch(n: int = 16, res: int = 2048, photons: int = 64 << 20, samples: int = 1):
    """The at-scale batch: c4_scene(seed) for seeds 0 .. n-1."""
    return [c4_scene(s, res, photons, samples) for s in range(n)]
