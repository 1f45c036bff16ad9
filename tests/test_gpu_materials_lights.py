"""GPU parity with the CPU oracle for every BSDF class and every light kind.

Scenes (``scenes.materials_scene``): one sphere per BSDF class of the
reference -- Fresnel_Specular glass, Specular_Reflection mirror,
Microfacet_Reflection_GGX rough metal, Fresnel_Microfacet_GGX with
specular > 0, Fresnel_Specular_Microfacet_GGX and a specular-0 plastic --
over a glossy plastic floor (specular > 0, so the merge takes the general
BSDF evaluation, ``src/Path_Vertices.h:46-68``), lit by a rotated area light,
or by a rotated area light, a sun, a point light and a sphere light
(``scenes.mixed_lights``) under compat_point_light / compat_light_pdf 0 and 1
(``src/Lights.h:78-81``, ``src/Light_list.cpp:10-30``).

Bars, as in test_gpu_parity.py: photon lists bit-exact per path (>= 99.9 % of
paths; the rest are closest-hit ties); camera pass on a given photon list and
two-iteration renders within 2e-4 relative + 1e-7 absolute per pixel (the
merge sums in another order)."""
import numpy as np
import pytest

from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine, photons_from_oracle

pytestmark = pytest.mark.gpu

CENTRES = np.asarray([(1.8 * (k - 2.5), 0.3 * (k % 2), 0.75) for k in range(6)])


def _object_of(pos):
    """0-5: within 0.72 of sphere k's centre (radius 0.7); 6: floor."""
    lab = np.full(len(pos), 6)
    for k in range(6):
        lab[np.linalg.norm(pos - CENTRES[k], axis=1) < 0.72] = k
    return lab


def _compare_photons(ph_g, uid_g, ph_o):
    uid_o = ph_o[:, 10].astype(np.uint32)
    common, ig, io = np.intersect1d(uid_g, uid_o, return_indices=True)
    assert len(common) >= 0.999 * max(len(uid_g), len(uid_o)), (len(common), len(uid_g), len(uid_o))
    exact = np.all(ph_g[ig].view(np.uint32) == photons_from_oracle(ph_o[io]).view(np.uint32), axis=1)
    assert exact.mean() >= 0.999, exact.mean()
    return common


def _close(img_g, img_o, max_bad=0.0):
    tol = 2e-4 * np.abs(img_o) + 1e-7
    bad = np.abs(img_g - img_o) > tol
    assert bad.mean() <= max_bad, (bad.mean(), np.abs(img_g - img_o).max())


def test_materials_first_hit(oracle):
    args = scenes.materials_scene()
    eng = Engine(args)
    orc = oracle.OracleScene(args)
    rng = np.random.default_rng(5)
    o = rng.uniform((-6, -3, 0.2), (6, 3, 3.0), (20000, 3))
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    tri_g, t_g = eng.first_hit(rays)
    tri_o, t_o = orc.first_hit(rays, brute=True)
    assert np.array_equal(tri_g >= 0, tri_o >= 0)
    same = (tri_g == tri_o) & (tri_o >= 0)
    assert np.array_equal(t_g[same], t_o[same].astype(np.float32))
    assert (~same & (tri_o >= 0)).sum() <= 0.001 * (tri_o >= 0).sum()
    eng.release()


def test_every_bsdf_class_photon_list_bit_exact(oracle):
    args = scenes.materials_scene()
    eng = Engine(args, photons=131072, seed=7)
    orc = oracle.OracleScene(args, photons=131072, seed=7)
    ph_g, uid_g = eng.trace_photons(0)
    ph_o = orc.trace_light(0)
    _compare_photons(ph_g, uid_g, ph_o)
    # photons sit on every non-delta sphere (rough metal, both plastics with
    # specular > 0, the diffuse one) and on the floor; paths of length > 1 carry
    # a nonzero dVM, so the recurrence of Sample_Scattering_light is exercised
    counts = np.bincount(_object_of(ph_o[:, :3]), minlength=7)
    assert np.all(counts[[2, 3, 4, 5, 6]] > 10), counts
    assert np.any(ph_o[:, 9] != 0.0)
    assert len(np.unique(ph_o[:, 10].astype(np.int64) & 7)) >= 3
    eng.release()


@pytest.mark.parametrize("exact,group", [(1, 8), (0, 8), (1, 1), (0, 1)])
def test_every_bsdf_class_camera_pass(oracle, exact, group):
    # group 8: the grouped merge (the default), 1: one query per lane -- both through the general BSDF evaluation
    args = scenes.materials_scene()
    eng = Engine(args, photons=131072, seed=3, exact_merge=exact, group_queries=group)
    orc = oracle.OracleScene(args, photons=131072, seed=3)
    ph = orc.trace_light(0)
    img_o = orc.camera_pass(0, ph)
    img_g = eng.camera_pass(0, photons_from_oracle(ph))
    assert (img_o > 0).mean() > 0.05
    _close(img_g, img_o)
    eng.release()


def test_every_bsdf_class_render(oracle):
    args = scenes.materials_scene()
    eng = Engine(args, photons=65536, seed=11)
    orc = oracle.OracleScene(args, photons=65536, seed=11)
    eng.render_iteration()
    eng.render_iteration()
    img_g = eng.image(scaled=True)
    img_o = orc.render(2)
    assert img_o.sum() > 0
    _close(img_g, img_o, max_bad=0.002)
    eng.release()


@pytest.mark.parametrize("compat_point,compat_pdf", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_mixed_lights_photon_list_bit_exact(oracle, compat_point, compat_pdf):
    args = scenes.materials_scene(lights=scenes.mixed_lights())
    opts = dict(photons=131072, seed=5, compat_point_light=compat_point, compat_light_pdf=compat_pdf)
    eng = Engine(args, **opts)
    orc = oracle.OracleScene(args, **opts)
    ph_g, uid_g = eng.trace_photons(0)
    ph_o = orc.trace_light(0)
    assert len(ph_o) > 300
    _compare_photons(ph_g, uid_g, ph_o)
    eng.release()


@pytest.mark.parametrize("compat", [0, 1])
def test_mixed_lights_camera_pass(oracle, compat):
    args = scenes.materials_scene(lights=scenes.mixed_lights())
    opts = dict(photons=131072, seed=9, compat_point_light=compat, compat_light_pdf=compat)
    eng = Engine(args, **opts)
    orc = oracle.OracleScene(args, **opts)
    ph = orc.trace_light(0)
    img_o = orc.camera_pass(0, ph)
    img_g = eng.camera_pass(0, photons_from_oracle(ph))
    assert img_o.sum() > 0
    _close(img_g, img_o)
    eng.release()
