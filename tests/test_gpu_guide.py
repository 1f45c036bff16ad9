"""The learned emission proposal on the GPU against its CPU oracle.

Bars (BF16 products accumulate in FP32 in a different order than numpy, and
a value near a BF16 rounding boundary may round the other way):
  logits           relative L2 error < 1e-4, max error < 1e-3 of max |logit|
  loss             relative error < 1e-4
  gradients        relative L2 error < 3e-3 per tensor (vs the BF16-mode oracle;
                   the weight gradients are sums of FP32 atomics, so their bits vary)
  Adam step        |dp| < 2e-6 (FP32 arithmetic of the same formula)
  table CDF        max |dCDF| < 1e-5
  guided photons   bit-exact per path (>= 99.9% of paths), as the plain photon list
"""
import numpy as np
import pytest

from oracle import guide
from realcausticsengine_amd import scenes
from realcausticsengine_amd.engine import Engine, photons_from_oracle

pytestmark = pytest.mark.gpu


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 6)).astype(np.float32)
    bins = rng.integers(0, 1024, n).astype(np.int32)
    bins[::9] = -1
    w = rng.exponential(1.0, n).astype(np.float32)
    w[::4] = 0.0
    return x, bins, w


@pytest.fixture(scope="module")
def eng():
    e = Engine(scenes.water_scene(32, 16, 16, samples=1), photons=16384, seed=0, write_exr=0)
    e.prepare()
    yield e
    e.release()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle(eng):
    p = guide.init_params(3).astype(np.float32)
    eng.guide_set_params(p)
    x, _, _ = _rows(1000, 1)
    z = eng.guide_forward(x)
    zo, _ = guide.forward(p.astype(np.float64), x, "bf16")
    print("forward rel", _rel(z, zo), "max", np.abs(z - zo).max(), "scale", np.abs(zo).max(), flush=True)
    assert _rel(z, zo) < 1e-4
    assert np.abs(z - zo).max() < 1e-3 * np.abs(zo).max()


def test_gradients_match_oracle(eng):
    p = guide.init_params(5).astype(np.float32)
    eng.guide_set_params(p)
    x, bins, w = _rows(4096, 2)
    loss, g = eng.guide_train_step(x, bins, w, apply=False)
    lo, go = guide.loss_and_grad(p.astype(np.float64), x, bins, w, "bf16")
    _, ge = guide.loss_and_grad(p.astype(np.float64), x, bins, w, "exact")
    print("loss", loss, lo, flush=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    for k, (o, n) in guide.OFF.items():
        r, re = _rel(g[o:o + n], go[o:o + n]), _rel(g[o:o + n], ge[o:o + n])
        print(f"grad {k}: rel vs bf16 oracle {r:.2e}, vs exact {re:.2e}", flush=True)
        assert r < 3e-3, k
        assert re < 0.1, k


def test_adam_step_matches_oracle(eng):
    p = guide.init_params(7).astype(np.float32)
    eng.guide_set_params(p)
    x, bins, w = _rows(2048, 3)
    _, g = eng.guide_train_step(x, bins, w, apply=True)
    p1 = eng.guide_params()
    ref = guide.Adam(p.astype(np.float64), lr=1e-3).step(g.astype(np.float64))
    assert np.abs(p1 - ref).max() < 2e-6
    assert np.abs(p1 - p).max() > 5e-4  # the step moved the weights


def test_table_matches_oracle(eng):
    p = guide.init_params(11).astype(np.float32)
    p[guide.OFF["W4"][0]:guide.OFF["W4"][0] + guide.OFF["W4"][1]] *= 20.0  # peaked distributions
    eng.guide_set_params(p)
    cdf, inp = eng.guide_table()
    assert len(cdf) == eng.get("guide_slots") * 16 * 16 == 256
    assert np.all(cdf[:, 0] == 0.0) and np.all(cdf[:, -1] == 1.0)
    assert np.all(np.diff(cdf, axis=1) >= 0.0)
    zo, _ = guide.forward(p.astype(np.float64), inp, "bf16")
    co = guide.table_cdf(zo)
    print("table max |dcdf|", np.abs(cdf - co).max(), flush=True)
    assert np.abs(cdf - co).max() < 1e-5
    # inputs: cell centres of the 2 x 2 light (z = 2) with its axis -z
    np.testing.assert_allclose(inp[:, 3:], np.tile([0.0, 0.0, -1.0], (len(inp), 1)), atol=1e-6)


def test_guided_photons_match_oracle(oracle):
    args = scenes.water_scene(32, 16, 16, samples=1)
    n = 32768
    eng = Engine(args, photons=n, seed=0, write_exr=0, guide=1, guide_warmup=0, guide_train=0, guide_alpha=0.2)
    eng.prepare()
    p = guide.init_params(13).astype(np.float32)
    p[guide.OFF["W4"][0]:guide.OFF["W4"][0] + guide.OFF["W4"][1]] *= 20.0
    eng.guide_set_params(p)
    ph_g, uid_g = eng.trace_photons(0)
    cdf, _ = eng.guide_table()
    orc = oracle.OracleScene(args, photons=n, seed=0)
    plain = orc.trace_light(0)
    orc.set_guide(cdf, slots=1, cells=16, slot_of_light=[0], alpha=0.2)
    ph_o = orc.trace_light(0)
    uid_o = ph_o[:, 10].astype(np.uint32)
    assert len(ph_o) > 1000
    common, ig, io = np.intersect1d(uid_g, uid_o, return_indices=True)
    assert len(common) >= 0.999 * max(len(uid_g), len(uid_o))
    exact = np.all(ph_g[ig].view(np.uint32) == photons_from_oracle(ph_o[io]).view(np.uint32), axis=1)
    print("guided photons", len(ph_g), "exact", exact.mean(), "plain photons", len(plain), flush=True)
    assert exact.mean() >= 0.999, exact.mean()
    # guiding changed the photon list
    cp, ip, _ = np.intersect1d(plain[:, 10].astype(np.uint32), uid_o, return_indices=True)
    assert len(cp) < 0.95 * len(uid_o) or not np.allclose(plain[ip, :3], ph_o[np.searchsorted(uid_o, cp), :3])
    eng.release()


def test_training_records_are_plain_bins(oracle):
    args = scenes.water_scene(32, 16, 16, samples=1)
    eng = Engine(args, photons=16384, seed=5, write_exr=0, guide=1, guide_warmup=1000000, guide_batch=4096)
    eng.prepare()
    eng.render_iteration()
    x, bins = eng.guide_records(4096)
    assert len(bins) == 4096
    want = np.zeros(4096, dtype=np.int64)
    u1 = np.zeros(4096)
    for i in range(4096):
        w0 = oracle.philox([0, i, 0, 0], [5, 0])
        w1 = oracle.philox([1, i, 0, 0], [5, 0])
        want[i] = (int(w0[3]) >> 27) | ((int(w1[0]) >> 27) << 5)
        u1[i] = (int(w0[1]) >> 8) * 2.0 ** -24
    np.testing.assert_array_equal(bins, want)
    np.testing.assert_allclose(x[:, 3:], np.tile([0.0, 0.0, -1.0], (4096, 1)), atol=1e-6)
    # the recorded position is an affine function of the position sample
    A = np.stack([u1, np.ones_like(u1)], axis=1)
    coef, res, _, _ = np.linalg.lstsq(A, x[:, 0].astype(np.float64), rcond=None)
    assert np.sqrt(res[0] / 4096) < 1e-5 and abs(coef[0]) > 0.1
    assert eng.stats()["guide_hit_fraction"] > 0.0
    eng.release()


def _run(guided: bool, iters: int, photons: int):
    args = scenes.water_scene(128, 64, 64, alpha=1.0, samples=iters)
    eng = Engine(args, photons=photons, seed=0, write_exr=0, guide=1, guide_batch=photons,
                 guide_warmup=3 if guided else 10 ** 9, compat_rows=0)
    eng.prepare()
    imgs, hit = [], []
    prev = np.zeros((64, 64, 3))
    for i in range(iters):
        eng.render_iteration()
        acc = eng.image(scaled=False).astype(np.float64)
        imgs.append(acc - prev)
        prev = acc
        hit.append(eng.stats()["guide_hit_fraction"])
    eng.release()
    return np.array(imgs), np.array(hit)


def test_trained_proposal_gains_over_plain():
    """Same scene, seed and photon count: the guided run sends more light paths
    into camera-visible merge cells, keeps the image's expectation (the mixture
    weight 1/q), and lowers the per-iteration pixel variance."""
    iters, photons = 40, 65536
    g_img, g_hit = _run(True, iters, photons)
    p_img, p_hit = _run(False, iters, photons)
    skip = 10  # training iterations before comparing
    gh, ph = g_hit[skip:].mean(), p_hit[skip:].mean()
    gm, pm = g_img[skip:].mean(axis=0), p_img[skip:].mean(axis=0)
    gv, pv = g_img[skip:].var(axis=0).mean(), p_img[skip:].var(axis=0).mean()
    print(f"hit fraction guided {gh:.4f} plain {ph:.4f}; image sum guided {gm.sum():.4f} plain {pm.sum():.4f}; "
          f"pixel variance guided {gv:.4e} plain {pv:.4e} (ratio {gv / pv:.3f})", flush=True)
    assert gh > 1.2 * ph
    assert abs(gm.sum() / pm.sum() - 1.0) < 0.03
    assert gv < pv
