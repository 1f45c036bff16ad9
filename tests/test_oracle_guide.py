"""Pins the CPU oracle of the learned emission proposal (oracle/guide.py).

Every function is checked against something independent of its own
formulas: central finite differences of the float64 loss, torch.autograd
and torch.optim.Adam on the same network in float64, torch's BF16 rounding,
the C oracle's own_sincos and numpy's sin/cos.
"""
import numpy as np
import pytest
import torch

from oracle import binding, guide


def _case(n=24, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 6))
    bins = rng.integers(0, guide.N_BINS, n)
    bins[::7] = -1  # rows without a target
    w = rng.exponential(1.0, n)
    w[::5] = 0.0
    return x, bins, w


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-3, 1000), [0.0, 1.0, -1.0, 1 + 2 ** -8,
                        1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, 65504.0]]).astype(np.float32)
    ours = guide.bf16(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(ours, ref)


def test_own_sincos_restatement_matches_c_oracle_and_sin():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-1, 1, 300) * f for f in guide.FREQS]).astype(np.float32)
    s, c = guide.own_sincos_f32(x)
    cs = np.array([binding.own_math(0, float(v)) for v in x[::7]], dtype=np.float32)
    cc = np.array([binding.own_math(1, float(v)) for v in x[::7]], dtype=np.float32)
    np.testing.assert_array_equal(s[::7], cs)
    np.testing.assert_array_equal(c[::7], cc)
    # |arg| <= 569: the float32 argument reduction keeps 1e-5 absolute accuracy
    assert np.max(np.abs(s - np.sin(x.astype(np.float64)))) < 1e-5
    assert np.max(np.abs(c - np.cos(x.astype(np.float64)))) < 1e-5


def test_encode_layout():
    x = np.array([[0.25, -0.5, 0.1, 0.0, 0.0, -1.0]])
    f = guide.encode(x, "exact")[0]
    for i in range(6):
        for k in range(16):
            arg = float(guide.FREQS[k]) * x[0, i]
            assert f[i * 32 + 2 * k] == pytest.approx(np.sin(arg), abs=1e-12)
            assert f[i * 32 + 2 * k + 1] == pytest.approx(np.cos(arg), abs=1e-12)
    fb = guide.encode(x, "bf16")[0]
    assert np.max(np.abs(fb - f)) < 2 ** -8  # BF16 keeps 8 significant bits of values <= 1


def test_backward_matches_finite_differences():
    x, bins, w = _case(n=12)
    p = guide.init_params(5)
    loss, g = guide.loss_and_grad(p, x, bins, w, "exact")
    rng = np.random.default_rng(7)
    # coordinates in every tensor, including ones that ReLU may switch off
    idx = np.concatenate([rng.integers(o, o + n, 6) for o, n in guide.OFF.values()])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (guide.loss_and_grad(pp, x, bins, w, "exact")[0] - guide.loss_and_grad(pm, x, bins, w, "exact")[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def _torch_loss(p, x, bins, w):
    P = {k: torch.from_numpy(v.copy()).requires_grad_(True) for k, v in guide.unpack(p).items()}
    X = torch.from_numpy(guide.encode(x, "exact"))
    h = X
    for l in (1, 2, 3):
        h = torch.relu(h @ P[f"W{l}"] + P[f"b{l}"])
    z = h @ P["W4"] + P["b4"]
    valid = torch.from_numpy(bins >= 0)
    wt = torch.from_numpy(w) * valid
    ce = torch.nn.functional.cross_entropy(z, torch.from_numpy(np.where(bins >= 0, bins, 0)), reduction="none")
    loss = (wt * ce).sum() / torch.from_numpy(w).sum()
    loss.backward()
    return float(loss.detach()), guide.pack({k: v.grad.numpy() for k, v in P.items()}), z.detach().numpy()


def test_forward_and_backward_match_torch_autograd():
    x, bins, w = _case(n=40, seed=11)
    p = guide.init_params(9)
    loss, g = guide.loss_and_grad(p, x, bins, w, "exact")
    z, _ = guide.forward(p, x, "exact")
    tl, tg, tz = _torch_loss(p, x, bins, w)
    np.testing.assert_allclose(z, tz, rtol=1e-12, atol=1e-12)
    assert loss == pytest.approx(tl, rel=1e-12)
    np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-14)


def test_bf16_mode_is_close_to_exact():
    x, bins, w = _case(n=64, seed=2)
    p = guide.init_params(4)
    le, ge = guide.loss_and_grad(p, x, bins, w, "exact")
    lb, gb = guide.loss_and_grad(p, x, bins, w, "bf16")
    assert lb == pytest.approx(le, rel=2e-2)
    for k, (o, n) in guide.OFF.items():
        rel = np.linalg.norm(gb[o:o + n] - ge[o:o + n]) / max(np.linalg.norm(ge[o:o + n]), 1e-30)
        assert rel < 0.1, k


def test_adam_matches_torch():
    rng = np.random.default_rng(2)
    p0 = rng.normal(0, 1, 1000)
    opt = guide.Adam(p0, lr=1e-3)
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    topt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(6):
        g = rng.normal(0, 1, 1000) * (10.0 ** -step)
        opt.step(g)
        topt.zero_grad()
        tp.grad = torch.from_numpy(g.copy())
        topt.step()
    np.testing.assert_allclose(opt.p, tp.detach().numpy(), rtol=0, atol=1e-12)


def test_table_cdf():
    z = np.random.default_rng(0).normal(0, 3, (5, guide.N_BINS))
    c = guide.table_cdf(z)
    assert np.all(c[:, 0] == 0) and np.allclose(c[:, -1], 1.0)
    assert np.all(np.diff(c, axis=1) >= 0)
    p = np.exp(z - z.max(1, keepdims=True))
    np.testing.assert_allclose(np.diff(c, axis=1), p / p.sum(1, keepdims=True), rtol=1e-12, atol=1e-15)
