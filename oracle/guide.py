"""CPU oracle of the learned emission proposal (test infrastructure only).

Restates, in numpy, the network of realcausticsengine_amd/csrc/rc_guide.h:
Fourier encoding (6 inputs x 16 frequencies pi*2^(k/2) x {sin, cos} = 192
features), three ReLU layers of 64, 1024 logits; the weighted softmax
cross-entropy  L = sum_i w_i (logsumexp(z_i) - z_i[b_i]) / sum_i w_i,  its
gradient by back-propagation, and Adam (Kingma & Ba 2015, Algorithm 1).
The reference has no such component (its emission is plain
Area_Light::sample_light, src/Lights.cpp:41-55), so nothing here follows a
reference line; the functions are pinned instead by

* finite differences of the exact (float64) loss       -> backward
* torch.autograd on the same network in float64         -> forward + backward
* torch.optim.Adam                                       -> adam
* the float64 sin/cos, and own_sincos's own tests        -> encode

Two modes: ``exact`` (float64 throughout) and ``bf16`` (the rounding points of
the GPU kernels: BF16 features, weights, hidden activations, dZ and the
back-propagated activations; FP32 logits), which the GPU is compared against
with a stated tolerance.  Only tests/, ``__graft_entry__.smoke`` and bench.py's
cpu_baseline leg may import this module.
"""
from __future__ import annotations

import numpy as np

N_IN, N_FREQ, N_FEAT, N_HID, N_BINS = 6, 16, 192, 64, 1024
FREQS = np.array([np.float32(np.pi * 2.0 ** (k / 2.0)) for k in range(N_FREQ)], dtype=np.float32)

OFF = {}
_o = 0
for _name, _size in (("W1", N_FEAT * N_HID), ("b1", N_HID), ("W2", N_HID * N_HID), ("b2", N_HID),
                     ("W3", N_HID * N_HID), ("b3", N_HID), ("W4", N_HID * N_BINS), ("b4", N_BINS)):
    OFF[_name] = (_o, _size)
    _o += _size
N_PARAMS = _o  # 87,232
SHAPES = {"W1": (N_FEAT, N_HID), "W2": (N_HID, N_HID), "W3": (N_HID, N_HID), "W4": (N_HID, N_BINS),
          "b1": (N_HID,), "b2": (N_HID,), "b3": (N_HID,), "b4": (N_BINS,)}


def unpack(flat: np.ndarray) -> dict:
    return {k: flat[o:o + n].reshape(SHAPES[k]) for k, (o, n) in OFF.items()}


def pack(d: dict) -> np.ndarray:
    out = np.zeros(N_PARAMS, dtype=np.float64)
    for k, (o, n) in OFF.items():
        out[o:o + n] = np.asarray(d[k], dtype=np.float64).reshape(-1)
    return out


def bf16(a) -> np.ndarray:
    """Round float32 values to BF16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def own_sincos_f32(x: np.ndarray):
    """float32 restatement of own_sincos2 (realcausticsengine_amd/csrc/rc_vec.h)."""
    f = np.float32
    x = np.asarray(x, dtype=np.float32)
    fk = x * f(0.6366197466850281)
    k = np.where(fk >= 0, fk + f(0.5), fk - f(0.5)).astype(np.int32)  # truncation toward zero
    kf = k.astype(np.float32)
    r = (x - kf * f(1.5703125)) - kf * f(0.0004838267923332751)
    r2 = r * r
    ps = f(-0.00019841270113829523) + r2 * f(2.7557318844628753e-06)
    ps = f(0.008333333767950535) + r2 * ps
    ps = f(-0.1666666716337204) + r2 * ps
    s = r + (r * r2) * ps
    pc = f(2.4801587642286904e-05) + r2 * f(-2.755731998149713e-07)
    pc = f(-0.0013888889225199819) + r2 * pc
    pc = f(0.0416666679084301) + r2 * pc
    pc = f(-0.5) + r2 * pc
    c = f(1.0) + r2 * pc

    def quad(q):
        return np.where(q == 0, s, np.where(q == 1, c, np.where(q == 2, -s, -c)))

    return quad(k & 3), quad((k + 1) & 3)


def encode(x: np.ndarray, mode: str = "bf16") -> np.ndarray:
    """Features [n, 192]: column i*32 + 2k (+1) = sin (cos) of FREQS[k] * x_i."""
    x = np.asarray(x)[:, :N_IN]
    n = x.shape[0]
    out = np.zeros((n, N_IN, N_FREQ, 2), dtype=np.float64)
    if mode == "exact":
        arg = x.astype(np.float64)[:, :, None] * FREQS.astype(np.float64)[None, None, :]
        out[..., 0], out[..., 1] = np.sin(arg), np.cos(arg)
        return out.reshape(n, N_FEAT)
    arg = x.astype(np.float32)[:, :, None] * FREQS[None, None, :]
    s, c = own_sincos_f32(arg)
    out[..., 0], out[..., 1] = bf16(s), bf16(c)
    return out.reshape(n, N_FEAT)


def _r(a, mode):
    return bf16(a).astype(np.float64) if mode == "bf16" else np.asarray(a, dtype=np.float64)


def forward(params: np.ndarray, x: np.ndarray, mode: str = "bf16"):
    """Returns (logits [n, 1024], cache for backward)."""
    P = unpack(np.asarray(params, dtype=np.float64))
    W = {k: _r(P[k], mode) for k in ("W1", "W2", "W3", "W4")}
    X = encode(x, mode)
    acts = [X]
    h = X
    for l in (1, 2, 3):
        pre = h @ W[f"W{l}"]
        if mode == "bf16":  # FP32 accumulator + FP32 bias, ReLU, BF16
            h = bf16(np.maximum(pre.astype(np.float32) + P[f"b{l}"].astype(np.float32), np.float32(0)))
            h = h.astype(np.float64)
        else:
            h = np.maximum(pre + P[f"b{l}"], 0.0)
        acts.append(h)
    z = h @ W["W4"]
    if mode == "bf16":
        z = (z.astype(np.float32) + P["b4"].astype(np.float32)).astype(np.float64)
    else:
        z = z + P["b4"]
    return z, {"acts": acts, "W": W, "mode": mode}


def loss_and_grad(params: np.ndarray, x: np.ndarray, bins: np.ndarray, w: np.ndarray, mode: str = "bf16"):
    """Weighted cross-entropy and its gradient (flat, float64)."""
    z, cache = forward(params, x, mode)
    acts, W = cache["acts"], cache["W"]
    bins = np.asarray(bins)
    w = np.asarray(w, dtype=np.float64)
    valid = bins >= 0
    wsum = w.sum()
    wr = np.where(valid, w / wsum, 0.0) if wsum > 0 else np.zeros_like(w)
    m = z.max(axis=1, keepdims=True)
    e = np.exp(z - m)
    s = e.sum(axis=1, keepdims=True)
    p = e / s
    rows = np.arange(len(bins))
    bsafe = np.where(valid, bins, 0)
    loss = float(np.sum(wr * (np.log(s[:, 0]) + m[:, 0] - z[rows, bsafe])))
    onehot = np.zeros_like(p)
    onehot[rows[valid], bins[valid]] = 1.0
    dz = _r(wr[:, None] * (p - onehot), mode)
    g = {"W4": acts[3].T @ dz, "b4": dz.sum(axis=0)}
    d = dz
    for l in (3, 2, 1):
        nxt = "W4" if l == 3 else f"W{l + 1}"
        d = _r((d @ W[nxt].T) * (acts[l] > 0), mode)
        g[f"W{l}"] = acts[l - 1].T @ d
        g[f"b{l}"] = d.sum(axis=0)
    return loss, pack(g)


class Adam:
    """Adam, Kingma & Ba (2015) Algorithm 1, bias-corrected, on a flat vector."""

    def __init__(self, params: np.ndarray, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.p = np.array(params, dtype=np.float64)
        self.m = np.zeros_like(self.p)
        self.v = np.zeros_like(self.p)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps

    def step(self, g: np.ndarray) -> np.ndarray:
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.p = self.p - self.lr * mh / (np.sqrt(vh) + self.eps)
        return self.p


def table_cdf(logits: np.ndarray) -> np.ndarray:
    """Per-row CDF of softmax(logits): [rows, 1025], cdf[0] = 0, cdf[1024] = 1."""
    z = np.asarray(logits, dtype=np.float64)
    e = np.exp(z - z.max(axis=1, keepdims=True))
    c = np.cumsum(e, axis=1)
    out = np.zeros((z.shape[0], N_BINS + 1))
    out[:, 1:] = c / c[:, -1:]
    return out


def init_params(seed: int = 1) -> np.ndarray:
    """A He-uniform initialisation from numpy (any parameters serve the checks)."""
    rng = np.random.default_rng(seed)
    d = {}
    for k in ("W1", "W2", "W3", "W4"):
        fan_in = SHAPES[k][0]
        a = np.sqrt(6.0 / fan_in) * (0.1 if k == "W4" else 1.0)
        d[k] = rng.uniform(-a, a, SHAPES[k])
    for k in ("b1", "b2", "b3", "b4"):
        d[k] = rng.normal(0.0, 0.05, SHAPES[k])
    return pack(d)
