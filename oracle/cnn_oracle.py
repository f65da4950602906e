"""CPU oracle of the sparse-angle residual CNN -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` import this module, as the checker of the GPU kernels in
``projects_amd/csrc/k_cnn.hip``.  The reference has no learned component; the
network is the one the benchmark names: five 3x3 convolutions (1 -> 32 -> 32
-> 32 -> 32 -> 1 channels), ReLU after the first four, a residual connection
from the input, BF16 weights and activations with FP32 accumulation, mean
squared error against the target, Adam.

Rounding points (the GPU rounds at exactly these places, round-to-nearest-even):

* the normalised input x/scale is rounded to BF16 before the first convolution;
* every weight used by a convolution is the BF16 rounding of the FP32 master;
* a1..a4 = bf16(relu(conv + bias));
* the residual r = conv5(a4) + b5 and the output y = x/scale + r stay FP32;
* activation gradients dz4..dz1 are rounded to BF16 after the ReLU mask.

Accumulation here is float64, so GPU-vs-oracle differences are the GPU's FP32
accumulation order plus the BF16 roundings it then feeds (tolerances in the
tests say which).  ``emulate=False`` drops every rounding: the network in
float64, the form the finite-difference and library checks pin.

Layouts: activations [B, H, W, C] (channels last, as on the GPU); weights
W1 [32][9], W2..W4 [32 co][32 ci][9], W5 [32][9] (taps ky*3+kx, correlation:
z[y, x] = sum W[tap] * in[y+ky-1, x+kx-1], zero outside the image).
"""
from __future__ import annotations

import numpy as np

C = 32
PARAM_SHAPES = [("W1", (C, 9)), ("b1", (C,)), ("W2", (C, C, 9)), ("b2", (C,)), ("W3", (C, C, 9)), ("b3", (C,)),
                ("W4", (C, C, 9)), ("b4", (C,)), ("W5", (C, 9)), ("b5", (1,))]


def bf16(x) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float32 values."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(f), f, out)


def init_params(seed: int) -> dict:
    """He-normal weights (std sqrt(2 / fan_in)), zero biases; the last layer
    scaled by 0.1 so the untrained network starts close to the identity."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in PARAM_SHAPES:
        if name.startswith("b"):
            p[name] = np.zeros(shape, np.float32)
            continue
        fan_in = 9 * (1 if name == "W1" else C)
        w = rng.standard_normal(shape) * np.sqrt(2.0 / fan_in)
        if name == "W5":
            w *= 0.1
        p[name] = w.astype(np.float32)
    return p


def _shift(a: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """out[:, y, x] = a[:, y+dy, x+dx], zero outside."""
    B, H, W = a.shape[:3]
    out = np.zeros_like(a)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd] = a[:, ys, xs]
    return out


def conv3x3(a: np.ndarray, w: np.ndarray) -> np.ndarray:
    """a [B,H,W,Ci], w [Co][Ci][9] -> [B,H,W,Co], float64."""
    a = a.astype(np.float64)
    w = w.astype(np.float64)
    out = 0.0
    for t in range(9):
        ky, kx = divmod(t, 3)
        out = out + _shift(a, ky - 1, kx - 1) @ w[:, :, t].T
    return out


def conv3x3_transpose(d: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Gradient of conv3x3 with respect to its input: d [B,H,W,Co] -> [B,H,W,Ci]."""
    d = d.astype(np.float64)
    w = w.astype(np.float64)
    out = 0.0
    for t in range(9):
        ky, kx = divmod(t, 3)
        out = out + _shift(d @ w[:, :, t], 1 - ky, 1 - kx)
    return out


def conv3x3_wgrad(a: np.ndarray, d: np.ndarray) -> np.ndarray:
    """Gradient of conv3x3 with respect to its weights: -> [Co][Ci][9]."""
    a = a.astype(np.float64)
    d = d.astype(np.float64)
    Ci, Co = a.shape[-1], d.shape[-1]
    g = np.zeros((Co, Ci, 9))
    dm = d.reshape(-1, Co)
    for t in range(9):
        ky, kx = divmod(t, 3)
        g[:, :, t] = dm.T @ _shift(a, ky - 1, kx - 1).reshape(-1, Ci)
    return g


def forward(p: dict, x: np.ndarray, scale: float = 1.0, emulate: bool = True) -> dict:
    """x [B,H,W] -> cache with a0 (the network input), a1..a4, r, y (normalised units)."""
    rnd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    wq = {k: (rnd(v) if k.startswith("W") else np.asarray(v, np.float64)) for k, v in p.items()}
    xn = np.asarray(x, np.float32) / np.float32(scale) if emulate else np.asarray(x, np.float64) / scale
    a0 = rnd(xn)[..., None]
    cache = {"xn": xn, "a0": a0, "wq": wq}
    a = a0
    for l in range(1, 5):
        w = wq[f"W{l}"] if l > 1 else wq["W1"][:, None, :]
        z = conv3x3(a, w) + wq[f"b{l}"]
        a = rnd(np.maximum(z, 0.0))
        cache[f"a{l}"] = a
    r = conv3x3(a, wq["W5"][None, :, :])[..., 0] + wq["b5"][0]
    cache["r"] = r.astype(np.float32) if emulate else r
    cache["y"] = (xn + cache["r"]).astype(np.float32) if emulate else xn + r
    return cache


def loss_and_backward(p: dict, cache: dict, target: np.ndarray, scale: float = 1.0, emulate: bool = True):
    """Mean squared error of y against target/scale and the gradients of every parameter."""
    rnd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    wq = cache["wq"]
    tn = np.asarray(target, np.float64) / scale
    y = np.asarray(cache["y"], np.float64)
    n = y.size
    loss = float(np.mean((y - tn) ** 2))
    dy = 2.0 * (y - tn) / n
    if emulate:
        dy = dy.astype(np.float32).astype(np.float64)
    g = {}
    g["b5"] = np.array([dy.sum()])
    a4 = cache["a4"]
    g["W5"] = conv3x3_wgrad(a4, dy[..., None])[0]
    da = conv3x3_transpose(dy[..., None], wq["W5"][None, :, :])
    for l in range(4, 0, -1):
        dz = rnd(da * (cache[f"a{l}"] > 0))
        a_prev = cache[f"a{l - 1}"]
        gw = conv3x3_wgrad(a_prev, dz)
        g[f"W{l}"] = gw[:, 0, :] if l == 1 else gw
        g[f"b{l}"] = dz.reshape(-1, C).sum(axis=0)
        if l > 1:
            da = conv3x3_transpose(dz, wq[f"W{l}"])
    return loss, g


def adam(p: dict, g: dict, m: dict, v: dict, t: int, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999,
         eps: float = 1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected), in place on float64 copies."""
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return p, m, v
