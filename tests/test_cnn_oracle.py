"""Pins of the CNN oracle (oracle/cnn_oracle.py) on the CPU.

Each check reduces the oracle to something independent of it: torch's CPU
conv2d and Adam (library routines), the adjoint identities that define the
input and weight gradients, central finite differences of the whole loss, and
torch's own float -> bfloat16 conversion.
"""
import numpy as np
import pytest

from oracle import cnn_oracle as K

torch = pytest.importorskip("torch")
F = torch.nn.functional


def rand(rng, *shape):
    return rng.standard_normal(shape)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3.0e38, -1e-40], np.float32)])
    ours = K.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("ci,co,h,w", [(1, 32, 7, 9), (32, 32, 6, 5), (32, 1, 5, 8), (3, 4, 1, 1)])
def test_conv3x3_is_torch_conv2d(ci, co, h, w):
    rng = np.random.default_rng(ci * 100 + co)
    a = rand(rng, 2, h, w, ci)
    wt = rand(rng, co, ci, 9)
    ours = K.conv3x3(a, wt)
    ref = F.conv2d(torch.from_numpy(a).permute(0, 3, 1, 2), torch.from_numpy(wt).reshape(co, ci, 3, 3), padding=1)
    assert np.allclose(ours, ref.permute(0, 2, 3, 1).numpy(), rtol=1e-12, atol=1e-12)


def test_transpose_and_wgrad_are_adjoints():
    rng = np.random.default_rng(3)
    a, d, w = rand(rng, 2, 6, 7, 5), rand(rng, 2, 6, 7, 4), rand(rng, 4, 5, 9)
    lhs = np.sum(K.conv3x3(a, w) * d)
    assert np.isclose(lhs, np.sum(a * K.conv3x3_transpose(d, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * K.conv3x3_wgrad(a, d)), rtol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = {k: v.astype(np.float64) for k, v in K.init_params(4).items()}
    p["W5"] = p["W5"] * 10  # a residual that matters
    for k in p:
        if k.startswith("b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    x, t = rand(rng, 2, 6, 5), rand(rng, 2, 6, 5)
    cache = K.forward(p, x, 1.5, emulate=False)
    _, g = K.loss_and_backward(p, cache, t, 1.5, emulate=False)

    def loss(q):
        return K.loss_and_backward(q, K.forward(q, x, 1.5, emulate=False), t, 1.5, emulate=False)[0]

    h = 1e-6
    for k in p:  # a few coordinates of every tensor
        flat = p[k].reshape(-1)
        for i in rng.choice(flat.size, min(4, flat.size), replace=False):
            q = {kk: vv.copy() for kk, vv in p.items()}
            q[k].reshape(-1)[i] += h
            lp = loss(q)
            q[k].reshape(-1)[i] -= 2 * h
            lm = loss(q)
            fd = (lp - lm) / (2 * h)
            an = g[k].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-8, (k, i, fd, an)


def test_forward_backward_match_torch_autograd():
    """The whole network (no rounding) against torch autograd in float64."""
    rng = np.random.default_rng(5)
    p = {k: v.astype(np.float64) for k, v in K.init_params(5).items()}
    x, t = rand(rng, 3, 8, 7), rand(rng, 3, 8, 7)
    cache = K.forward(p, x, 2.0, emulate=False)
    loss, g = K.loss_and_backward(p, cache, t, 2.0, emulate=False)
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    xt = torch.from_numpy(x)[:, None] / 2.0
    a = xt
    for l in range(1, 5):
        w = tp[f"W{l}"].reshape(32, 1 if l == 1 else 32, 3, 3)
        a = torch.relu(F.conv2d(a, w, tp[f"b{l}"], padding=1))
    r = F.conv2d(a, tp["W5"].reshape(1, 32, 3, 3), tp["b5"], padding=1)
    y = xt + r
    lt = torch.mean((y - torch.from_numpy(t)[:, None] / 2.0) ** 2)
    lt.backward()
    assert np.isclose(loss, lt.item(), rtol=1e-12)
    for k in p:
        assert np.allclose(g[k], tp[k].grad.numpy(), rtol=1e-9, atol=1e-12), k


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    p = {"a": rand(rng, 5, 3), "b": rand(rng, 7)}
    tp = [torch.tensor(p["a"], requires_grad=True), torch.tensor(p["b"], requires_grad=True)]
    opt = torch.optim.Adam(tp, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = {k: np.zeros_like(v) for k, v in p.items()}
    v = {k: np.zeros_like(x) for k, x in p.items()}
    for t in range(1, 6):
        g = {"a": rand(rng, 5, 3), "b": rand(rng, 7)}
        p, m, v = K.adam(p, g, m, v, t)
        tp[0].grad, tp[1].grad = torch.from_numpy(g["a"]), torch.from_numpy(g["b"])
        opt.step()
    assert np.allclose(p["a"], tp[0].detach().numpy(), rtol=1e-12, atol=1e-14)
    assert np.allclose(p["b"], tp[1].detach().numpy(), rtol=1e-12, atol=1e-14)


def test_emulated_rounding_points():
    """With rounding on, activations are BF16 values and y - xn is the FP32 residual."""
    rng = np.random.default_rng(7)
    p = K.init_params(7)
    x = rand(rng, 1, 9, 9).astype(np.float32)
    c = K.forward(p, x, 3.0)
    for l in range(1, 5):
        a = c[f"a{l}"].astype(np.float32)
        assert np.array_equal(K.bf16(a), a) and (a >= 0).all()
    assert c["y"].dtype == np.float32
