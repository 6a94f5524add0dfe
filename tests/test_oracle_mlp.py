"""Pins the MLP oracle (oracle/mlp.py): finite differences, torch autograd,
torch.optim.Adam, and the BF16 rounding against torch's bfloat16 cast."""

import numpy as np
import pytest
import torch

from oracle import mlp as om

DIMS = (13, 24, 16, 2)   # din, h1, h2, dout


def _params(seed=0):
    p = om.init_params(*DIMS, seed=seed)
    rng = np.random.default_rng(seed + 100)
    for k in ("b1", "b2", "b3"):   # non-zero biases so that every term matters
        p[k] = (p[k] + rng.normal(0, 0.3, p[k].shape)).astype(np.float32)
    return {k: np.asarray(v, np.float64) for k, v in p.items()}


def test_bf16_round_matches_torch_cast():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 5000), rng.normal(0, 1e-3, 5000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3.0e38, np.inf, -np.inf]])
    x = x.astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(om.bf16_round(x), want)


def test_backward_matches_finite_differences():
    p = _params()
    rng = np.random.default_rng(1)
    X = rng.normal(size=(7, DIMS[0]))
    G = rng.normal(size=(7, DIMS[3]))
    _, cache = om.forward(p, X)
    grads = om.backward(cache, G)
    h = 1e-6
    for name in om.PARAM_NAMES:
        flat = p[name].reshape(-1)
        picks = rng.choice(flat.size, size=min(12, flat.size), replace=False)
        for t in picks:
            old = flat[t]
            flat[t] = old + h
            zp = (om.forward(p, X)[0] * G).sum()
            flat[t] = old - h
            zm = (om.forward(p, X)[0] * G).sum()
            flat[t] = old
            fd = (zp - zm) / (2 * h)
            assert abs(grads[name].reshape(-1)[t] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, t)


def test_backward_matches_torch_autograd():
    p = _params(3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(9, DIMS[0]))
    G = rng.normal(size=(9, DIMS[3]))
    _, cache = om.forward(p, X)
    grads = om.backward(cache, G)
    tp = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    x = torch.tensor(X, dtype=torch.float64)
    gelu = torch.nn.functional.gelu
    h1 = gelu(x @ tp["W1"].T + tp["b1"])
    h2 = gelu(h1 @ tp["W2"].T + tp["b2"])
    z = h2 @ tp["W3"].T + tp["b3"]
    (z * torch.tensor(G)).sum().backward()
    for k in om.PARAM_NAMES:
        np.testing.assert_allclose(grads[k], tp[k].grad.numpy(), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("neg", [0, 3])
def test_loss_grad_matches_finite_differences(neg):
    rng = np.random.default_rng(5)
    B, a, b, lam = 11, 1.577, 0.895, 0.3
    Z = rng.normal(0, 2.0, size=(2 * B, 2))
    Y = rng.normal(0, 2.0, size=(2 * B, 2))
    dZ, (ce, reg) = om.loss_grad(Z, Y, B, neg, a, b, lam, rep_eps=0.0, floor=0.0)
    assert abs(ce + reg - om.edge_loss(Z, Y, B, neg, a, b, lam)) < 1e-12
    h = 1e-6
    fd = np.zeros_like(Z)
    for i in range(Z.shape[0]):
        for q in range(2):
            Zp, Zm = Z.copy(), Z.copy()
            Zp[i, q] += h
            Zm[i, q] -= h
            fd[i, q] = (om.edge_loss(Zp, Y, B, neg, a, b, lam) - om.edge_loss(Zm, Y, B, neg, a, b, lam)) / (2 * h)
    # Negatives move the head only (umap.py:384-386), so the head rows are the
    # full gradient and the tail rows the gradient of the loss without them.
    np.testing.assert_allclose(dZ[:B], fd[:B], rtol=1e-5, atol=1e-7)
    fd0 = np.zeros_like(Z)
    for i in range(B, 2 * B):
        for q in range(2):
            Zp, Zm = Z.copy(), Z.copy()
            Zp[i, q] += h
            Zm[i, q] -= h
            fd0[i, q] = (om.edge_loss(Zp, Y, B, 0, a, b, lam) - om.edge_loss(Zm, Y, B, 0, a, b, lam)) / (2 * h)
    np.testing.assert_allclose(dZ[B:], fd0[B:], rtol=1e-5, atol=1e-7)


def test_loss_grad_attraction_is_the_layout_law():
    # With one edge, no negatives and lam = 0, the head's step -B * dZ equals the
    # layout's unclipped attractive update per unit step (umap.py:367-372).
    a, b = 1.577, 0.895
    Z = np.array([[0.3, -0.2], [1.1, 0.5]])
    dZ, _ = om.loss_grad(Z, Z, 1, 0, a, b, 0.0)
    d2 = ((Z[0] - Z[1]) ** 2).sum()
    coef = -2 * a * b * d2 ** (b - 1) / (1 + a * d2 ** b)
    np.testing.assert_allclose(-dZ[0], coef * (Z[0] - Z[1]), rtol=1e-12)
    np.testing.assert_allclose(dZ[1], -dZ[0], rtol=1e-12)


def test_adam_step_matches_torch():
    rng = np.random.default_rng(6)
    p0 = rng.normal(size=50)
    tp = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.normal(size=50)
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = om.adam_step(p, g, m, v, step)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=0, atol=1e-12)


def test_flatten_roundtrip_with_padding():
    p = _params()
    flat = om.flatten(p, dp=64)
    assert flat.size == DIMS[1] * 64 + DIMS[1] + DIMS[2] * DIMS[1] + DIMS[2] + DIMS[3] * DIMS[2] + DIMS[3]
    q = om.unflatten(flat, *DIMS, dp=64)
    for k in om.PARAM_NAMES:
        np.testing.assert_array_equal(q[k], p[k])
