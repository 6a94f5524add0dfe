"""CPU oracle of the learned form (parametric MLP embedding). Test
infrastructure only: the product never imports it.

The reference has no parametric UMAP; its MLP precedent is the MMAE
autoencoder (mlx_vis/_mmae/mmae.py:22-60, AdamW loop at 258-281). The loss
head follows the layout's gradient law (mlx_vis/_umap/umap.py:367-380)
without the +-4 clip.

Network: H1 = gelu(X W1^T + b1), H2 = gelu(H1 W2^T + b2), Z = H2 W3^T + b3,
GELU with the exact erf. Everything here is FP64. With ``bf16=True`` the
values the device rounds to BF16 are rounded the same way (round to nearest
even) at the same points: X, W1, W2, both pre-activations (after the bias),
both activations, and the two backpropagated pre-activation gradients. The
remaining difference to the device is FP32 versus FP64 accumulation.

Pins (tests/test_oracle_mlp.py):
  - backward() against central finite differences of a linear functional
    sum(G * Z) in FP64 (every parameter block), and against torch-CPU
    autograd in float64;
  - loss_grad() with rep_eps = 0 and no floor against central finite
    differences of edge_loss() (the cross-entropy it is the gradient of);
  - adam_step() against torch.optim.Adam on the same gradients.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

SQRT1_2 = 1.0 / math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def bf16_round(x) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = np.where(finite, r, u & 0xFFFF0000).astype(np.uint32)
    return out.view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x * SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x * SQRT1_2)) + x * INV_SQRT_2PI * np.exp(-0.5 * x * x)


PARAM_NAMES = ("W1", "b1", "W2", "b2", "W3", "b3")


def init_params(din, h1, h2, dout, seed=0):
    """Kaiming-uniform style init (the device has no initialiser; weights are
    passed in)."""
    rng = np.random.default_rng(seed)

    def lin(fan_out, fan_in):
        bound = 1.0 / math.sqrt(fan_in)
        return (rng.uniform(-bound, bound, (fan_out, fan_in)).astype(np.float32),
                rng.uniform(-bound, bound, fan_out).astype(np.float32))

    W1, b1 = lin(h1, din)
    W2, b2 = lin(h2, h1)
    W3, b3 = lin(dout, h2)
    return {"W1": W1, "b1": b1, "W2": W2, "b2": b2, "W3": W3, "b3": b3}


def flatten(params, dp=None) -> np.ndarray:
    """Parameters in the device blob order (W1 rows padded to dp columns)."""
    W1 = np.asarray(params["W1"], dtype=np.float64)
    if dp is not None and dp > W1.shape[1]:
        W1 = np.pad(W1, ((0, 0), (0, dp - W1.shape[1])))
    return np.concatenate([W1.ravel()] + [np.asarray(params[k], dtype=np.float64).ravel()
                                          for k in PARAM_NAMES[1:]])


def unflatten(flat, din, h1, h2, dout, dp=None):
    dp = dp or din
    o = 0
    out = {}
    for name, shape in (("W1", (h1, dp)), ("b1", (h1,)), ("W2", (h2, h1)), ("b2", (h2,)),
                        ("W3", (dout, h2)), ("b3", (dout,))):
        cnt = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + cnt], dtype=np.float64).reshape(shape)
        o += cnt
    out["W1"] = out["W1"][:, :din]
    return out


def forward(params, X, bf16=False):
    """Returns (Z, cache)."""
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    X = r(X)
    W1, W2 = r(params["W1"]), r(params["W2"])
    b1 = np.asarray(params["b1"], dtype=np.float64)
    b2 = np.asarray(params["b2"], dtype=np.float64)
    W3 = np.asarray(params["W3"], dtype=np.float64)
    b3 = np.asarray(params["b3"], dtype=np.float64)
    A1 = r(X @ W1.T + b1)
    H1 = r(gelu(A1))
    A2 = r(H1 @ W2.T + b2)
    H2 = r(gelu(A2))
    Z = H2 @ W3.T + b3
    return Z, {"X": X, "W1": W1, "W2": W2, "W3": W3, "A1": A1, "H1": H1, "A2": A2, "H2": H2}


def backward(cache, dZ, bf16=False):
    """Gradients of sum(dZ * Z) with respect to every parameter."""
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    dZ = np.asarray(dZ, dtype=np.float64)
    g = {"W3": dZ.T @ cache["H2"], "b3": dZ.sum(axis=0)}
    dA2 = r((dZ @ cache["W3"]) * gelu_grad(cache["A2"]))
    g["W2"] = dA2.T @ cache["H1"]
    g["b2"] = dA2.sum(axis=0)
    dA1 = r((dA2 @ cache["W2"]) * gelu_grad(cache["A1"]))
    g["W1"] = dA1.T @ cache["X"]
    g["b1"] = dA1.sum(axis=0)
    return g


def negatives(B: int, neg: int):
    """Negative tail rows of edge e: tails of edges (e + 1 + 97 s) mod B, skipping e."""
    out = []
    for s in range(neg):
        te = (np.arange(B) + 1 + 97 * s) % B
        out.append(te)
    return np.stack(out, axis=1) if neg else np.zeros((B, 0), dtype=np.int64)


def loss_grad(Z, Yb, B, neg, a, b, lam, rep_eps=0.001, floor=1e-6):
    """z-gradient of the edge loss plus regression, and the two loss terms.

    Rows [0, B) of Z are heads, [B, 2B) tails; Yb the regression targets of
    the same rows. Attraction coefficient -2ab d^(2(b-1)) / (1 + a d^(2b)),
    repulsion 2b / ((rep_eps + d^2)(1 + a d^(2b))) on the head only, both
    divided by B (umap.py:367-380 without the clip)."""
    Z = np.asarray(Z, dtype=np.float64)
    Yb = np.asarray(Yb, dtype=np.float64)
    zi, zj = Z[:B], Z[B:]
    dZ = np.zeros_like(Z)
    diff = zi - zj
    d2 = np.maximum((diff ** 2).sum(axis=1), floor)
    pw = d2 ** b
    coef = -2.0 * a * b * d2 ** (b - 1.0) / (1.0 + a * pw)
    ce = np.log1p(a * pw).sum()
    gq = coef[:, None] * diff / B
    dZ[:B] -= gq
    dZ[B:] += gq
    for s, te in enumerate(negatives(B, neg).T):
        keep = te != np.arange(B)
        dn = zi - zj[te]
        n2 = np.maximum((dn ** 2).sum(axis=1), floor)
        pwn = n2 ** b
        c = 2.0 * b / ((rep_eps + n2) * (1.0 + a * pwn))
        q = 1.0 / (1.0 + a * pwn)
        ce -= np.where(keep, np.log(np.maximum(1.0 - q, 1e-30)), 0.0).sum()
        dZ[:B] -= np.where(keep[:, None], c[:, None] * dn / B, 0.0)
    res = Z - Yb
    reg = (res ** 2).sum()
    dZ += 2.0 * lam * res / B
    return dZ, (ce / B, lam * reg / B)


def edge_loss(Z, Yb, B, neg, a, b, lam):
    """The loss whose gradient loss_grad(rep_eps=0, floor=0) is (cross-entropy
    of the attractive edges and the in-batch negatives, plus regression)."""
    Z = np.asarray(Z, dtype=np.float64)
    zi, zj = Z[:B], Z[B:]
    d2 = ((zi - zj) ** 2).sum(axis=1)
    L = np.log1p(a * d2 ** b).sum()
    for te in negatives(B, neg).T:
        keep = te != np.arange(B)
        n2 = ((zi - zj[te]) ** 2).sum(axis=1)
        L -= np.where(keep, np.log(1.0 - 1.0 / (1.0 + a * n2 ** b)), 0.0).sum()
    return L / B + lam * ((Z - Yb) ** 2).sum() / B


def adam_step(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015), as torch.optim.Adam."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    c1 = 1.0 - b1 ** step
    c2 = 1.0 - b2 ** step
    p = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return p, m, v


def batch_grad(params, X, Y, idx, B, neg, a, b, lam, bf16=False, rep_eps=0.001, floor=1e-6):
    """Full gradient of one batch (heads idx[:B], tails idx[B:2B])."""
    Xb = np.asarray(X)[idx]
    Yb = np.asarray(Y)[idx]
    Z, cache = forward(params, Xb, bf16=bf16)
    dZ, losses = loss_grad(Z, Yb, B, neg, a, b, lam, rep_eps=rep_eps, floor=floor)
    return backward(cache, dZ, bf16=bf16), losses, Z, dZ
