"""CPU oracle of the edge-gradient family beyond UMAP (test infrastructure
only): CNE's contrastive gradients (mlx_vis/_cne/cne.py:192-313), its Adam
loop (cne.py:315-372), PaCMAP/LocalMAP pair laws (mlx_vis/_pacmap/pacmap.py:
275-354) and t-SNE's attractive term (mlx_vis/_tsne/tsne.py:511-517), in FP64
with np.add.at scatters.

Pins (tests/test_oracle_cne.py): every CNE gradient is the exact gradient of
its loss (InfoNCE: sum_e -log(s_p / (s_p + sum_k s_k)); NCE with q = 1/m:
sum_e -log(s_p / (s_p + q)) - sum_k log(1 - s_k / (s_k + q)); NEG:
sum_e -log s_p - sum_k log(1 - s_k)) and is checked against central finite
differences; the PaCMAP laws are the gradients of w d/(10 + d),
w d/(10^4 + d) and w/(1 + d) (d = |y_s - y_t|^2 + 1), also checked by finite
differences; build_edges against a brute-force set construction.
"""

from __future__ import annotations

import numpy as np


def build_edges(knn_indices, n):
    """CNE._build_edges — cne.py:192-217: both directions, unique, no self-loops."""
    k = knn_indices.shape[1]
    src = np.repeat(np.arange(n, dtype=np.int32), k)
    dst = knn_indices.ravel().astype(np.int32)
    all_src = np.concatenate([src, dst])
    all_dst = np.concatenate([dst, src])
    keys = all_src.astype(np.int64) * n + all_dst.astype(np.int64)
    _, unique_idx = np.unique(keys, return_index=True)
    edges = np.stack([all_src[unique_idx], all_dst[unique_idx]], axis=1)
    return edges[edges[:, 0] != edges[:, 1]]


def _sims(Y, ei, ej, neg):
    dp = Y[ei] - Y[ej]
    sp = 1.0 / (1.0 + (dp * dp).sum(-1))
    dn = Y[ei][:, None, :] - Y[neg]
    sn = 1.0 / (1.0 + (dn * dn).sum(-1))
    return dp, sp, dn, sn


def cne_grad(Y, ei, ej, neg, loss):
    """Gradient of the summed CNE loss (cne.py:219-313), FP64."""
    Y = np.asarray(Y, dtype=np.float64)
    m = neg.shape[1]
    dp, sp, dn, sn = _sims(Y, ei, ej, neg)
    if loss == "infonce":
        Z = sp + sn.sum(1)
        wa = (1.0 - sp / Z) * 2.0 * sp
        wr = (sn / Z[:, None]) * 2.0 * sn
    elif loss == "nce":
        q = 1.0 / m
        wa = 2.0 * q * sp / (sp + q)
        wr = 2.0 * sn * sn / (sn + q)
    else:
        wa = 2.0 * sp
        wr = 2.0 * sn * sn / np.maximum(1.0 - sn, 1e-8)
    ga = wa[:, None] * dp
    grad = np.zeros_like(Y)
    np.add.at(grad, ei, ga - (wr[:, :, None] * dn).sum(1))
    np.add.at(grad, ej, -ga)
    np.add.at(grad, neg.reshape(-1), (wr[:, :, None] * dn).reshape(-1, Y.shape[1]))
    return grad


def cne_loss(Y, ei, ej, neg, loss):
    """The loss whose gradient cne_grad is."""
    Y = np.asarray(Y, dtype=np.float64)
    m = neg.shape[1]
    _, sp, _, sn = _sims(Y, ei, ej, neg)
    if loss == "infonce":
        return float((-np.log(sp) + np.log(sp + sn.sum(1))).sum())
    if loss == "nce":
        q = 1.0 / m
        return float((-np.log(sp / (sp + q)) - np.log(1.0 - sn / (sn + q)).sum(1)).sum())
    return float((-np.log(sp) - np.log(1.0 - sn).sum(1)).sum())


def adam_cne(Y, g, m, v, itr, lr, b1=0.9, b2=0.999, eps=1e-8):
    """The Adam step of CNE._optimize (cne.py:355-360)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    Y = Y - lr * (m / (1.0 - b1 ** itr)) / (np.sqrt(v / (1.0 - b2 ** itr)) + eps)
    return Y, m, v


def pair_grad(Y, src, dst, law, w=1.0, w_pair=None, scale=1.0):
    """Per-pair force laws of PaCMAP/LocalMAP (pacmap.py:275-354) and t-SNE's
    attractive term (tsne.py:511-517), FP64."""
    Y = np.asarray(Y, dtype=np.float64)
    df = Y[src] - Y[dst]
    d2 = (df * df).sum(1)
    d = d2 + 1.0
    ww = np.full(src.shape[0], w, dtype=np.float64) if w_pair is None else np.asarray(w_pair, dtype=np.float64)
    if law == "pacmap_nn":
        c = ww * 20.0 / ((10.0 + d) ** 2)
    elif law == "pacmap_mn":
        c = ww * 20000.0 / ((10000.0 + d) ** 2)
    elif law == "pacmap_fp":
        c = -(ww * 2.0 / ((1.0 + d) ** 2))
    elif law == "localmap_nn":
        c = ww * 20.0 * scale / (((10.0 + d) ** 2) * np.sqrt(d))
    elif law == "tsne_attr":
        c = 4.0 * ww / (1.0 + d2)
    else:
        raise ValueError(law)
    g = c[:, None] * df
    grad = np.zeros_like(Y)
    np.add.at(grad, src, g)
    if law != "tsne_attr":
        np.add.at(grad, dst, -g)
    return grad


def pair_loss(Y, src, dst, law, w=1.0):
    """Losses whose gradients the PaCMAP laws are."""
    df = np.asarray(Y, dtype=np.float64)[src] - np.asarray(Y, dtype=np.float64)[dst]
    d = (df * df).sum(1) + 1.0
    if law == "pacmap_nn":
        return float((w * d / (10.0 + d)).sum())
    if law == "pacmap_mn":
        return float((w * d / (10000.0 + d)).sum())
    if law == "pacmap_fp":
        return float((w / (1.0 + d)).sum())
    raise ValueError(law)
