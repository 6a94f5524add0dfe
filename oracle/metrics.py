"""Embedding-quality and recall metrics on the CPU (test infrastructure only).

The reference defines none (it has no metric code); these are the textbook
definitions, pinned against sklearn where sklearn has one.
"""

from __future__ import annotations

import numpy as np


def knn_sets_overlap(a_idx: np.ndarray, b_idx: np.ndarray) -> np.ndarray:
    """Per-row |A_i intersect B_i| / k for two (n, k) neighbour arrays."""
    a = np.sort(np.asarray(a_idx), axis=1)
    b = np.sort(np.asarray(b_idx), axis=1)
    k = a.shape[1]
    out = np.empty(a.shape[0], dtype=np.float64)
    for i in range(a.shape[0]):
        out[i] = np.intersect1d(a[i], b[i], assume_unique=False).size / k
    return out


def recall(approx_idx, exact_idx) -> float:
    """Mean |approx_i intersect exact_i| / k."""
    return float(knn_sets_overlap(approx_idx, exact_idx).mean())


def knn_preservation(X_high, Y_low, k: int = 15, rows=None) -> float:
    """Mean overlap of the k-NN sets in the input and the embedding."""
    from .knn import exact_knn

    X_high = np.asarray(X_high, dtype=np.float32)
    Y_low = np.asarray(Y_low, dtype=np.float32)
    if rows is None:
        hi, _ = exact_knn(X_high, X_high, k, self_base=0)
        lo, _ = exact_knn(Y_low, Y_low, k, self_base=0)
    else:
        rows = np.asarray(rows)
        hi, _ = exact_knn(X_high[rows], X_high, k, q_self=rows)
        lo, _ = exact_knn(Y_low[rows], Y_low, k, q_self=rows)
    return recall(lo, hi)


def trust_excess(X_high, lo_idx, rows=None) -> np.ndarray:
    """Per-row sum of max(r(i, j) - k, 0) over the given low-dimensional
    neighbour lists lo_idx (m, k): r(i, j) is the rank of j among all other
    points ordered by (input-space squared distance, index), 1 = nearest; the
    query point itself sorts last (rank n). FP64 expansion distances (exact
    for small-integer or dyadic inputs)."""
    X = np.asarray(X_high, dtype=np.float64)
    lo = np.asarray(lo_idx, dtype=np.int64)
    n = X.shape[0]
    rows = np.arange(lo.shape[0]) if rows is None else np.asarray(rows)
    k = lo.shape[1]
    sqx = (X * X).sum(1)
    out = np.zeros(rows.size, dtype=np.int64)
    for s in range(0, rows.size, 256):
        r = rows[s:s + 256]
        Dh = sqx[r, None] + sqx[None, :] - 2.0 * X[r] @ X.T
        Dh[np.arange(r.size), r] = np.inf
        order_h = np.argsort(Dh, axis=1, kind="stable")
        ranks = np.empty_like(order_h)
        ranks[np.arange(r.size)[:, None], order_h] = np.arange(1, n + 1)[None, :]
        rk = np.take_along_axis(ranks, lo[s:s + 256], axis=1)
        out[s:s + r.size] = np.maximum(rk - k, 0).sum(axis=1)
    return out


def trustworthiness(X_high, Y_low, k: int = 15, rows=None) -> float:
    """Trustworthiness (Venna & Kaski 2001), optionally over a row sample.

    T = 1 - 2 / (m k (2n - 3k - 1)) * sum_i sum_{j in U_k(i)} (r(i, j) - k),
    U_k(i): the embedding's k neighbours of i that are not among its k input
    neighbours; r(i, j): rank of j among i's input neighbours (1 = nearest).
    m = number of rows averaged (n for the full statistic).
    """
    Y = np.asarray(Y_low, dtype=np.float64)
    n = Y.shape[0]
    rows = np.arange(n) if rows is None else np.asarray(rows)
    sqy = (Y * Y).sum(1)
    nn_l = np.empty((rows.size, k), dtype=np.int64)
    for s in range(0, rows.size, 256):
        r = rows[s:s + 256]
        Dl = sqy[r, None] + sqy[None, :] - 2.0 * Y[r] @ Y.T
        Dl[np.arange(r.size), r] = np.inf
        nn_l[s:s + r.size] = np.argsort(Dl, axis=1, kind="stable")[:, :k]
    total = int(trust_excess(X_high, nn_l, rows).sum())
    m = rows.size
    return float(1.0 - 2.0 / (m * k * (2.0 * n - 3.0 * k - 1.0)) * total)
