"""Brute-force kNN oracle (test infrastructure only; see oracle/__init__.py).

Canonical order (what the product must reproduce bit for bit): ascending
exact squared distance, ties broken by the lower index. The exact distance of
a pair is the IEEE double sum, in order t = 0..d-1, of
((double)q_t - (double)x_t)^2 without fused multiply-add. This is the order
the reference's argsort (mlx_vis/_knn.py:71-72) produces on exact distances;
the reference's own FP32 norm-expansion values (_knn.py:60-63) are only an
approximation of them and are restated separately in ``reference_knn``.
"""

from __future__ import annotations

import numpy as np


def dist64_seq(q: np.ndarray, xs: np.ndarray) -> np.ndarray:
    """Canonical squared distances between rows q (m, d) and xs (m, c, d).

    Sequential over t, unfused: acc += (q_t - x_t)^2 in float64.
    """
    q = np.asarray(q, dtype=np.float32)
    xs = np.asarray(xs, dtype=np.float32)
    m, c, d = xs.shape
    acc = np.zeros((m, c), dtype=np.float64)
    qd = q.astype(np.float64)
    for t in range(d):
        diff = qd[:, t][:, None] - xs[:, :, t].astype(np.float64)
        acc += diff * diff
    return acc


def _self_index(nq: int, q_self, self_base) -> np.ndarray:
    if q_self is not None:
        return np.asarray(q_self, dtype=np.int64).reshape(nq)
    if self_base is not None and self_base >= 0:
        return np.arange(nq, dtype=np.int64) + int(self_base)
    return np.full(nq, -1, dtype=np.int64)


def exact_knn(Q, X, k: int, q_self=None, self_base=None, row_chunk: int = 512,
              return_squared: bool = True):
    """Canonical kNN of each row of Q among the rows of X.

    A float64 BLAS screen nominates k + m candidates per row; their canonical
    distances are computed by ``dist64_seq``; a row is accepted when the k-th
    canonical distance lies below the smallest screened value outside the
    candidates minus the float64 screen error bound, else m grows (up to all
    rows, where the result is exhaustive).

    Returns (idx int64 (nq, k), d64 float64 (nq, k)) with squared distances
    (or Euclidean if return_squared=False).
    """
    Q = np.ascontiguousarray(np.asarray(Q, dtype=np.float32))
    X = np.ascontiguousarray(np.asarray(X, dtype=np.float32))
    nq, d = Q.shape
    n = X.shape[0]
    selfs = _self_index(nq, q_self, self_base)
    Xd = X.astype(np.float64)
    sqx = np.einsum("ij,ij->i", Xd, Xd)
    out_i = np.empty((nq, k), dtype=np.int64)
    out_d = np.empty((nq, k), dtype=np.float64)
    eps64 = 2.0 ** -53
    for s in range(0, nq, row_chunk):
        e = min(nq, s + row_chunk)
        Qd = Q[s:e].astype(np.float64)
        sqq = np.einsum("ij,ij->i", Qd, Qd)
        D = sqq[:, None] + sqx[None, :] - 2.0 * (Qd @ Xd.T)
        np.maximum(D, 0.0, out=D)
        rows = np.arange(e - s)
        sel = selfs[s:e]
        has_self = sel >= 0
        D[rows[has_self], sel[has_self]] = np.inf
        err = (2.0 * d + 8.0) * eps64 * 4.0 * (sqq[:, None] + sqx.max())
        pending = np.arange(e - s)
        m = max(2 * k, k + 16)
        while pending.size:
            width = min(n, m)
            Dp = D[pending]
            if width < n:
                part = np.argpartition(Dp, width - 1, axis=1)[:, :width]
                outside = np.partition(Dp, width, axis=1)[:, width] if width < n else None
            else:
                part = np.tile(np.arange(n), (pending.size, 1))
                outside = None
            cand_x = X[part]  # (p, width, d)
            d64 = dist64_seq(Q[s + pending], cand_x)
            valid = np.isfinite(Dp[np.arange(pending.size)[:, None], part])
            d64 = np.where(valid, d64, np.inf)
            order = np.lexsort((part, d64), axis=1)
            cs = np.take_along_axis(part, order, 1)[:, :k]
            ds = np.take_along_axis(d64, order, 1)[:, :k]
            if outside is None:
                ok = np.ones(pending.size, dtype=bool)
            else:
                ok = ds[:, k - 1] * (1 + 1e-12) < outside - err[pending, 0]
            out_i[s + pending[ok]] = cs[ok]
            out_d[s + pending[ok]] = ds[ok]
            pending = pending[~ok]
            m *= 4
    if not return_squared:
        out_d = np.sqrt(out_d)
    return out_i, out_d


def reference_knn(X, k: int, return_euclidean: bool = True):
    """Op-for-op FP32 restatement of the reference brute path, _knn.py:47-88.

    Norm expansion in FP32 (_knn.py:49, 60-63), +1e30 on the diagonal
    (_knn.py:64-69), a full-row stable argsort keeping k (_knn.py:71-72) per
    row chunk of min(n, max(1000, 5e8 // (4 n))) rows (_knn.py:52), sqrt of the
    clamped values (_knn.py:86-87). Used as the CPU baseline; its near-tie
    order follows FP32 rounding and is not the canonical order.
    """
    X = np.asarray(X, dtype=np.float32)
    n = X.shape[0]
    sq = np.einsum("ij,ij->i", X, X, dtype=np.float32)
    chunk = min(n, max(1000, 500_000_000 // (n * 4)))
    idx_out = np.zeros((n, k), dtype=np.int32)
    d_out = np.zeros((n, k), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(s + chunk, n)
        D = np.maximum(sq[s:e, None] + sq[None, :] - np.float32(2.0) * (X[s:e] @ X.T), np.float32(0.0))
        D[np.arange(e - s), np.arange(s, e)] += np.float32(1e30)
        idx = np.argsort(D, axis=1, kind="stable")[:, :k]
        idx_out[s:e] = idx
        d_out[s:e] = np.take_along_axis(D, idx, axis=1)
    if return_euclidean:
        return idx_out, np.sqrt(np.maximum(d_out, 0)).astype(np.float32)
    return idx_out, d_out


def knn_loops(Q, X, k: int, q_self=None):
    """Pure-Python canonical kNN for tiny inputs (pins exact_knn)."""
    Q = np.asarray(Q, dtype=np.float32)
    X = np.asarray(X, dtype=np.float32)
    res_i, res_d = [], []
    for r in range(Q.shape[0]):
        pairs = []
        for j in range(X.shape[0]):
            if q_self is not None and q_self[r] == j:
                continue
            acc = 0.0
            for t in range(Q.shape[1]):
                diff = float(Q[r, t]) - float(X[j, t])
                acc = acc + diff * diff
            pairs.append((acc, j))
        pairs.sort()
        res_i.append([j for _, j in pairs[:k]])
        res_d.append([v for v, _ in pairs[:k]])
    return np.array(res_i, dtype=np.int64), np.array(res_d, dtype=np.float64)


def reference_knn_rows(X, rows, k: int, return_euclidean: bool = True):
    """``reference_knn`` for a subset of query rows (the same per-row work:
    FP32 norm expansion against all of X, +1e30 self mask, full stable argsort)."""
    X = np.asarray(X, dtype=np.float32)
    rows = np.asarray(rows)
    sq = np.einsum("ij,ij->i", X, X, dtype=np.float32)
    D = np.maximum(sq[rows, None] + sq[None, :] - np.float32(2.0) * (X[rows] @ X.T), np.float32(0.0))
    D[np.arange(rows.size), rows] += np.float32(1e30)
    idx = np.argsort(D, axis=1, kind="stable")[:, :k]
    d = np.take_along_axis(D, idx, axis=1)
    return idx.astype(np.int32), (np.sqrt(np.maximum(d, 0)) if return_euclidean else d).astype(np.float32)
