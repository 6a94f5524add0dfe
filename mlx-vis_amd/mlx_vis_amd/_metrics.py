"""Embedding-quality metrics on the GPU: trustworthiness and k-NN preservation.

The reference defines no metric; these are the textbook statistics
(trustworthiness: Venna & Kaski 2001, the sum sklearn.manifold.trustworthiness
computes; preservation: the mean overlap |kNN_X(i) & kNN_Y(i)| / k). Both use
exact neighbour lists from the brute kNN kernel; trustworthiness counts, for
every embedding neighbour, the points closer to the query in the input space
(libmivis ``mv_trust_excess``), without materialising the distance matrix.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi, _device
from ._knn import knn_brute_device
from ._parametric import knn_preservation_device


def _rows_tensor(rows, n: int, dev) -> torch.Tensor | None:
    if rows is None:
        return None
    r = torch.as_tensor(np.asarray(rows.cpu() if isinstance(rows, torch.Tensor) else rows), dtype=torch.int64)
    if r.dim() != 1:
        raise ValueError("rows must be a 1-D index array")
    if r.numel() and (int(r.min()) < 0 or int(r.max()) >= n):
        raise ValueError(f"rows must lie in [0, {n})")
    return r.to(device=dev, dtype=torch.int32).contiguous()


def _embedding_neighbours(Y: torch.Tensor, k: int, rows: torch.Tensor | None) -> torch.Tensor:
    if rows is None:
        idx, _ = knn_brute_device(Y, k)
    else:
        idx, _ = knn_brute_device(Y, k, Y.index_select(0, rows.long()).contiguous(), q_self=rows)
    return idx


def trust_excess_device(X: torch.Tensor, lo: torch.Tensor, rows: torch.Tensor | None = None) -> torch.Tensor:
    """Per-query sum of max(r - k, 0) over the neighbour list ``lo`` (m, k),
    r the rank of the neighbour among the input-space neighbours of the query
    (rows[q], or q when rows is None) in (distance, index) order. int64 (m,)."""
    _device.require_device("trust_excess_device", X, lo, rows)
    if X.dim() != 2 or lo.dim() != 2:
        raise ValueError("X must be (n, d) and lo (m, k)")
    n, d = X.shape
    m, k = lo.shape
    if rows is not None and rows.numel() != m:
        raise ValueError("rows must have one entry per row of lo")
    if rows is None and m > n:
        raise ValueError("lo has more rows than X")
    if not 1 <= k <= 64:
        raise ValueError("trustworthiness supports 1 <= k <= 64")
    X = _device.f32_rows(X)
    lo = _device.i32_vec(lo)
    rows = _device.i32_vec(rows)
    out = torch.empty(m, dtype=torch.int64, device=X.device)
    need = ctypes.c_size_t()
    _capi.call("mv_trust_ws_bytes", m, k, ctypes.byref(need))
    ws = _device.workspace(need.value)
    _capi.call("mv_trust_excess", X.data_ptr(), n, d, _capi.ptr(rows), m, lo.data_ptr(), k, k, out.data_ptr(),
               ws.data_ptr(), ws.numel(), _device.stream())
    return out


def trustworthiness(X, Y, k: int = 15, rows=None) -> float:
    """Trustworthiness of the embedding Y of X, over all points or a sample.

    T = 1 - 2 / (m k (2n - 3k - 1)) * sum_i sum_{j in U(i)} (r(i, j) - k),
    U(i): the k nearest neighbours of i in Y that are not among its k nearest
    in X, r(i, j): rank of j among i's neighbours in X (1 = nearest), m the
    number of rows averaged (n for the full statistic). Requires k < n / 2
    (the normalisation's range, as in sklearn)."""
    Xd = _device.as_device_f32(X)
    Yd = _device.as_device_f32(Y)
    n = Xd.shape[0]
    if Yd.shape[0] != n:
        raise ValueError("X and Y must have the same number of rows")
    if not 1 <= k < n / 2:
        raise ValueError(f"k must satisfy 1 <= k < n / 2, got k={k}, n={n}")
    r = _rows_tensor(rows, n, Xd.device)
    lo = _embedding_neighbours(Yd, k, r)
    m = lo.shape[0]
    if m == 0:
        raise ValueError("no rows to average over")
    total = int(trust_excess_device(Xd, lo, r).sum().item())
    return float(1.0 - 2.0 / (m * k * (2.0 * n - 3.0 * k - 1.0)) * total)


def knn_preservation(X, Y, k: int = 15, rows=None) -> float:
    """Mean |kNN_X(i) & kNN_Y(i)| / k over all points or the given rows, from
    exact (FP32 brute) neighbour lists in both spaces."""
    Xd = _device.as_device_f32(X)
    Yd = _device.as_device_f32(Y)
    n = Xd.shape[0]
    if Yd.shape[0] != n:
        raise ValueError("X and Y must have the same number of rows")
    if not 1 <= k < n:
        raise ValueError(f"k must satisfy 1 <= k < n, got k={k}, n={n}")
    r = _rows_tensor(rows, n, Xd.device)
    hi = _embedding_neighbours(Xd, k, r)
    lo = _embedding_neighbours(Yd, k, r)
    if hi.shape[0] == 0:
        raise ValueError("no rows to average over")
    counts, _ = knn_preservation_device(hi, lo)
    return float(counts.sum().item()) / (hi.shape[0] * k)
