"""Shared kNN operator: brute force (libmivis MFMA screen + exact re-rank) or
NNDescent.

Mirrors ``compute_knn`` of mlx_vis/_knn.py:7-88: same arguments, same
``auto`` rule (brute for n <= 20,000, else NNDescent, _knn.py:32-33), same
NNDescent seed default (42, _knn.py:40), same return types (int32 indices,
float32 distances, Euclidean unless ``return_euclidean=False``).

The brute path returns the canonical neighbours: ascending exact distance,
ties broken by the lower index (the reference's argsort order,
_knn.py:71-72); the screen ``dtype`` ("fp32", "bf16", "fp8", "bf16x3": split BF16) changes speed,
never the result.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi, _device


def knn_brute_device(
    X: torch.Tensor,
    k: int,
    Q: torch.Tensor | None = None,
    *,
    dtype: str = "fp32",
    rerank: int | None = None,
    q_self: torch.Tensor | None = None,
    exclude_self: bool | None = None,
    euclid: bool = True,
    stats: torch.Tensor | None = None,
):
    """Device-level brute kNN.

    X: (n, d) CUDA tensor. Q: (nq, d) queries (default X itself). Other
    dtypes and strided views are converted to contiguous float32 first.
    Self exclusion: with ``Q is None`` the query row i never returns i (the
    +1e30 diagonal of _knn.py:64-69); ``q_self`` (nq,) int32 names the row of X
    each query must not return (-1 for none).
    rerank: the screen nominates rerank * k candidates per row (at most 128
    with the symmetric self-kNN screen, 64 otherwise). Default: 2, and for
    the BF16 screen enough for 128, the candidate set its certificate needs on
    clustered data whose 15th and 30th neighbours are close.
    k: up to 64 for any query set; up to 1,024 for self kNN (Q is None), which
    then selects from the symmetric screen's n x n matrix (n up to ~113,000).
    Returns (idx int32 (nq, k), dist float32 (nq, k)) CUDA tensors.
    """
    if dtype not in _capi.DTYPES:
        raise ValueError(f"unknown kNN dtype {dtype!r}; expected one of {sorted(_capi.DTYPES)}")
    if rerank is None:
        rerank = max(2, -(-128 // int(k))) if dtype == "bf16" else 2
    _device.require_device("knn_brute_device", X, Q, stats)
    same = Q is None
    if X.dim() != 2 or (not same and (Q.dim() != 2 or X.shape[1] != Q.shape[1])):
        raise ValueError(f"bad shapes X {tuple(X.shape)} Q {None if same else tuple(Q.shape)}")
    X = _device.f32_rows(X)
    Q = X if same else _device.f32_rows(Q)
    n, d = X.shape
    nq = Q.shape[0]
    if exclude_self is None:
        exclude_self = same and q_self is None
    self_base = 0 if (exclude_self and q_self is None) else -1
    if q_self is not None:
        q_self = q_self.to(device=X.device, dtype=torch.int32).contiguous()
        if q_self.numel() != nq:
            raise ValueError("q_self must have one entry per query row")
    out_idx = torch.empty((nq, k), dtype=torch.int32, device=X.device)
    out_dist = torch.empty((nq, k), dtype=torch.float32, device=X.device)
    lib = _capi.lib()
    import ctypes

    def ws_bytes():
        nbytes = ctypes.c_size_t(0)
        _capi.check(
            lib.mv_knn_ws_bytes(nq, n, d, int(k), _capi.DTYPES[dtype], int(rerank), int(same), ctypes.byref(nbytes)),
            "mv_knn_ws_bytes",
        )
        return nbytes.value

    nbytes = ws_bytes()
    # The symmetric self-kNN screen keeps the n x n matrix in its workspace;
    # when that would take more than half of the free device memory, this call
    # uses the fused screen instead (same results, no matrix).
    sym_prev = None
    if same and n >= 4096 and k <= 64:   # k above 64 needs the matrix
        free, _ = torch.cuda.mem_get_info(X.device)
        if nbytes > free // 2:
            sym_prev = lib.mv_knn_set_symmetric(0)
            nbytes = ws_bytes()
    try:
        ws = _device.workspace(nbytes)
        _capi.check(
            lib.mv_knn_brute(
                Q.data_ptr(), nq, X.data_ptr(), n, d, int(k), _capi.DTYPES[dtype], int(rerank),
                _capi.ptr(q_self), self_base, int(bool(euclid)),
                out_idx.data_ptr(), out_dist.data_ptr(), _capi.ptr(stats),
                ws.data_ptr(), ws.numel(), _device.stream(),
            ),
            "mv_knn_brute",
        )
    finally:
        if sym_prev is not None:
            lib.mv_knn_set_symmetric(sym_prev)
    return out_idx, out_dist


def compute_knn(X, k, method="auto", verbose=False, random_state=None, return_euclidean=True,
                dtype="fp32"):
    """Compute k-nearest neighbours (drop-in for mlx_vis._knn.compute_knn).

    Args:
        X: (n, d) array (numpy, torch, or anything ``np.asarray`` accepts).
        k: number of neighbours.
        method: "auto" (brute for n <= 20000, else NNDescent; brute as well when
            k is too large for the device NNDescent's candidate slots),
            "brute", "nndescent".
        random_state: NNDescent seed (default 42).
        return_euclidean: Euclidean distances if True, squared if False.
        dtype: brute-force screen precision ("fp32", "bf16", "fp8", "bf16x3").

    Returns:
        (indices int32 (n, k), distances float32 (n, k)) numpy arrays.
    """
    if isinstance(X, torch.Tensor):
        n = X.shape[0]
    else:
        X = np.asarray(X)
        n = X.shape[0]
    if method == "auto":
        from ._nndescent import nndescent_fits

        method = "brute" if n <= 20000 or not nndescent_fits(min(k, n - 1)) else "nndescent"
    if method == "nndescent":
        from ._nndescent import NNDescent

        nn = NNDescent(k=k, verbose=verbose,
                       random_state=random_state if random_state is not None else 42)
        indices, dists = nn.build(X)
        if not return_euclidean:
            dists = dists ** 2
        return indices, dists
    if method != "brute":
        raise ValueError(f"Unknown knn method: {method}")
    Xd = _device.as_device_f32(X)
    idx, dist = knn_brute_device(Xd, k, dtype=dtype, euclid=bool(return_euclidean))
    return idx.cpu().numpy(), dist.cpu().numpy()
