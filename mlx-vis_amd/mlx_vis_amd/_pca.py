"""PCA preprocessing (mlx_vis/_umap/umap.py:92-113) with the array-sized work
on the device: the covariance (X - mean)^T (X - mean) / (n - 1) and the
projection (libmivis mv_pca_cov, mv_pca_project); the eigendecomposition of
the d x d covariance runs on the host in FP64, as the reference runs it on
MLX's CPU stream (umap.py:100)."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi, _device


def pca_device(Xd: torch.Tensor, pca_dim: int, verbose: bool = False) -> torch.Tensor:
    """(n, d) device tensor -> (n, pca_dim) projection onto the leading
    eigenvectors of the covariance, ordered by decreasing eigenvalue."""
    _device.require_device("pca_device", Xd)
    Xd = _device.f32_rows(Xd)
    n, d = Xd.shape
    if not 1 <= pca_dim <= d:
        raise ValueError(f"pca_dim must lie in [1, {d}], got {pca_dim}")
    lib = _capi.lib()
    nb = ctypes.c_size_t(0)
    _capi.check(lib.mv_pca_ws_bytes(n, d, ctypes.byref(nb)), "mv_pca_ws_bytes")
    ws = _device.workspace(nb.value)
    mean = torch.empty(d, dtype=torch.float64, device=Xd.device)
    cov = torch.empty((d, d), dtype=torch.float64, device=Xd.device)
    _capi.check(lib.mv_pca_cov(Xd.data_ptr(), n, d, mean.data_ptr(), cov.data_ptr(), ws.data_ptr(), ws.numel(),
                               _device.stream()), "mv_pca_cov")
    vals, vecs = np.linalg.eigh(cov.cpu().numpy())
    proj = np.ascontiguousarray(vecs[:, -pca_dim:][:, ::-1], dtype=np.float32)
    if verbose:
        print(f"Variance retained: {vals[-pca_dim:].sum() / vals.sum() * 100:.1f}%")
    V = torch.from_numpy(proj).to(Xd.device)
    mean32 = mean.float().contiguous()
    out = torch.empty((n, pca_dim), dtype=torch.float32, device=Xd.device)
    _capi.check(lib.mv_pca_project(Xd.data_ptr(), n, d, mean32.data_ptr(), V.data_ptr(), pca_dim, out.data_ptr(),
                                   _device.stream()), "mv_pca_project")
    return out
