"""NNDescent on MI355X: drop-in for ``mlx_vis.NNDescent``
(mlx_vis/_nndescent/nndescent.py:14-403).

Same constructor (k=15, n_iters=20, max_candidates=None, delta=0.015,
random_state=None, verbose=False), same ``build(X) -> (indices int32 (n, k),
Euclidean distances float32 (n, k))`` with ``neighbor_graph`` set, same
clamp k = min(k, n - 1).

The random draws are the reference's NumPy legacy stream in the same order:
seed, init indices (nndescent.py:54-56), the 100-D projection (75), the 300-D
projection (82, drawn eagerly). The whole iteration schedule then runs inside
libmivis (mv_nnd_build).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi, _device


# Largest k of the device NNDescent: one wavefront per row holds the row's k
# neighbours, one per lane (two per lane from k = 65); its candidates (the
# current k, mc * k neighbours of neighbours and k reverse ones) are taken
# 1,024 at a time.
NND_K_MAX = 128


def nndescent_fits(k: int, max_candidates: int | None = None) -> bool:
    """Whether the device NNDescent handles this k (any max_candidates)."""
    return 1 <= k <= NND_K_MAX


class NNDescent:
    """Approximate k-NN graph via NNDescent on the GPU."""

    def __init__(self, k: int = 15, n_iters: int = 20, max_candidates: int | None = None, delta: float = 0.015,
                 random_state: int | None = None, verbose: bool = False):
        self.k = k
        self.n_iters = n_iters
        self.max_candidates = max_candidates
        self.delta = delta
        self.random_state = random_state
        self.verbose = verbose
        self.neighbor_graph = None
        self.iterations_ = None
        self.changed_ = None

    def _draws(self, n: int, d: int, k: int):
        rs = np.random.RandomState(self.random_state) if self.random_state is not None else np.random.RandomState()
        idx = rs.randint(0, n - 1, (n, k), dtype=np.int32)
        rows = np.arange(n, dtype=np.int32)[:, None]
        idx = np.where(idx >= rows, idx + 1, idx).astype(np.int32)
        p1 = min(100, d)
        p2 = min(300, d)
        # float32 draws divided by a NumPy float64 scalar promote to float64;
        # the reference's mx.array(...) stores them as float32 (nndescent.py:75, 82)
        P1 = (rs.randn(d, p1).astype(np.float32) / np.sqrt(p1)).astype(np.float32) if p1 < d else None
        P2 = (rs.randn(d, p2).astype(np.float32) / np.sqrt(p2)).astype(np.float32) if p2 < d else None
        return idx, P1, P2

    def build_device(self, Xd: torch.Tensor):
        """Device-level build: returns (indices int32, Euclidean distances float32) CUDA tensors."""
        _device.require_device("NNDescent.build_device", Xd)
        if Xd.dim() != 2:
            raise ValueError(f"X must be (n, d), got shape {tuple(Xd.shape)}")
        Xd = _device.f32_rows(Xd)
        n, d = Xd.shape
        k = min(self.k, n - 1)
        if k < 1:
            raise ValueError("NNDescent needs at least 2 points")
        idx0, P1, P2 = self._draws(n, d, k)
        dev = Xd.device
        init = torch.from_numpy(idx0).to(dev)
        P1d = torch.from_numpy(np.ascontiguousarray(P1, dtype=np.float32)).to(dev) if P1 is not None else None
        P2d = torch.from_numpy(np.ascontiguousarray(P2, dtype=np.float32)).to(dev) if P2 is not None else None
        p1 = 0 if P1 is None else P1.shape[1]
        p2 = 0 if P2 is None else P2.shape[1]
        lib = _capi.lib()
        nbytes = ctypes.c_size_t(0)
        _capi.check(lib.mv_nnd_ws_bytes(n, d, k, p1, p2, ctypes.byref(nbytes)), "mv_nnd_ws_bytes")
        ws = _device.workspace(nbytes.value)
        out_idx = torch.empty((n, k), dtype=torch.int32, device=dev)
        out_dist = torch.empty((n, k), dtype=torch.float32, device=dev)
        changed = (ctypes.c_int64 * max(self.n_iters, 1))()
        iters = ctypes.c_int(0)
        _capi.check(lib.mv_nnd_build(
            Xd.data_ptr(), n, d, k, int(self.n_iters), int(self.max_candidates or 0), float(self.delta),
            init.data_ptr(), _capi.ptr(P1d), p1, _capi.ptr(P2d), p2, int(bool(self.verbose)),
            out_idx.data_ptr(), out_dist.data_ptr(), changed, ctypes.byref(iters), ws.data_ptr(), ws.numel(),
            _device.stream()), "mv_nnd_build")
        self.iterations_ = iters.value
        self.changed_ = [int(changed[i]) for i in range(iters.value)]
        return out_idx, out_dist

    def build(self, X) -> tuple[np.ndarray, np.ndarray]:
        """Build the approximate k-NN graph.

        Returns:
            (indices int32 (n, k), distances float32 (n, k)), Euclidean.
        """
        Xd = _device.as_device_f32(X)
        idx, dist = self.build_device(Xd)
        self.neighbor_graph = (idx.cpu().numpy(), dist.cpu().numpy())
        return self.neighbor_graph
