"""Spectral initialisation: drop-in for UMAP._spectral_init (umap.py:298-357).

The power iteration runs in libmivis (mv_umap_spectral); the final scaling is
the reference's host NumPy code path: drop the first vector, scale to
max |v| = 10, add N(0, 1e-4) noise from RandomState(random_state or 42), then
min-max each column to [0, 10] (umap.py:339-348). Any failure (including a
non-finite result) falls back to N(0, 1) * 0.01 as the reference's
``except Exception`` does (umap.py:354-357).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi, _device


def spectral_init(g, n: int, n_components: int, random_state=None, verbose: bool = False,
                  iters: int = 100) -> torch.Tensor:
    dev = g.knn_idx.device
    try:
        nc = n_components + 1
        start = np.random.RandomState(random_state).standard_normal((n, nc)).astype(np.float32)
        V = torch.from_numpy(start).to(dev)
        lib = _capi.lib()
        nbytes = ctypes.c_size_t(0)
        _capi.check(lib.mv_umap_spectral_ws_bytes(n, g.k, nc, ctypes.byref(nbytes)), "mv_umap_spectral_ws_bytes")
        ws = _device.workspace(nbytes.value)
        _capi.check(lib.mv_umap_spectral(g.w_sym.data_ptr(), g.eps.data_ptr(), g.knn_idx.data_ptr(), n, g.k, nc,
                                         int(iters), V.data_ptr(), ws.data_ptr(), ws.numel(), _device.stream()),
                    "mv_umap_spectral")
        emb = V[:, 1:nc].cpu().numpy()
        if not np.isfinite(emb).all():
            raise FloatingPointError("non-finite spectral embedding")
        rng = np.random.RandomState(random_state or 42)
        expansion = 10.0 / np.max(np.abs(emb))
        emb = (emb * expansion).astype(np.float32)
        emb += rng.normal(scale=0.0001, size=emb.shape).astype(np.float32)
        emb = 10.0 * (emb - emb.min(axis=0)) / (emb.max(axis=0) - emb.min(axis=0) + 1e-10)
        if verbose:
            print("Using spectral initialization")
        return torch.from_numpy(np.ascontiguousarray(emb, dtype=np.float32)).to(dev)
    except (_capi.MivisError, ValueError):
        raise
    except Exception as e:  # noqa: BLE001 - mirrors the reference's fallback
        if verbose:
            print(f"Spectral init failed ({e}), using random initialization")
        rs = np.random.RandomState(random_state)
        Y = (rs.standard_normal((n, n_components)) * 0.01).astype(np.float32)
        return torch.from_numpy(Y).to(dev)
