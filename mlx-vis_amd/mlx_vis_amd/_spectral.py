"""Spectral initialisation: drop-in for UMAP._spectral_init (umap.py:298-357).

The start vectors (drawn by the reference from MLX's generator, umap.py:325,
a stream that cannot be reproduced here) are standard normals from
mv_random_normal, keyed by random_state; the power iteration runs in libmivis
(mv_umap_spectral); the final scaling (umap.py:339-348: drop the first vector,
scale to max |v| = 10, add N(0, 1e-4) noise, min-max each column to [0, 10])
runs on the device too (mv_umap_spectral_embed), with the noise drawn from
Philox keyed by random_state where the reference draws it from NumPy's
RandomState(random_state or 42). A non-finite spectral result falls back to
N(0, 1) * 0.01 as the reference's ``except Exception`` does (umap.py:354-357).
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
        lib = _capi.lib()
        V = torch.empty((n, nc), dtype=torch.float32, device=dev)
        seed = 0x5EC7 if random_state is None else int(random_state)
        seed = (seed * 0x9E3779B97F4A7C15 + 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF
        _capi.check(lib.mv_random_normal(V.data_ptr(), n * nc, seed, _device.stream()), "mv_random_normal")
        nbytes = ctypes.c_size_t(0)
        _capi.check(lib.mv_umap_spectral_ws_bytes(n, g.k, nc, ctypes.byref(nbytes)), "mv_umap_spectral_ws_bytes")
        ws = _device.workspace(nbytes.value)
        _capi.check(lib.mv_umap_spectral(g.w_sym.data_ptr(), g.eps.data_ptr(), g.knn_idx.data_ptr(), n, g.k, nc,
                                         int(iters), V.data_ptr(), ws.data_ptr(), ws.numel(), _device.stream()),
                    "mv_umap_spectral")
        Y = torch.empty((n, n_components), dtype=torch.float32, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        seed2 = (seed ^ 0x6A09E667F3BCC909) * 0xBF58476D1CE4E5B9 & 0xFFFFFFFFFFFFFFFF
        _capi.check(lib.mv_umap_spectral_embed(V.data_ptr(), n, nc, seed2, Y.data_ptr(), flag.data_ptr(),
                                               ws.data_ptr(), ws.numel(), _device.stream()),
                    "mv_umap_spectral_embed")
        if int(flag.item()) != 0:
            raise FloatingPointError("non-finite spectral embedding")
        if verbose:
            print("Using spectral initialization")
        return Y
    except (_capi.MivisError, ValueError):
        raise
    except Exception as e:  # noqa: BLE001 - mirrors the reference's fallback
        if verbose:
            print(f"Spectral init failed ({e}), using random initialization")
        rs = np.random.RandomState(random_state)
        Y = (rs.standard_normal((n, n_components)) * 0.01).astype(np.float32)
        return torch.from_numpy(Y).to(dev)
