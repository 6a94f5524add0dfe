"""Synthetic workloads of the benchmark (host NumPy, fixed draw order).

gaussian_iid          n x d i.i.d. N(0, 1) float32 (10,000 x 128, seed 0).
gaussian_mixture_784  mixture of 10 isotropic Gaussians in 784-D: centres
                      N(0, 4^2) per dimension, sigma 1, equal weights, then
                      per-feature min-max to [0, 1] (the formula of
                      mlx_vis/_normalize.py:28-31) (70,000 x 784, seed 1).
mixture_fresh_784     fresh points from the same centres (new labels and
                      noise from their own seed), scaled with the min/max of
                      the seed-1 set.
gaussian_mixture_128  1,000,000 x 128 mixture of 50 Gaussians: centres
                      N(0, 3^2), sigma 1, Dirichlet(1) weights (seed 2).
"""

from __future__ import annotations

import numpy as np


def gaussian_iid(n: int = 10_000, d: int = 128, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal((n, d), dtype=np.float32)


def _mixture_raw(rng, centers, n):
    labels = rng.integers(0, centers.shape[0], size=n)
    X = rng.standard_normal((n, centers.shape[1]), dtype=np.float32)
    X += centers[labels]
    return X, labels


def mixture_784_centers(seed: int = 1, k: int = 10, d: int = 784) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((k, d)) * 4.0).astype(np.float32)


def gaussian_mixture_784(n: int = 70_000, seed: int = 1, return_labels: bool = False, return_minmax=False):
    rng = np.random.default_rng(seed)
    centers = (rng.standard_normal((10, 784)) * 4.0).astype(np.float32)
    X, labels = _mixture_raw(rng, centers, n)
    lo = X.min(axis=0)
    hi = X.max(axis=0)
    X = (X - lo) / (hi - lo + np.float32(1e-8))
    out = [X.astype(np.float32, copy=False)]
    if return_labels:
        out.append(labels)
    if return_minmax:
        out.append((lo, hi))
    return out[0] if len(out) == 1 else tuple(out)


def mixture_fresh_784(n: int = 1_000_000, seed: int = 3, base_seed: int = 1, base_n: int = 70_000,
                      return_labels: bool = False):
    _, (lo, hi) = gaussian_mixture_784(base_n, base_seed, return_minmax=True)
    centers = mixture_784_centers(base_seed)
    rng = np.random.default_rng(seed)
    X, labels = _mixture_raw(rng, centers, n)
    X = (X - lo) / (hi - lo + np.float32(1e-8))
    X = X.astype(np.float32, copy=False)
    return (X, labels) if return_labels else X


def gaussian_mixture_128(n: int = 1_000_000, seed: int = 2, return_labels: bool = False):
    rng = np.random.default_rng(seed)
    weights = rng.dirichlet(np.ones(50))
    centers = (rng.standard_normal((50, 128)) * 3.0).astype(np.float32)
    labels = rng.choice(50, size=n, p=weights)
    X = rng.standard_normal((n, 128), dtype=np.float32)
    X += centers[labels]
    return (X, labels) if return_labels else X
