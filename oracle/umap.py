"""UMAP oracle: NumPy restatement of mlx_vis/_umap/umap.py (test infrastructure
only; see oracle/__init__.py).

Arithmetic follows the reference's dtypes: the fuzzy set and the schedule are
FP32 (MLX float32 arrays, NumPy float32 arrays with weak Python scalars), the
a/b fit is FP64 NumPy. ``sgd_step`` is available in FP64 (the checker of the
single-step test) and FP32.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def fuzzy_rows(knn_dists: np.ndarray, k: int):
    """rho, sigma bisection and membership weights — umap.py:203-240.

    Returns (weights (n, k) f32, sigma (n,) f32, rho (n,) f32).
    """
    d = np.asarray(knn_dists, dtype=F32)
    n = d.shape[0]
    target = F32(np.log2(k))                                   # umap.py:204
    pos = np.where(d > 0, d, F32(np.inf))                      # umap.py:209-210
    rho = np.maximum(pos.min(axis=1), F32(1e-8)).astype(F32)   # umap.py:211
    lo = np.full(n, 1e-20, dtype=F32)                          # umap.py:215
    hi = np.full(n, 1e3, dtype=F32)
    sigma = np.ones(n, dtype=F32)
    sh = np.maximum(d - rho[:, None], F32(0)).astype(F32)      # umap.py:218
    tail = sh[:, 1:]                                           # umap.py:219
    for _ in range(64):                                        # umap.py:221
        vals = np.exp(-tail / sigma[:, None]).astype(F32)
        s = np.zeros(n, dtype=F32)
        for j in range(tail.shape[1]):                         # sequential FP32 sum
            s = (s + vals[:, j]).astype(F32)
        conv = np.abs(s - target) < F32(1e-5)
        too_high = (s > target) & ~conv
        too_low = (s < target) & ~conv
        hi = np.where(too_high, sigma, hi)
        lo = np.where(too_low, sigma, lo)
        sigma = np.where(too_high, (lo + sigma) / F32(2), sigma).astype(F32)
        sigma = np.where(too_low, np.where(hi >= F32(1e3), sigma * F32(2), (sigma + hi) / F32(2)),
                         sigma).astype(F32)
        if conv.all():
            break
    w = np.exp(-sh / np.maximum(sigma[:, None], F32(1e-10))).astype(F32)   # umap.py:239
    return w, sigma, rho


def symmetrize(knn_idx: np.ndarray, w: np.ndarray) -> np.ndarray:
    """w_sym = w + w_rev - w*w_rev on the forward edges — umap.py:243-262.

    int64 keys row*n + col, a stable argsort, and a left-sided binary search
    for the reverse key (umap.py:8-22 is np.searchsorted(side='left')).
    """
    n, k = knn_idx.shape
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    cols = knn_idx.reshape(-1).astype(np.int64)
    vals = np.asarray(w, dtype=F32).reshape(-1)
    fwd = rows * n + cols
    rev = cols * n + rows
    order = np.argsort(fwd, kind="stable")
    sk = fwd[order]
    sv = vals[order]
    pos = np.minimum(np.searchsorted(sk, rev, side="left"), sk.size - 1)
    matched = sk[pos] == rev
    w_rev = np.where(matched, sv[pos], F32(0)).astype(F32)
    return (vals + w_rev - vals * w_rev).astype(F32).reshape(n, k)


def prune(w_sym: np.ndarray, n_epochs: int):
    """Keep mask w_sym >= max(w_sym)/n_epochs — umap.py:265-270."""
    thr = F32(w_sym.max()) / F32(n_epochs)
    return w_sym >= thr


def fuzzy_simplicial_set(knn_idx, knn_dists, n_epochs: int):
    """The reference's (rows, cols, vals) edge list, compacted in row-major order."""
    n, k = knn_idx.shape
    w, _, _ = fuzzy_rows(knn_dists, k)
    ws = symmetrize(knn_idx, w)
    keep = prune(ws, n_epochs).reshape(-1)
    rows = np.repeat(np.arange(n, dtype=np.int32), k)[keep]
    cols = knn_idx.reshape(-1).astype(np.int32)[keep]
    return rows, cols, ws.reshape(-1)[keep]


def find_ab_params(spread: float, min_dist: float):
    """Gauss-Newton least squares of 1/(1 + a x^{2b}) — umap.py:272-296."""
    xv = np.linspace(0, spread * 3, 300)
    yv = np.where(xv < min_dist, 1.0, np.exp(-(xv - min_dist) / spread))
    a, b = 1.0, 1.0
    for _ in range(100):
        x2b = np.power(xv, 2 * b)
        denom = 1.0 + a * x2b
        residual = 1.0 / denom - yv
        da = -x2b / (denom * denom)
        db = -a * 2.0 * np.log(np.maximum(xv, 1e-20)) * x2b / (denom * denom)
        step = np.linalg.lstsq(np.column_stack([da, db]), -residual, rcond=None)[0]
        a += step[0]
        b += step[1]
        if np.sum(step ** 2) < 1e-12:
            break
    return float(a), float(b)


def epochs_per_sample(weights: np.ndarray, n_epochs: int) -> np.ndarray:
    """umap.py:400-403 in FP32 (float32 array / weak Python scalars)."""
    w = np.asarray(weights, dtype=F32)
    max_w = float(w.max())
    n_samples = (n_epochs * (w / F32(max_w))).astype(F32)
    return np.where(n_samples > 0, F32(n_epochs) / n_samples, F32(-1.0)).astype(F32)


def schedule(weights: np.ndarray, n_epochs: int):
    """Active edge sets per epoch — umap.py:404-414. Returns a list of index arrays."""
    eps = epochs_per_sample(weights, n_epochs)
    epn = eps.copy()
    sets = []
    for epoch in range(n_epochs):
        active = np.where(epn <= epoch)[0]
        if active.size:
            epn[active] = (epn[active] + eps[active]).astype(F32)
        sets.append(active)
    return sets


def sgd_step(Y, ef, et, nf, nt, alpha, a, b, dtype=np.float64):
    """One synchronous _sgd_step — umap.py:361-387 (np.add.at scatters)."""
    Y = np.asarray(Y, dtype=dtype)
    a = dtype(a)
    b = dtype(b)
    alpha = dtype(alpha)
    out = Y.copy()
    diff = Y[ef] - Y[et]
    d2 = np.maximum((diff * diff).sum(axis=1, keepdims=True), dtype(1e-6))
    coeff = dtype(-2.0) * a * b * np.power(d2, b - dtype(1.0)) / (dtype(1.0) + a * np.power(d2, b))
    pos = np.clip(coeff * diff, -4.0, 4.0).astype(dtype) * alpha
    nd = Y[nf] - Y[nt]
    n2 = np.maximum((nd * nd).sum(axis=1, keepdims=True), dtype(1e-6))
    ncoeff = dtype(2.0) * b / ((dtype(0.001) + n2) * (dtype(1.0) + a * np.power(n2, b)))
    neg = np.clip(ncoeff * nd, -4.0, 4.0).astype(dtype) * alpha
    np.add.at(out, ef, pos)
    np.add.at(out, et, -pos)
    np.add.at(out, nf, neg)
    return out


def optimize(rows, cols, vals, Y, a, b, n_epochs, neg_rate=5, lr=1.0, seed=0, dtype=np.float32):
    """The epoch loop of umap.py:389-449 with a NumPy RNG for the negatives
    (the reference draws them from MLX's RNG, umap.py:436)."""
    n = Y.shape[0]
    rng = np.random.default_rng(seed)
    sets = schedule(vals, n_epochs)
    Y = np.asarray(Y, dtype=dtype).copy()
    for epoch in range(n_epochs):
        act = sets[epoch]
        if act.size == 0:
            continue
        ef = rows[act]
        et = cols[act]
        A = act.size
        nf = ef[np.arange(neg_rate * A) % A]
        nt = rng.integers(0, n, size=neg_rate * A)
        alpha = lr * (1.0 - epoch / n_epochs)
        Y = sgd_step(Y, ef, et, nf, nt, np.float32(alpha), np.float32(a), np.float32(b), dtype=dtype)
    return Y


def spectral_power(rows, cols, vals, n, V0, iters=100):
    """Power iteration with modified Gram-Schmidt — umap.py:305-337 (FP64)."""
    deg = np.zeros(n)
    np.add.at(deg, rows, vals.astype(np.float64))
    dis = 1.0 / np.sqrt(np.maximum(deg, 1e-10))
    wn = vals * dis[rows] * dis[cols]
    V = np.asarray(V0, dtype=np.float64).copy()
    kk = V.shape[1]
    for _ in range(iters):
        out = np.zeros_like(V)
        np.add.at(out, rows, wn[:, None] * V[cols])
        V = out
        for j in range(kk):
            for i in range(j):
                V[:, j] -= (V[:, j] @ V[:, i]) * V[:, i]
            V[:, j] /= np.sqrt(V[:, j] @ V[:, j] + 1e-10)
    return V


def spectral_scale(V, noise):
    """Scaling of the spectral vectors to [0, 10] — umap.py:339-348, FP32 as
    the reference's NumPy code: drop column 0, scale to max |v| = 10, add the
    noise (already N(0, 1); scaled by 1e-4 here), min-max each column."""
    e = np.asarray(V, dtype=F32)[:, 1:]
    expansion = F32(10.0) / np.max(np.abs(e))
    e = (e * expansion).astype(F32)
    e = e + (np.asarray(noise, dtype=F32) * F32(1e-4)).astype(F32)
    lo, hi = e.min(axis=0), e.max(axis=0)
    return (F32(10.0) * (e - lo) / (hi - lo + F32(1e-10))).astype(F32)


def sgd_step_index_add(Y, ef, et, nf, nt, alpha, a, b):
    """_sgd_step (umap.py:361-387) in FP32 with torch-CPU index_add_ for the
    three scatter-adds (the reference's `.at[].add`), for the CPU baseline:
    same arithmetic as ``sgd_step(dtype=np.float32)``, faster scatters than
    np.add.at. Y: (n, dim) float32 torch tensor (CPU); index arrays int64
    tensors; returns the new Y."""
    import torch

    a = torch.tensor(a, dtype=torch.float32)
    b = torch.tensor(b, dtype=torch.float32)
    diff = Y[ef] - Y[et]
    d2 = torch.clamp((diff * diff).sum(dim=1, keepdim=True), min=1e-6)
    coeff = -2.0 * a * b * torch.pow(d2, b - 1.0) / (1.0 + a * torch.pow(d2, b))
    pos = torch.clamp(coeff * diff, -4.0, 4.0) * alpha
    nd = Y[nf] - Y[nt]
    n2 = torch.clamp((nd * nd).sum(dim=1, keepdim=True), min=1e-6)
    ncoeff = 2.0 * b / ((0.001 + n2) * (1.0 + a * torch.pow(n2, b)))
    neg = torch.clamp(ncoeff * nd, -4.0, 4.0) * alpha
    out = Y.clone()
    out.index_add_(0, ef, pos)
    out.index_add_(0, et, -pos)
    out.index_add_(0, nf, neg)
    return out


def fit_transform_random_init(X, n_neighbors=15, n_epochs=200, min_dist=0.1, neg_rate=5, seed=0):
    """The reference's fit_transform (umap.py:75-144) with a random-uniform
    [-10, 10] init, restated for the CPU baseline: the FP32 brute kNN with
    full-row argsort (knn.reference_knn), the fuzzy set, a/b, the schedule and
    the epoch loop with torch-CPU scatters. Returns the (n, 2) embedding."""
    import torch

    from .knn import reference_knn

    X = np.asarray(X, dtype=F32)
    n = X.shape[0]
    idx, dist = reference_knn(X, n_neighbors)
    rows, cols, vals = fuzzy_simplicial_set(idx, dist, n_epochs)
    a, b = find_ab_params(1.0, min_dist)
    rng = np.random.default_rng(seed)
    Y = torch.from_numpy(rng.uniform(-10.0, 10.0, size=(n, 2)).astype(F32))
    sets = schedule(vals, n_epochs)
    r_t, c_t = torch.from_numpy(rows.astype(np.int64)), torch.from_numpy(cols.astype(np.int64))
    for epoch in range(n_epochs):
        act = sets[epoch]
        if act.size == 0:
            continue
        A = act.size
        at = torch.from_numpy(act.astype(np.int64))
        ef, et = r_t[at], c_t[at]
        nf = ef[torch.arange(neg_rate * A) % A]
        nt = torch.from_numpy(rng.integers(0, n, size=neg_rate * A))
        Y = sgd_step_index_add(Y, ef, et, nf, nt, float(1.0 - epoch / n_epochs), a, b)
    return Y.numpy()
