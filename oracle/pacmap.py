"""CPU oracle of PaCMAP / LocalMAP (test infrastructure only): pair-count
rules, the three pair samplers and the optimiser step of
mlx_vis/_pacmap/pacmap.py, restated sequentially in NumPy (FP64 for the step).

- ``decide_num_pairs``  — PaCMAP._decide_num_pairs, pacmap.py:455-495
- ``sample_neighbors``  — _sample_neighbors, pacmap.py:65-81
- ``sample_mn_pairs``   — _sample_MN_pairs, pacmap.py:84-122 (distances in
  FP64 where the reference uses MLX FP32; the pick is the second closest of
  six candidates)
- ``sample_fp_pairs``   — _sample_FP_pairs, pacmap.py:125-177, row by row with
  the same Generator calls in the same order, so that the same seed gives the
  same pairs
- ``step``              — _step_with_mn / _step_no_mn / _step_no_mn_local,
  pacmap.py:275-354 (per-pair laws from oracle/cne.pair_grad, then the Adam
  form with the bias corrections folded into lr_t, pacmap.py:300-303)
- ``lr_t``, ``weights`` — the schedule of _fit_transform_impl, pacmap.py:583-640

Pins (tests/test_oracle_pacmap.py): the pair laws are finite-difference
gradients of their losses (tests/test_oracle_cne.py); the step against a
hand-written Adam update; the samplers against their defining properties
(brute-force second-closest, exclusion sets, scaled-distance order) and the
FP sampler against a direct transcription of the reference's loop on small n.
"""

from __future__ import annotations

import math

import numpy as np

from . import cne as oc


def decide_num_pairs(n, n_neighbors, mn_ratio, fp_ratio):
    if n_neighbors is None:
        n_neighbors = 10 if n <= 10_000 else int(round(10 + 15 * (np.log10(n) - 4)))
    else:
        n_neighbors = int(n_neighbors)
    n_mn = int(round(n_neighbors * mn_ratio))
    n_fp = int(round(n_neighbors * fp_ratio))
    n_neighbors = min(n_neighbors, n - 1)
    n_mn = min(n_mn, n - 1)
    n_fp = min(n_fp, max(0, n - 1 - n_neighbors))
    if n_neighbors + n_mn + n_fp >= n:
        denom = 1.0 + mn_ratio + fp_ratio
        n_neighbors = max(1, int(n / denom))
        n_mn = int(n_neighbors * mn_ratio)
        n_fp = int(n_neighbors * fp_ratio)
        n_neighbors = min(n_neighbors, n - 1)
        n_mn = min(n_mn, n - 1)
        n_fp = min(n_fp, max(0, n - 1 - n_neighbors))
    if n_neighbors < 1:
        raise ValueError("The number of nearest neighbours can't be less than 1")
    return n_neighbors, n_mn, n_fp


def sample_neighbors(knn_dist, knn_idx, n_neighbors):
    """Neighbour pairs by scaled distance d^2 / (sigma_i sigma_j), sigma = mean
    of the 4th-6th neighbour distances; row by row, stable order."""
    n = knn_dist.shape[0]
    sig = np.maximum(knn_dist[:, 3:6].astype(np.float64).mean(axis=1), 1e-10)
    out = np.empty((n * n_neighbors, 2), dtype=np.int32)
    for i in range(n):
        scaled = [(float(knn_dist[i, c]) ** 2 / (sig[i] * sig[knn_idx[i, c]]), c) for c in range(knn_idx.shape[1])]
        scaled.sort(key=lambda t: t[0])
        for j in range(n_neighbors):
            out[i * n_neighbors + j] = (i, knn_idx[i, scaled[j][1]])
    return out


def sample_mn_pairs(X, n_mn, rng):
    """For each of n_mn rounds and each row: six candidates (i + offset) % n
    with offsets drawn in [1, n); the second closest becomes the pair."""
    n = X.shape[0]
    if n_mn == 0:
        return np.empty((0, 2), dtype=np.int32)
    X = np.asarray(X, dtype=np.float64)
    out = np.empty((n * n_mn, 2), dtype=np.int32)
    for j in range(n_mn):
        offsets = rng.integers(1, n, size=(n, 6)).astype(np.int32)
        for i in range(n):
            cand = (i + offsets[i]) % n
            d = ((X[cand] - X[i]) ** 2).sum(axis=1)
            order = sorted(range(6), key=lambda t: (d[t], t))
            out[j * n + i] = (i, cand[order[1]])
    return out


def sample_fp_pairs(n, pair_neighbors, n_neighbors, n_fp, rng):
    """Further pairs, row by row: rounds of uniform draws, skipping self, the
    row's neighbours and repeats, until n_fp are chosen (pacmap.py:125-177)."""
    if n_fp == 0:
        return np.empty((0, 2), dtype=np.int32)
    out = np.empty((n * n_fp, 2), dtype=np.int32)
    rows = pair_neighbors[:, 1].reshape(n, n_neighbors) if n_neighbors else None
    factor = max(16, 4 * n_fp)
    for i in range(n):
        chosen, seen = [], set()
        reject = {i}
        if rows is not None:
            reject.update(int(x) for x in rows[i])
        rounds = 0
        while len(chosen) < n_fp and rounds < 24:
            remaining = n_fp - len(chosen)
            cands = rng.integers(0, n, size=max(factor, remaining * factor), dtype=np.int32)
            for c in cands.tolist():
                if c in reject or c in seen:
                    continue
                chosen.append(c)
                seen.add(c)
                if len(chosen) == n_fp:
                    break
            rounds += 1
        if len(chosen) < n_fp:
            unavailable = np.fromiter(reject.union(seen), dtype=np.int32)
            pool = np.setdiff1d(np.arange(n, dtype=np.int32), unavailable)
            need = n_fp - len(chosen)
            if pool.size:
                chosen.extend(rng.choice(pool, size=need, replace=pool.size < need).astype(np.int32).tolist())
            else:
                pool = np.setdiff1d(np.arange(n, dtype=np.int32), np.array([i], dtype=np.int32))
                chosen.extend(rng.choice(pool, size=need, replace=pool.size < need).astype(np.int32).tolist())
        out[i * n_fp:(i + 1) * n_fp, 0] = i
        out[i * n_fp:(i + 1) * n_fp, 1] = chosen[:n_fp]
    return out


def lr_t(lr, itr, b1=0.9, b2=0.999):
    """Step size of iteration itr (0-based): lr sqrt(1 - b2^(t+1)) / (1 - b1^(t+1))."""
    return lr * math.sqrt(1.0 - b2 ** (itr + 1)) / (1.0 - b1 ** (itr + 1))


def weights(itr, phases):
    """(w_nb, w_mn, w_fp) of iteration itr; w_mn None once mid-near pairs stop."""
    p1, p2, _ = phases
    if itr < p1:
        return 2.0, (1.0 - itr / p1) * 1000.0 + (itr / p1) * 3.0, 1.0
    if itr < p1 + p2:
        return 3.0, 3.0, 1.0
    return 1.0, None, 1.0


def step(Y, m, v, nn, mn, fp, w_nb, w_mn, w_fp, lr_step, local_scale=None):
    """One optimiser step, FP64. mn is None (or w_mn None) without mid-near
    pairs; local_scale selects LocalMAP's neighbour law."""
    Y = np.asarray(Y, dtype=np.float64)
    if local_scale is None:
        g = oc.pair_grad(Y, nn[:, 0], nn[:, 1], "pacmap_nn", w=w_nb)
    else:
        g = oc.pair_grad(Y, nn[:, 0], nn[:, 1], "localmap_nn", w=w_nb, scale=local_scale)
    if mn is not None and w_mn is not None and len(mn):
        g = g + oc.pair_grad(Y, mn[:, 0], mn[:, 1], "pacmap_mn", w=w_mn)
    if len(fp):
        g = g + oc.pair_grad(Y, fp[:, 0], fp[:, 1], "pacmap_fp", w=w_fp)
    m = 0.9 * m + 0.1 * g
    v = 0.999 * v + 0.001 * g * g
    Y = Y - lr_step * m / (np.sqrt(v) + 1e-7)
    return Y, m, v
