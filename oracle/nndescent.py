"""NNDescent oracle: NumPy restatement of mlx_vis/_nndescent/nndescent.py
(test infrastructure only; see oracle/__init__.py).

Same NumPy random stream and draw order (seed, init indices, P, P2), same
schedule (projections, FP16 iterations 1-4, reverse candidates, full and
converging paths, early stop). MLX sort tie orders are taken as stable, and
equal distances in the top-k selection are ordered by id.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def gather_dists(X, sq, ids, half=False, rows=None):
    """_gather_dists / _gather_dists_half — nndescent.py:573-641."""
    rows = np.arange(ids.shape[0]) if rows is None else rows
    out = np.empty(ids.shape, dtype=F32)
    for s in range(0, ids.shape[0], 2048):
        r = rows[s:s + 2048]
        c = ids[s:s + 2048]
        if half:
            Xh = X.astype(np.float16)
            dots = np.einsum("id,icd->ic", Xh[r].astype(F32), Xh[c].astype(F32)).astype(np.float16).astype(F32)
        else:
            dots = np.einsum("id,icd->ic", X[r], X[c]).astype(F32)
        out[s:s + 2048] = np.maximum(sq[r][:, None] + sq[c] - F32(2.0) * dots, F32(0.0))
    return out


def reverse_candidates(indices):
    """_compute_reverse_candidates — nndescent.py:406-436 (stable argsort)."""
    n, k = indices.shape
    src = np.repeat(np.arange(n, dtype=np.int32), k)
    dst = indices.reshape(-1)
    order = np.argsort(dst, kind="stable")
    rs, rd = src[order], dst[order]
    new_group = np.concatenate([[True], rd[1:] != rd[:-1]])
    pos = np.arange(n * k)
    starts = np.maximum.accumulate(np.where(new_group, pos, 0))
    within = pos - starts
    keep = within < k
    out = np.zeros(n * k, dtype=np.int32)
    out[(rd * k + within)[keep]] = rs[keep]
    return out.reshape(n, k)


def dedup_topk(cands, dists, k):
    """_dedup_topk_block — nndescent.py:439-458: later duplicates (in stable
    id order) get 1e30; the k smallest by (distance, id)."""
    order = np.argsort(cands, axis=1, kind="stable")
    sc = np.take_along_axis(cands, order, 1)
    sd = np.take_along_axis(dists, order, 1).copy()
    dup = np.zeros_like(sc, dtype=bool)
    dup[:, 1:] = sc[:, 1:] == sc[:, :-1]
    sd[dup] = F32(1e30)
    sel = np.lexsort((sc, sd), axis=1)[:, :k]
    return np.take_along_axis(sc, sel, 1).astype(np.int32), np.take_along_axis(sd, sel, 1)


def schedule_params(k, mc, uf):
    """(mc_nn, j_nn) of an iteration — nndescent.py:100-101: how many of a
    row's neighbours are expanded and how many of each one's neighbours are
    taken."""
    sc = min(1.0, uf ** 0.5 * 3)
    mc_nn = max(3, int(min(k, mc) * sc))
    j_nn = max(k // 2, min(k, int(k * sc)))
    return mc_nn, j_nn


def converging_step(X, sq, idx, dists, flags, uf, k, mc):
    """One iteration of the converging path (update_frac < 0.5) —
    nndescent.py:103-220: NEW neighbours first (stable), mc_new = max(3,
    int(k * uf)) of them expanded by j_nn of their neighbours, reverse
    candidates while uf >= 0.10, only rows with a NEW flag (the others keep
    their row), FP32 full-d distances of the new candidates merged with the
    STORED distances of the current ones. Returns (idx, dists)."""
    n = idx.shape[0]
    _, j_nn = schedule_params(k, mc, uf)
    new_order = np.argsort((~flags).astype(np.int32), axis=1, kind="stable")
    new_first = np.take_along_axis(idx, new_order, 1)
    mc_new = max(3, int(k * uf))
    active = np.where(flags.any(axis=1))[0]
    if active.size == 0:
        # no row changed (reachable only with delta <= 0; the reference's row
        # loop then has a zero step, nndescent.py:135-136): nothing to update
        return idx.copy(), dists.copy()
    use_rev = uf >= 0.10
    rev = reverse_candidates(idx) if use_rev else None
    nn = idx[new_first[active, :mc_new].reshape(-1)].reshape(active.size, -1, k)[:, :, :j_nn]
    nc = nn.reshape(active.size, -1)
    if use_rev:
        nc = np.concatenate([nc, rev[active]], axis=1)
    nd = gather_dists(X, sq, nc, rows=active)
    ac = np.concatenate([idx[active], nc], axis=1)
    ad = np.concatenate([dists[active], nd], axis=1)
    ad = np.where(ac == active[:, None], F32(1e30), ad)
    ri, rd = dedup_topk(ac, ad, k)
    new_idx = idx.copy()
    new_d = dists.copy()
    new_idx[active] = ri
    new_d[active] = rd
    return new_idx, new_d


def full_step(Xd, sqd, idx, dists, uf, it, k, mc):
    """One iteration of the full path (update_frac >= 0.5) —
    nndescent.py:221-378 without the iteration-4 recompute: every row expands
    its first mc_nn neighbours by j_nn of theirs, plus reverse candidates
    while uf >= 0.10; distances in the space (Xd, sqd) given for this
    iteration, FP16 dots in iterations 1-4, the current neighbours' distances
    recomputed in iterations 0-4 (stored ones reused after). Returns (idx,
    dists)."""
    n = idx.shape[0]
    mc_nn, j_nn = schedule_params(k, mc, uf)
    use_half = it in (1, 2, 3, 4)
    recompute = it == 0 or use_half
    use_rev = uf >= 0.10
    nn = idx[idx[:, :mc_nn].reshape(-1)].reshape(n, -1, k)[:, :, :j_nn].reshape(n, -1)
    nc = np.concatenate([nn, reverse_candidates(idx)], axis=1) if use_rev else nn
    nd = gather_dists(Xd, sqd, nc, half=use_half)
    cd = gather_dists(Xd, sqd, idx, half=use_half) if recompute else dists
    ac = np.concatenate([idx, nc], axis=1)
    ad = np.concatenate([cd, nd], axis=1)
    ad = np.where(ac == np.arange(n)[:, None], F32(1e30), ad)
    return dedup_topk(ac, ad, k)


def nndescent(X, k=15, n_iters=20, max_candidates=None, delta=0.015, random_state=None, return_trace=False):
    """NNDescent.build — nndescent.py:34-403. Returns (idx, Euclidean dists)
    (and the changed count of every iteration run with return_trace)."""
    X = np.asarray(X, dtype=F32)
    n, d = X.shape
    k = min(k, n - 1)
    mc = max_candidates or k
    rs = np.random.RandomState(random_state) if random_state is not None else np.random.RandomState()
    sq = np.einsum("ij,ij->i", X, X).astype(F32)
    idx = rs.randint(0, n - 1, (n, k), dtype=np.int32)
    idx = np.where(idx >= np.arange(n, dtype=np.int32)[:, None], idx + 1, idx).astype(np.int32)
    dists = gather_dists(X, sq, idx)
    si = np.argsort(dists, axis=1, kind="stable")
    idx = np.take_along_axis(idx, si, 1)
    dists = np.take_along_axis(dists, si, 1)
    p1, p2 = min(100, d), min(300, d)
    if p1 < d:
        P = (rs.randn(d, p1).astype(F32) / np.sqrt(p1)).astype(F32)   # mx.array stores float32
        X_low = (X @ P).astype(F32)
        sq_low = np.einsum("ij,ij->i", X_low, X_low).astype(F32)
    else:
        X_low, sq_low = X, sq
    P2 = (rs.randn(d, p2).astype(F32) / np.sqrt(p2)).astype(F32) if p2 < d else None
    uf = 1.0
    flags = np.ones((n, k), dtype=bool)
    trace = []
    for it in range(n_iters):
        if uf < 0.5:
            new_idx, new_d = converging_step(X, sq, idx, dists, flags, uf, k, mc)
        else:
            if it == 0:
                Xd, sqd = X_low, sq_low
            elif it == 1 and P2 is not None:
                Xd = (X @ P2).astype(F32)
                sqd = np.einsum("ij,ij->i", Xd, Xd).astype(F32)
            else:
                Xd, sqd = X, sq
            new_idx, new_d = full_step(Xd, sqd, idx, dists, uf, it, k, mc)
            if it == 4:
                new_d = gather_dists(X, sq, new_idx)
        flags = new_idx != idx
        changed = int(flags.sum())
        uf = changed / (n * k)
        trace.append(changed)
        idx, dists = new_idx, new_d
        if uf < delta:
            break
    out = (idx, np.sqrt(np.maximum(dists, 0)).astype(F32))
    return (out + (trace,)) if return_trace else out
