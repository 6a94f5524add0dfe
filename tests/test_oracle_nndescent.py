"""Pins the NNDescent oracle (CPU only): invariants and recall."""

import numpy as np
import pytest

from oracle import knn as oknn
from oracle import metrics as om
from oracle import nndescent as onn


def _data(n=1500, d=40, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((6, d)) * 3
    return (c[rng.integers(0, 6, n)] + rng.standard_normal((n, d))).astype(np.float32)


def test_reverse_candidates_definition():
    rng = np.random.default_rng(1)
    n, k = 30, 4
    idx = np.stack([rng.choice(np.delete(np.arange(n), i), k, replace=False) for i in range(n)]).astype(np.int32)
    rev = onn.reverse_candidates(idx)
    for j in range(n):
        srcs = [i for i in range(n) for c in range(k) if idx[i, c] == j][:k]
        want = srcs + [0] * (k - len(srcs))
        assert rev[j].tolist() == want


def test_dedup_topk_rules():
    cands = np.array([[5, 3, 5, 9, 3, 1]], dtype=np.int32)
    dists = np.array([[2.0, 1.0, 0.5, 3.0, 0.1, 4.0]], dtype=np.float32)
    ri, rd = onn.dedup_topk(cands, dists, 4)
    # id 3 keeps its first slot (1.0), id 5 keeps its first slot (2.0)
    assert ri.tolist() == [[3, 5, 9, 1]]
    np.testing.assert_allclose(rd, [[1.0, 2.0, 3.0, 4.0]])


def test_oracle_invariants_and_recall():
    X = _data()
    idx, dist = onn.nndescent(X, k=15, n_iters=12, random_state=0)
    n = X.shape[0]
    assert (idx != np.arange(n)[:, None]).all()
    assert all(len(set(r)) == 15 for r in idx)
    assert (np.diff(dist, axis=1) >= 0).all()
    exact, _ = oknn.exact_knn(X, X, 15, self_base=0)
    assert om.recall(idx, exact) > 0.9
    # Stored distances are distances of the stored ids; entries last computed
    # in an FP16 iteration (1-4) keep FP16-rounded dots, as in the reference
    # when it leaves the full path before its iteration-4 FP32 recompute.
    d2 = ((X[:, None, :] - X[idx]) ** 2).sum(-1)
    np.testing.assert_allclose(dist, np.sqrt(d2), rtol=2e-2, atol=0.1)


def test_gather_dists_exact_on_integer_points():
    """oracle gather_dists (nndescent.py:573-641) against pairwise squared
    distances computed in exact integer arithmetic: with small-integer
    coordinates every dot product is exact in FP32 and in FP16, so both the
    FP32 and the FP16-dot variant must return the exact values; on general
    floats the FP32 variant stays within FP32 rounding of the FP64 values."""
    rng = np.random.default_rng(13)
    n, d, c = 300, 8, 7
    Xi = rng.integers(0, 4, size=(n, d))
    X = Xi.astype(np.float32)
    sq = (X * X).sum(1).astype(np.float32)
    ids = rng.integers(0, n, size=(n, c))
    exact = ((Xi[:, None, :] - Xi[ids]) ** 2).sum(-1).astype(np.float32)
    np.testing.assert_array_equal(onn.gather_dists(X, sq, ids), exact)
    np.testing.assert_array_equal(onn.gather_dists(X, sq, ids, half=True), exact)
    rows = np.array([5, 0, 299, 17])
    want_rows = ((Xi[rows][:, None, :] - Xi[ids[:4]]) ** 2).sum(-1).astype(np.float32)
    np.testing.assert_array_equal(onn.gather_dists(X, sq, ids[:4], rows=rows), want_rows)
    Xf = rng.standard_normal((n, 16)).astype(np.float32)
    sqf = (Xf * Xf).sum(1).astype(np.float32)
    ref = ((Xf[:, None, :].astype(np.float64) - Xf[ids].astype(np.float64)) ** 2).sum(-1)
    got = onn.gather_dists(Xf, sqf, ids)
    assert np.all(np.abs(got - ref) <= 1e-5 * (sqf[:, None] + sqf[ids]) + 1e-6)


# --------------------------------------------------------------------------
# The iteration schedule of NNDescent.build (nndescent.py:94-398), pinned by
# what each iteration must return rather than by restating its code. The data
# is i.i.d. N(0, 1) in 320-D (2,000 rows, k = 15, seed 0): it converges slowly
# enough that iterations 0-5 run the full path (update_frac >= 0.5) and
# iteration 6 the converging path, so each kind of distance shows in the
# result of a run stopped after it.

_N, _D, _K = 2000, 320, 15


def _iid():
    return np.random.default_rng(1).standard_normal((_N, _D)).astype(np.float32)


def _projections(X, seed=0):
    """The reference's NumPy draws in order (nndescent.py:48, 54, 75, 82):
    init indices, then P (d x 100), then P2 (d x 300), both / sqrt(dim)."""
    rs = np.random.RandomState(seed)
    rs.randint(0, X.shape[0] - 1, (X.shape[0], _K), dtype=np.int32)
    P = (rs.randn(X.shape[1], 100).astype(np.float32) / np.sqrt(100)).astype(np.float32)
    P2 = (rs.randn(X.shape[1], 300).astype(np.float32) / np.sqrt(300)).astype(np.float32)
    return P, P2


def _d64(A, idx):
    """FP64 direct-difference squared distances of the pairs (i, idx[i, c])."""
    A = A.astype(np.float64)
    return ((A[:, None, :] - A[idx]) ** 2).sum(-1)


def _run(n_iters, delta=1e-9):
    idx, dist, trace = onn.nndescent(_iid(), k=_K, n_iters=n_iters, delta=delta, random_state=0, return_trace=True)
    return idx, dist.astype(np.float64) ** 2, [c / (_N * _K) for c in trace]


def test_trace_keeps_the_full_path_through_iteration_5():
    _, _, uf = _run(7)
    assert len(uf) == 7
    assert min(uf[:5]) >= 0.5 and uf[5] < 0.5      # iterations 0-5 full, 6 converging


def test_iteration_0_distances_live_in_the_100d_projection():
    """n_iters = 1: the returned distances are those of X P (nndescent.py:
    68-78, 227-228), not of X; a missing projection fails both checks."""
    X = _iid()
    P, _ = _projections(X)
    idx, d2, _ = _run(1)
    proj = _d64((X.astype(np.float64) @ P.astype(np.float64)).astype(np.float32), idx)
    full = _d64(X, idx)
    sq = (X.astype(np.float64) @ P.astype(np.float64)) ** 2
    scale = sq.sum(1)[:, None] + sq.sum(1)[idx]
    assert np.all(np.abs(d2 - proj) <= 2e-5 * scale + 1e-4)
    assert np.median(np.abs(d2 - full) / full) > 0.05      # JL distortion of 100 dimensions


def test_iteration_1_distances_are_fp16_dots_in_the_300d_projection():
    """n_iters = 2 (d > 300): FP16 dots in X P2 (nndescent.py:229-236,
    615-641): within FP16 rounding of the exact projected distances, visibly
    not FP32 (some pair deviates beyond FP32 rounding), and not the 100-D or
    full-D distances."""
    X = _iid()
    P, P2 = _projections(X)
    idx, d2, _ = _run(2)
    Xp2 = (X @ P2).astype(np.float32)
    proj = _d64(Xp2, idx)
    sqn = (Xp2.astype(np.float64) ** 2).sum(1)
    scale = sqn[:, None] + sqn[idx]
    err = np.abs(d2 - proj)
    assert np.all(err <= 4e-3 * scale)
    assert err.max() > 1e-4 * scale[err.argmax() // _K, err.argmax() % _K]   # FP16, not FP32
    assert np.median(np.abs(d2 - _d64(X, idx)) / _d64(X, idx)) > 0.02
    Xp1 = (X @ P).astype(np.float32)
    assert np.median(np.abs(d2 - _d64(Xp1, idx)) / _d64(Xp1, idx)) > 0.02


def test_iteration_3_distances_are_fp16_dots_in_full_dimension():
    """n_iters = 4: iteration 3 is an FP16 iteration in the full space."""
    X = _iid()
    idx, d2, _ = _run(4)
    full = _d64(X, idx)
    sqn = (X.astype(np.float64) ** 2).sum(1)
    scale = sqn[:, None] + sqn[idx]
    err = np.abs(d2 - full)
    assert np.all(err <= 4e-3 * scale)
    assert err.max() > 1e-4 * scale.max() / 2


def test_iteration_4_recomputes_fp32_and_later_iterations_stay_fp32():
    """n_iters = 5: the FP32 recompute after iteration 4 (nndescent.py:
    376-378) makes every distance exact to FP32 rounding; iteration 5 (full
    path, stored distances reused) and iteration 6 (converging path, FP32
    distances of new candidates against stored ones) keep that."""
    X = _iid()
    sqn = (X.astype(np.float64) ** 2).sum(1)
    for n_iters in (5, 6, 7):
        idx, d2, _ = _run(n_iters)
        full = _d64(X, idx)
        scale = sqn[:, None] + sqn[idx]
        assert np.all(np.abs(d2 - full) <= 2e-6 * scale), n_iters


def test_early_stop_is_strictly_below_delta():
    """update_frac < delta stops (nndescent.py:395): delta equal to the first
    iteration's update_frac continues, the next float up stops after one."""
    _, _, uf = _run(3)
    _, _, t_eq = _run(3, delta=uf[0])
    _, _, t_up = _run(3, delta=float(np.nextafter(uf[0], 2.0)))
    assert len(t_eq) >= 2 and len(t_up) == 1
    _, _, t_one = _run(3, delta=1.5)
    assert len(t_one) == 1


def test_reverse_candidates_only_while_update_frac_at_least_0_10(monkeypatch):
    """Reverse candidates (nndescent.py:241, 121-127) in every iteration whose
    incoming update_frac is >= 0.10, on the full and the converging path, and
    in no other."""
    calls = []
    real = onn.reverse_candidates
    monkeypatch.setattr(onn, "reverse_candidates", lambda idx: calls.append(1) or real(idx))
    seen = []
    real_conv, real_full = onn.converging_step, onn.full_step

    def conv(X, sq, idx, dists, flags, uf, k, mc):
        seen.append(("conv", uf, len(calls)))
        return real_conv(X, sq, idx, dists, flags, uf, k, mc)

    def full(Xd, sqd, idx, dists, uf, it, k, mc):
        seen.append(("full", uf, len(calls)))
        return real_full(Xd, sqd, idx, dists, uf, it, k, mc)

    monkeypatch.setattr(onn, "converging_step", conv)
    monkeypatch.setattr(onn, "full_step", full)
    _run(9)
    seen.append(("end", None, len(calls)))
    paths = set()
    for (path, uf, c0), (_, _, c1) in zip(seen, seen[1:]):
        assert (c1 - c0) == (1 if uf >= 0.10 else 0), (path, uf)
        paths.add((path, uf >= 0.10))
    assert paths == {("full", True), ("conv", True), ("conv", False)}


def _crafted_state(k=12, seed=0):
    """Row 0 at the origin, its true nearest neighbour g = 1 right next to it,
    everyone else far away; every row's list holds far points only (no row
    lists 0 or 1), with FP32 distances stored."""
    rng = np.random.default_rng(seed)
    n = 200
    X = rng.uniform(0.5, 1.0, (n, 2)).astype(np.float32)
    X[0] = (0.0, 0.0)
    X[1] = (0.001, 0.0)
    far = np.arange(2, n)
    idx = np.stack([rng.choice(far[far != i], k, replace=False) for i in range(n)]).astype(np.int32)
    sq = (X * X).sum(1).astype(np.float32)
    dists = onn.gather_dists(X, sq, idx)
    return X, sq, idx, dists


@pytest.mark.parametrize("p,uf,found", [(3, 0.35, True), (4, 0.35, False), (4, 0.42, True), (2, 0.12, True)])
def test_converging_path_expands_mc_new_new_neighbours(p, uf, found):
    """mc_new = max(3, int(k * update_frac)) NEW neighbours are expanded
    (nndescent.py:107, 148): g, listed by row 0's neighbour at position p,
    is found iff p < mc_new (k = 12: uf 0.35 -> 4, 0.42 -> 5, 0.12 -> 3)."""
    X, sq, idx, dists = _crafted_state()
    k = idx.shape[1]
    idx[idx[0, p], 0] = 1
    dists = onn.gather_dists(X, sq, idx)
    flags = np.ones_like(idx, dtype=bool)
    ri, _ = onn.converging_step(X, sq, idx, dists, flags, uf, k, k)
    assert (1 in ri[0]) == found


@pytest.mark.parametrize("q,found", [(5, True), (6, False)])
def test_converging_path_takes_j_nn_of_each_neighbour(q, found):
    """j_nn = max(k // 2, min(k, int(k * min(1, 3 sqrt(uf))))) of each expanded
    neighbour's list (nndescent.py:101, 149): at k = 12, uf = 0.02 that is 6,
    so g at position q of the neighbour's list is found iff q < 6."""
    X, sq, idx, dists = _crafted_state()
    k = idx.shape[1]
    idx[idx[0, 0], q] = 1
    dists = onn.gather_dists(X, sq, idx)
    flags = np.ones_like(idx, dtype=bool)
    ri, _ = onn.converging_step(X, sq, idx, dists, flags, 0.02, k, k)
    assert (1 in ri[0]) == found


def test_converging_path_skips_rows_without_new_flags_and_keeps_stored_distances():
    """Rows whose flags are all OLD are not updated even when a better
    candidate is reachable (nndescent.py:109-114, 211-216); an updated row
    merges the STORED distances of its current neighbours (nndescent.py:
    191-195), not recomputed ones."""
    X, sq, idx, dists = _crafted_state()
    k = idx.shape[1]
    idx[idx[0, 0], 0] = 1
    idx[idx[5, 0], 0] = 1
    dists = onn.gather_dists(X, sq, idx)
    stored = dists.copy()
    stored[5] *= np.float32(0.5)          # stored values that a recompute would replace
    flags = np.ones_like(idx, dtype=bool)
    flags[0] = False                      # row 0: all OLD
    ri, rd = onn.converging_step(X, sq, idx, stored, flags, 0.3, k, k)
    np.testing.assert_array_equal(ri[0], idx[0])
    np.testing.assert_array_equal(rd[0], stored[0])
    kept = [c for c in range(k) if idx[5, c] in ri[5]]
    assert kept
    for c in kept:
        assert rd[5][list(ri[5]).index(idx[5, c])] == stored[5, c]
