"""Pins the NNDescent oracle (CPU only): invariants and recall."""

import numpy as np

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
