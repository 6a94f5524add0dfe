"""Pins the kNN oracle against textbook references (CPU only)."""

import numpy as np
import pytest
from scipy.spatial.distance import cdist
from sklearn.neighbors import NearestNeighbors

from oracle import knn as oknn


def test_dist64_matches_cdist():
    rng = np.random.default_rng(0)
    q = rng.standard_normal((7, 33)).astype(np.float32)
    xs = rng.standard_normal((7, 5, 33)).astype(np.float32)
    got = oknn.dist64_seq(q, xs)
    want = np.stack([cdist(q[i:i + 1].astype(np.float64), xs[i].astype(np.float64), "sqeuclidean")[0]
                     for i in range(7)])
    np.testing.assert_allclose(got, want, rtol=1e-13)


def test_exact_knn_matches_sklearn():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((600, 24)).astype(np.float32)
    idx, d = oknn.exact_knn(X, X, 10, self_base=0)
    nn = NearestNeighbors(n_neighbors=11, algorithm="brute").fit(X.astype(np.float64))
    sd, si = nn.kneighbors(X.astype(np.float64))
    # drop self (column 0 for continuous data)
    assert (si[:, 0] == np.arange(600)).all()
    np.testing.assert_array_equal(idx, si[:, 1:])
    np.testing.assert_allclose(np.sqrt(d), sd[:, 1:], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_exact_knn_matches_loops_with_ties(seed):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 3, size=(12, 3)).astype(np.float32)  # many exact ties/duplicates
    X = np.concatenate([base, base[:5]], axis=0)
    Q = X[:9]
    qs = np.arange(9)
    idx, d = oknn.exact_knn(Q, X, 6, q_self=qs)
    li, ld = oknn.knn_loops(Q, X, 6, q_self=qs)
    np.testing.assert_array_equal(idx, li)
    np.testing.assert_array_equal(d, ld)


def test_exact_knn_no_self_exclusion_and_queries():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((50, 7)).astype(np.float32)
    Q = rng.standard_normal((9, 7)).astype(np.float32)
    idx, d = oknn.exact_knn(Q, X, 5)
    li, ld = oknn.knn_loops(Q, X, 5)
    np.testing.assert_array_equal(idx, li)
    np.testing.assert_allclose(d, ld, rtol=1e-15)


def test_exact_knn_widening_path():
    # Forces the candidate set to widen: many points at one distance.
    X = np.zeros((200, 4), dtype=np.float32)
    X[1:, 0] = 1.0
    X[100:, 1] = 0.5
    idx, d = oknn.exact_knn(X[:1], X, 20, self_base=0)
    np.testing.assert_array_equal(idx[0], np.arange(1, 21))


def test_reference_knn_agrees_on_separated_data():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((400, 16)).astype(np.float32)
    ri, rd = oknn.reference_knn(X, 8)
    ei, ed = oknn.exact_knn(X, X, 8, self_base=0, return_squared=False)
    # FP32 expansion order differs from the canonical one only at near-ties.
    assert (ri == ei).mean() > 0.99
    np.testing.assert_allclose(rd, ed, rtol=2e-3, atol=2e-3)
