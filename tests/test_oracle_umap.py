"""Pins the UMAP oracle (CPU only)."""

import numpy as np
import pytest
from scipy.optimize import curve_fit

from oracle import umap as ou
from oracle import metrics as om


def test_ab_params_match_curve_fit():
    a, b = ou.find_ab_params(1.0, 0.1)
    xv = np.linspace(0, 3.0, 300)
    yv = np.where(xv < 0.1, 1.0, np.exp(-(xv - 0.1)))
    (ca, cb), _ = curve_fit(lambda x, a, b: 1.0 / (1.0 + a * x ** (2 * b)), xv, yv, p0=(1.0, 1.0))
    assert abs(a - ca) < 1e-4 and abs(b - cb) < 1e-4
    assert abs(a - 1.577) < 0.01 and abs(b - 0.895) < 0.01


def test_product_ab_matches_oracle():
    from mlx_vis_amd import find_ab_params
    for spread, md in [(1.0, 0.1), (1.0, 0.5), (2.0, 0.05)]:
        assert find_ab_params(spread, md) == ou.find_ab_params(spread, md)


def test_fuzzy_rows_solve_their_equation():
    rng = np.random.default_rng(0)
    d = np.sort(rng.uniform(0.5, 3.0, size=(200, 15)).astype(np.float32), axis=1)
    w, sigma, rho = ou.fuzzy_rows(d, 15)
    np.testing.assert_allclose(rho, d[:, 0])
    s = np.exp(-(d[:, 1:] - rho[:, None]) / sigma[:, None]).sum(1)
    np.testing.assert_allclose(s, np.log2(15), atol=2e-4)
    np.testing.assert_allclose(w[:, 0], 1.0)


def test_symmetrize_closed_form():
    rng = np.random.default_rng(1)
    n, k = 40, 5
    idx = np.stack([rng.choice(np.delete(np.arange(n), i), k, replace=False) for i in range(n)])
    w = rng.uniform(0.1, 1.0, size=(n, k)).astype(np.float32)
    W = np.zeros((n, n))
    for i in range(n):
        W[i, idx[i]] = w[i]
    S = W + W.T - W * W.T
    got = ou.symmetrize(idx, w)
    want = np.stack([S[i, idx[i]] for i in range(n)])
    np.testing.assert_allclose(got, want, rtol=1e-6)


def test_schedule_counts_match_closed_form():
    rng = np.random.default_rng(2)
    w = rng.uniform(0.05, 1.0, size=500).astype(np.float32)
    w[7] = 1.0
    n_epochs = 200
    sets = ou.schedule(w, n_epochs)
    counts = np.zeros(500, dtype=int)
    for s in sets:
        counts[s] += 1
    eps = ou.epochs_per_sample(w, n_epochs)
    # an edge is sampled at epochs eps, 2eps, ... < n_epochs: floor((n_epochs-1)/eps) times
    # up to FP32 rounding of the running sum
    expect = np.floor((n_epochs - 1) / eps.astype(np.float64)).astype(int)
    assert np.abs(counts - expect).max() <= 1
    assert sets[0].size == 0 and all(7 in s for s in sets[1:])


def _loss_grad_fd(Y, ef, et, nf, nt, a, b, alpha):
    """Central finite differences of the UMAP edge losses whose negative
    gradients are the unclipped _sgd_step updates."""
    def att(y):
        d2 = ((y[ef] - y[et]) ** 2).sum(1)
        return (np.log1p(a * d2 ** b)).sum()
    def rep(y):
        d2 = ((y[nf] - y[nt]) ** 2).sum(1)
        return 0.0
    h = 1e-6
    G = np.zeros_like(Y)
    for i in range(Y.shape[0]):
        for q in range(Y.shape[1]):
            Yp = Y.copy(); Yp[i, q] += h
            Ym = Y.copy(); Ym[i, q] -= h
            G[i, q] = (att(Yp) - att(Ym)) / (2 * h)
    return G


def test_sgd_step_attraction_is_loss_gradient():
    """Unclipped attraction updates equal -alpha * d/dy of log(1 + a d^{2b})."""
    rng = np.random.default_rng(3)
    Y = rng.uniform(-1, 1, size=(12, 2))
    ef = np.array([0, 1, 2, 3, 4]); et = np.array([5, 6, 7, 8, 9])
    a, b, alpha = 1.577, 0.895, 0.7
    out = ou.sgd_step(Y, ef, et, np.array([], int), np.array([], int), alpha, a, b)
    G = _loss_grad_fd(Y, ef, et, None, None, a, b, alpha)
    np.testing.assert_allclose(out - Y, -alpha * G, rtol=1e-5, atol=1e-7)


def test_sgd_step_repulsion_closed_form():
    Y = np.array([[0.0, 0.0], [0.3, -0.4], [10.0, 0.0]])
    a, b, alpha = 1.5, 0.9, 1.0
    out = ou.sgd_step(Y, np.array([], int), np.array([], int), np.array([0, 0]), np.array([1, 2]), alpha, a, b)
    d2 = 0.25
    c = 2 * b / ((0.001 + d2) * (1 + a * d2 ** b))
    g1 = np.clip(c * (Y[0] - Y[1]), -4, 4)
    d2b = 100.0
    c2 = 2 * b / ((0.001 + d2b) * (1 + a * d2b ** b))
    g2 = np.clip(c2 * (Y[0] - Y[2]), -4, 4)
    np.testing.assert_allclose(out[0], Y[0] + g1 + g2)
    np.testing.assert_allclose(out[1:], Y[1:])


def test_sgd_step_index_add_matches_sgd_step():
    """The CPU baseline's torch index_add_ step against the np.add.at step on
    the same edges and negatives, in FP64 (repeated heads, tails and
    negatives, some pairs closer than the 1e-6 floor, some clipped at 4)."""
    import torch

    rng = np.random.default_rng(7)
    n, E = 60, 400
    Y = rng.uniform(-3, 3, (n, 2))
    Y[5] = Y[6]   # a zero-distance pair
    ef = rng.integers(0, n, E)
    et = rng.integers(0, n, E)
    ef[0], et[0] = 5, 6
    nf = np.repeat(ef, 5)
    nt = rng.integers(0, n, 5 * E)
    a, b, alpha = 1.577, 0.895, 0.7
    want = ou.sgd_step(Y, ef, et, nf, nt, alpha, a, b, dtype=np.float64)
    t = lambda v: torch.from_numpy(np.asarray(v, dtype=np.int64))
    got = ou.sgd_step_index_add(torch.from_numpy(Y.astype(np.float32)), t(ef), t(et), t(nf), t(nt), alpha, a, b)
    np.testing.assert_allclose(got.numpy(), want, rtol=1e-5, atol=2e-5)
    # the step must move points (the comparison is not vacuous)
    assert np.abs(want - Y).max() > 0.1


def test_spectral_power_matches_eigsh_on_symmetric_graph():
    from scipy.sparse import coo_matrix
    from scipy.sparse.linalg import eigsh
    rng = np.random.default_rng(4)
    n = 120
    A = (rng.uniform(size=(n, n)) < 0.08).astype(float) * rng.uniform(0.2, 1, (n, n))
    A = np.triu(A, 1); A = A + A.T
    rows, cols = np.nonzero(A)
    vals = A[rows, cols]
    V = ou.spectral_power(rows, cols, vals, n, rng.standard_normal((n, 3)), iters=400)
    deg = A.sum(1); Dm = 1 / np.sqrt(np.maximum(deg, 1e-10))
    M = Dm[:, None] * A * Dm[None, :]
    w, U = np.linalg.eigh(M)
    top = U[:, np.argsort(-np.abs(w))[:1]]
    assert abs(abs(top[:, 0] @ V[:, 0]) - 1) < 1e-6


def test_spectral_scale_hand_worked():
    """umap.py:339-348 on a hand-worked case: column 0 dropped, max |v| -> 10,
    1e-4 noise scale, each column min-maxed to [0, 10]."""
    V = np.array([[9.0, 1.0, -2.0], [-7.0, 3.0, 0.0], [3.0, -4.0, 2.0]], dtype=np.float32)
    out = ou.spectral_scale(V, np.zeros((3, 2)))
    # e = V[:, 1:] * 10 / 4 = [[2.5, -5], [7.5, 0], [-10, 5]]
    np.testing.assert_allclose(out, [[12.5 / 17.5 * 10, 0.0], [10.0, 5.0], [0.0, 10.0]], rtol=1e-6, atol=1e-5)
    # noise N(0,1) enters as 1e-4 * noise: 5000 adds 0.5 -> e col0 = [2.5, 7.5, -9.5]
    noise = np.array([[0.0, 0.0], [0.0, 0.0], [5000.0, 0.0]])
    out = ou.spectral_scale(V, noise)
    np.testing.assert_allclose(out[:, 0], [12.0 / 17.0 * 10, 10.0, 0.0], rtol=1e-6, atol=1e-5)
    # invariant to positive rescaling of V and to column 0 (without noise)
    rng = np.random.default_rng(3)
    W = rng.standard_normal((50, 3)).astype(np.float32)
    a = ou.spectral_scale(W, np.zeros((50, 2)))
    W2 = W * np.float32(4.0)
    W2[:, 0] = rng.standard_normal(50)
    np.testing.assert_allclose(ou.spectral_scale(W2, np.zeros((50, 2))), a, atol=1e-5)
    # a sign flip of a column mirrors it: y -> 10 - y
    W3 = W.copy(); W3[:, 2] *= -1
    np.testing.assert_allclose(ou.spectral_scale(W3, np.zeros((50, 2)))[:, 1], 10.0 - a[:, 1], atol=1e-5)


def test_trustworthiness_matches_sklearn():
    from sklearn.manifold import trustworthiness
    rng = np.random.default_rng(5)
    X = rng.standard_normal((300, 10))
    Y = X[:, :2] + 0.1 * rng.standard_normal((300, 2))
    np.testing.assert_allclose(om.trustworthiness(X, Y, 7), trustworthiness(X, Y, n_neighbors=7), atol=1e-12)


def test_knn_preservation_bounds():
    rng = np.random.default_rng(6)
    X = rng.standard_normal((200, 5)).astype(np.float32)
    assert om.knn_preservation(X, X, 10) == 1.0
    assert om.knn_preservation(X, rng.standard_normal((200, 2)), 10) < 0.3


def test_trust_excess_brute_force_ranks_with_ties():
    """oracle trust_excess against ranks counted pair by pair: r(i, j) = 1 +
    #{l != i : (D(i, l), l) < (D(i, j), j)}, the query itself ranked n, on
    small-integer points with many equal distances."""
    rng = np.random.default_rng(11)
    n, d, k = 40, 3, 6
    X = rng.integers(0, 3, size=(n, d)).astype(np.float64)
    rows = np.array([5, 0, 39, 17, 5, 22])
    lo = rng.integers(0, n, size=(rows.size, k))
    lo[1, 2] = rows[1]                      # the query itself: rank n
    want = np.zeros(rows.size, dtype=np.int64)
    for q, i in enumerate(rows):
        D = [float(((X[i] - X[l]) ** 2).sum()) for l in range(n)]
        for j in lo[q]:
            if j == i:
                r = n
            else:
                r = 1 + sum(1 for l in range(n) if l != i and (D[l], l) < (D[j], j))
            want[q] += max(r - k, 0)
    np.testing.assert_array_equal(om.trust_excess(X, lo, rows), want)
