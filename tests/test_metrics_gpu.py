"""Embedding-quality metrics on the GPU (trustworthiness ranks, k-NN
preservation) against oracle/metrics.py.

Exactness: on small-integer and dyadic inputs every squared distance is exact
in FP32 and in the oracle's FP64, so the per-row rank sums must agree
bit-for-bit, ties included (both order equal distances by index). On general
floats near-equal distances may order differently in FP32 and FP64; there
the statistic is compared with an absolute tolerance of 1e-4.
"""

import numpy as np
import pytest
import torch

from oracle import metrics as om

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


@pytest.mark.parametrize("n,d,k,with_rows", [(1500, 7, 10, False), (1500, 37, 15, True), (130, 16, 64, True),
                                             (64, 1, 1, False)])
def test_trust_excess_exact_on_integer_points(n, d, k, with_rows):
    """Per-row rank excess bit-exact against the oracle; arbitrary neighbour
    lists (not kNN), a query equal to its own neighbour, ragged tiles."""
    from mlx_vis_amd import trust_excess_device

    rng = np.random.default_rng(n + d + k)
    X = rng.integers(0, 4, size=(n, d)).astype(np.float32)    # many equal distances
    if with_rows:
        rows = rng.integers(0, n, size=257).astype(np.int32)   # unsorted, with repeats
    else:
        rows = None
    m = n if rows is None else rows.size
    lo = rng.integers(0, n, size=(m, k)).astype(np.int32)
    qi = np.arange(m) if rows is None else rows
    lo[3, 0] = qi[3]                                           # the query itself: rank n
    print(f"trust_excess n={n} d={d} k={k} m={m}", flush=True)
    got = trust_excess_device(_dev(X), _dev(lo, torch.int32),
                              None if rows is None else _dev(rows, torch.int32)).cpu().numpy()
    want = om.trust_excess(X, lo, rows)
    np.testing.assert_array_equal(got, want)


def test_trustworthiness_dyadic_matches_oracle():
    """The full statistic, embedding neighbours from the GPU brute kNN, on
    dyadic data (exact arithmetic, ties in both spaces)."""
    import mlx_vis_amd as mv

    rng = np.random.default_rng(4)
    n, d = 2100, 50
    X = (rng.integers(-64, 64, size=(n, d)) / 16.0).astype(np.float32)
    Y = (X[:, :2] + rng.integers(-8, 8, size=(n, 2)) / 64.0).astype(np.float32)
    print("trustworthiness dyadic 2100 x 50", flush=True)
    t_gpu = mv.trustworthiness(X, Y, 15)
    t_ora = om.trustworthiness(X, Y, 15)
    assert abs(t_gpu - t_ora) <= 1e-12, (t_gpu, t_ora)
    rows = rng.permutation(n)[:300]
    t_gpu_r = mv.trustworthiness(X, Y, 15, rows=rows)
    t_ora_r = om.trustworthiness(X, Y, 15, rows=rows)
    assert abs(t_gpu_r - t_ora_r) <= 1e-12, (t_gpu_r, t_ora_r)


def test_trustworthiness_and_preservation_float_data():
    import mlx_vis_amd as mv

    rng = np.random.default_rng(7)
    n = 3000
    c = rng.normal(0, 3, (5, 40))
    X = (c[rng.integers(0, 5, n)] + rng.normal(size=(n, 40))).astype(np.float32)
    Y = (X[:, :2] + 0.3 * rng.normal(size=(n, 2))).astype(np.float32)
    print("trustworthiness float 3000 x 40", flush=True)
    t_gpu = mv.trustworthiness(X, Y, 15)
    t_ora = om.trustworthiness(X, Y, 15)
    assert abs(t_gpu - t_ora) <= 1e-4, (t_gpu, t_ora)
    p_gpu = mv.knn_preservation(X, Y, 15)
    p_ora = om.knn_preservation(X, Y, 15)
    assert abs(p_gpu - p_ora) <= 1e-3, (p_gpu, p_ora)
    # an embedding equal to the data keeps every neighbour: both statistics are
    # 1 (a near-tie at the 15th neighbour ordered differently by the FP32 ranks
    # than by the certified kNN costs 2 / (n k (2n - 3k - 1)) = 7.5e-9 per pair)
    assert mv.trustworthiness(X, X, 15) >= 1.0 - 1e-6
    assert mv.knn_preservation(X, X, 15) == 1.0


def test_trustworthiness_mixture_70000x784_row_sample():
    """The 70,000 x 784 mixture (the fit_transform workload) on 256 sampled
    rows, against the oracle's FP64 ranks on the same rows."""
    import mlx_vis_amd as mv
    from mlx_vis_amd.datasets import gaussian_mixture_784

    X = gaussian_mixture_784()
    rng = np.random.default_rng(9)
    Y = (X[:, :2] * 10 + rng.normal(size=(X.shape[0], 2))).astype(np.float32)
    rows = rng.permutation(X.shape[0])[:256]
    print("trustworthiness 70000 x 784, 256 rows", flush=True)
    t_gpu = mv.trustworthiness(X, Y, 15, rows=rows)
    t_ora = om.trustworthiness(X, Y, 15, rows=rows)
    assert abs(t_gpu - t_ora) <= 1e-4, (t_gpu, t_ora)


def test_metric_argument_errors():
    import mlx_vis_amd as mv

    X = np.zeros((50, 3), np.float32)
    with pytest.raises(ValueError):
        mv.trustworthiness(X, X[:, :2], 25)            # k >= n / 2
    with pytest.raises(ValueError):
        mv.trustworthiness(X, X[:, :2], 5, rows=[0, 50])
    with pytest.raises(ValueError):
        mv.trustworthiness(X, X[:40, :2], 5)
    with pytest.raises(ValueError):
        mv.trust_excess_device(_dev(X), _dev(np.zeros((5, 65), np.int32), torch.int32))
