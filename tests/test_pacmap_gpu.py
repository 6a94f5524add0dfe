"""PaCMAP / LocalMAP on the GPU against oracle/pacmap.py: the device
iterations (pair laws + Adam) on fixed pairs, and end-to-end embeddings."""

import numpy as np
import pytest

from oracle import metrics as om
from oracle import pacmap as op

pytestmark = pytest.mark.gpu


def _data(n=1500, d=20, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 5, (6, d))
    return (c[rng.integers(0, 6, n)] + rng.normal(size=(n, d))).astype(np.float32)


@pytest.mark.parametrize("local", [False, True])
def test_device_iterations_match_oracle(local):
    from mlx_vis_amd import LocalMAP, PaCMAP

    rng = np.random.default_rng(1)
    n = 2000
    Y0 = rng.normal(0, 3, (n, 2)).astype(np.float32)

    def pairs(m):
        s = np.repeat(np.arange(n), m)
        t = (s + 1 + rng.integers(0, n - 1, n * m)) % n
        return np.stack([s, t], 1).astype(np.int32)

    nn, mn, fp = pairs(10), pairs(5), pairs(20)
    phases = (3, 2, 3)
    est = LocalMAP(num_iters=phases, low_dist_thres=6.0) if local else PaCMAP(num_iters=phases)
    # no resampling inside these iterations (it needs itr > p1 + p2 and itr % 10 == 0)
    got = est._optimize(Y0, nn, mn, fp, np.random.default_rng(0), low_dist_thres=6.0 if local else None)
    Y, m, v = Y0.astype(np.float64), np.zeros((n, 2)), np.zeros((n, 2))
    for itr in range(sum(phases)):
        w_nb, w_mn, w_fp = op.weights(itr, phases)
        Y, m, v = op.step(Y, m, v, nn, mn, fp, w_nb, w_mn, w_fp, op.lr_t(1.0, itr),
                          local_scale=3.0 if (local and w_mn is None) else None)
    # FP32 device against FP64 oracle: each Adam step moves a coordinate by
    # about lr_t ~ 1 with FP32 rounding of |Y| ~ 10; the normalised update
    # m / sqrt(v) magnifies rounding where the gradient is small
    err = np.abs(got - Y) / (1.0 + np.abs(Y))
    print(f"max rel err {err.max():.3g}, 99.9% {np.quantile(err, 0.999):.3g}, median {np.median(err):.3g}")
    assert np.median(err) < 1e-5
    assert np.quantile(err, 0.999) < 3e-4
    assert err.max() < 2e-3


def test_pacmap_end_to_end_matches_oracle_loop():
    import mlx_vis_amd._pacmap as pm
    from mlx_vis_amd import PaCMAP

    X = _data()
    est = PaCMAP(random_state=3, num_iters=(100, 100, 250), knn_method="brute")
    emb = est.fit_transform(X)
    assert emb.shape == (1500, 2) and np.isfinite(emb).all()
    # oracle loop on the same pairs and the same initialisation
    nn, mn, fp = est.pairs_
    Xp = PaCMAP()._preprocess(X.copy())
    Y = est._init(Xp, "pca", None).astype(np.float64)
    m = np.zeros_like(Y)
    v = np.zeros_like(Y)
    for itr in range(450):
        w_nb, w_mn, w_fp = op.weights(itr, (100, 100, 250))
        Y, m, v = op.step(Y, m, v, nn, mn, fp, w_nb, w_mn, w_fp, op.lr_t(1.0, itr))
    t_gpu = om.trustworthiness(X, emb, 10)
    t_ora = om.trustworthiness(X, Y, 10)
    print(f"trustworthiness gpu={t_gpu:.4f} oracle={t_ora:.4f}")
    assert abs(t_gpu - t_ora) < 0.02
    assert t_gpu > 0.85
    # the samplers drew the reference's pairs for this seed
    rng = np.random.default_rng(3)
    np.testing.assert_array_equal(mn, pm.sample_mn_pairs(Xp, mn.shape[0] // 1500, rng))
    np.testing.assert_array_equal(fp, op.sample_fp_pairs(1500, nn, 10, fp.shape[0] // 1500, rng))


def test_localmap_and_callback():
    from mlx_vis_amd import LocalMAP

    X = _data(800, 12, 4)
    seen = []
    emb = LocalMAP(random_state=0, num_iters=(20, 20, 40), knn_method="brute").fit_transform(
        X, epoch_callback=lambda it, Y: seen.append(it))
    assert np.isfinite(emb).all()
    assert seen == list(range(0, 81))
    assert om.trustworthiness(X, emb, 10) > 0.8


def test_pacmap_high_dimensional_input_uses_pca():
    from mlx_vis_amd import PaCMAP

    X = _data(1200, 150, 5)
    est = PaCMAP(random_state=1, num_iters=(30, 30, 60))
    emb = est.fit_transform(X, init="random")
    assert est._pca_solution and emb.shape == (1200, 2) and np.isfinite(emb).all()
    with pytest.raises(ValueError):
        PaCMAP().fit_transform(X, init="bogus")


def test_pacmap_more_neighbours_than_64_candidates():
    """n_neighbors = 20: the reference asks compute_knn for k = n_neighbors + 50
    = 70 (pacmap.py:529), above the 64 of the candidate lists; the self kNN
    then selects from the symmetric screen's matrix."""
    from mlx_vis_amd import PaCMAP

    X = _data(1500, 20, 6)
    emb = PaCMAP(n_neighbors=20, random_state=2, num_iters=(40, 40, 80), knn_method="brute").fit_transform(X)
    assert emb.shape == (1500, 2) and np.isfinite(emb).all()
    assert om.trustworthiness(X, emb, 10) > 0.8
