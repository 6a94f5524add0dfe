"""PaCMAP / LocalMAP host logic and oracle (CPU): pair counts, the three
pair samplers of the package against the oracle's row-by-row restatement of
the reference (same seed, same pairs), and the oracle's force laws against
finite differences of the weighted pair losses."""

import numpy as np
import pytest

from oracle import cne as oc
from oracle import knn as oknn
from oracle import pacmap as op


def _pac():
    import mlx_vis_amd._pacmap as pm

    return pm


@pytest.mark.parametrize("n,nn,mn,fp", [(10, 10, 0.5, 2.0), (40, 10, 0.5, 2.0), (30000, None, 0.5, 2.0),
                                        (5, 3, 1.0, 1.0), (100000, None, 0.5, 2.0), (2, 10, 0.5, 2.0)])
def test_decide_num_pairs_matches_oracle(n, nn, mn, fp):
    assert _pac().decide_num_pairs(n, nn, mn, fp) == op.decide_num_pairs(n, nn, mn, fp)


def test_decide_num_pairs_hand_worked():
    # n = 40: 10 + 5 + 20 = 35 < 40, unchanged; n = 30: 10 + 5 + 19 = 34 >= 30 -> 8, 4, 16
    assert op.decide_num_pairs(40, 10, 0.5, 2.0) == (10, 5, 20)
    assert op.decide_num_pairs(30, 10, 0.5, 2.0) == (8, 4, 16)
    # n_neighbors=None: 10 up to 10,000 points, then 10 + 15 (log10 n - 4)
    assert op.decide_num_pairs(100000, None, 0.5, 2.0)[0] == 25


def test_sample_neighbors_matches_oracle_and_hand_example():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(300, 12)).astype(np.float32)
    idx, d2 = oknn.exact_knn(X, X, 25, self_base=0)
    dist = np.sqrt(d2).astype(np.float32)
    got = _pac().sample_neighbors(dist, idx, 10)
    np.testing.assert_array_equal(got, op.sample_neighbors(dist, idx, 10))
    # hand example: sigma_i = mean of columns 3..5
    d = np.array([[1, 2, 3, 4, 5, 6, 7], [1, 1, 1, 2, 2, 2, 9]], dtype=np.float32)
    ix = np.array([[1, 1, 1, 1, 1, 1, 1], [0, 0, 0, 0, 0, 0, 0]], dtype=np.int32)
    # sigma = (5, 2); row 0 scaled = d^2 / 10: ascending as d
    assert list(op.sample_neighbors(d, ix, 3)[:3, 1]) == [1, 1, 1]


def test_sample_mn_pairs_matches_oracle():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(500, 16)).astype(np.float32)
    got = _pac().sample_mn_pairs(X, 3, np.random.default_rng(7))
    want = op.sample_mn_pairs(X, 3, np.random.default_rng(7))
    np.testing.assert_array_equal(got, want)
    # every pick is the second closest of its six candidates (brute force)
    off = np.random.default_rng(7).integers(1, 500, size=(500, 6))
    cand = (np.arange(500)[:, None] + off) % 500
    d = ((X[cand].astype(np.float64) - X[:, None, :]) ** 2).sum(-1)
    second = cand[np.arange(500), np.argsort(d, axis=1, kind="stable")[:, 1]]
    np.testing.assert_array_equal(got[:500, 1], second)


@pytest.mark.parametrize("n,n_nb,n_fp,chunk", [(300, 10, 20, 4096), (300, 10, 20, 64), (50, 5, 10, 7),
                                               (31, 10, 20, 4096)])
def test_sample_fp_pairs_same_stream_as_reference_loop(n, n_nb, n_fp, chunk):
    pm = _pac()
    rng0 = np.random.default_rng(3)
    nb = np.stack([np.repeat(np.arange(n), n_nb), rng0.integers(0, n, n * n_nb)], axis=1).astype(np.int32)
    n_fp = min(n_fp, max(0, n - 1 - n_nb))
    got = pm.sample_fp_pairs(n, nb, n_nb, n_fp, np.random.default_rng(11), chunk=chunk)
    want = op.sample_fp_pairs(n, nb, n_nb, n_fp, np.random.default_rng(11))
    np.testing.assert_array_equal(got, want)
    # the exclusion rules hold: no self, no neighbour, no repeat within a row
    rows = got[:, 1].reshape(n, n_fp)
    nbr = nb[:, 1].reshape(n, n_nb)
    for i in range(n):
        assert i not in rows[i]
        assert not set(rows[i]) & set(nbr[i])
        assert len(set(rows[i])) == n_fp


def test_sample_fp_pairs_fallback_rows_match_reference_loop():
    pm = _pac()
    n, n_nb, n_fp = 200, 10, 20
    rng0 = np.random.default_rng(5)
    nbr = rng0.integers(0, n, (n, n_nb)).astype(np.int32)
    nb = np.stack([np.repeat(np.arange(n), n_nb), nbr.ravel()], axis=1).astype(np.int32)
    got = pm._fp_rows_one_by_one(range(n), n, nbr, n_fp, np.random.default_rng(9))
    want = op.sample_fp_pairs(n, nb, n_nb, n_fp, np.random.default_rng(9))[:, 1].reshape(n, n_fp)
    np.testing.assert_array_equal(got, want)


def test_resample_local_fp_pairs_rules():
    pm = _pac()
    rng = np.random.default_rng(2)
    n, n_nb, n_fp = 400, 5, 10
    Y = rng.normal(0, 3, (n, 2)).astype(np.float32)
    nb = np.stack([np.repeat(np.arange(n), n_nb), rng.integers(0, n, n * n_nb)], axis=1).astype(np.int32)
    fp = np.stack([np.repeat(np.arange(n), n_fp), rng.integers(0, n, n * n_fp)], axis=1).astype(np.int32)
    out = pm.resample_local_fp_pairs(nb, fp, Y, 2.0, np.random.default_rng(4))
    nbr = nb[:, 1].reshape(n, n_nb)
    changed = out[:, 1] != fp[:, 1]
    src = out[changed, 0]
    dst = out[changed, 1]
    assert (src != dst).all()
    assert (((Y[src] - Y[dst]) ** 2).sum(1) <= 4.0 + 1e-6).all()
    for s, t in zip(src, dst):
        assert t not in nbr[s]


@pytest.mark.parametrize("local", [False, True])
def test_oracle_step_gradient_is_the_weighted_loss_gradient(local):
    rng = np.random.default_rng(6)
    n = 12
    Y = rng.normal(0, 2, (n, 2))
    pairs = [np.stack([rng.integers(0, n, 15), rng.integers(0, n, 15)], 1) for _ in range(3)]
    pairs = [p[p[:, 0] != p[:, 1]] for p in pairs]
    nn, mn, fp = pairs
    w = (2.0, 300.0, 1.0)

    def loss(Z):
        Z = np.asarray(Z)
        tot = oc.pair_loss(Z, fp[:, 0], fp[:, 1], "pacmap_fp", w=w[2])
        tot += oc.pair_loss(Z, mn[:, 0], mn[:, 1], "pacmap_mn", w=w[1])
        if local:
            return tot, None   # LocalMAP's neighbour law: pinned by the line integral below
        return tot + oc.pair_loss(Z, nn[:, 0], nn[:, 1], "pacmap_nn", w=w[0]), None

    h = 1e-6
    g_fd = np.zeros_like(Y)
    for i in range(n):
        for q in range(2):
            Yp, Ym = Y.copy(), Y.copy()
            Yp[i, q] += h
            Ym[i, q] -= h
            g_fd[i, q] = (loss(Yp)[0] - loss(Ym)[0]) / (2 * h)
    if local:
        g_fd = g_fd + oc.pair_grad(Y, nn[:, 0], nn[:, 1], "localmap_nn", w=w[0], scale=5.0)
    # one step from m = v = 0 with lr_t = 1: m = 0.1 g, v = 0.001 g^2
    Y1, m1, v1 = op.step(Y, np.zeros_like(Y), np.zeros_like(Y), nn, mn, fp, *w, 1.0,
                         local_scale=5.0 if local else None)
    np.testing.assert_allclose(m1 / 0.1, g_fd, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(v1, 0.001 * g_fd ** 2, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(Y1, Y - m1 / (np.sqrt(v1) + 1e-7), rtol=1e-12)


def test_localmap_law_is_the_gradient_of_its_loss():
    # LocalMAP's neighbour law 20 w s / ((10 + d)^2 sqrt(d)) (y_s - y_t): its
    # work along a straight path equals the difference of the closed-form
    # potential f below (a field with a potential is path independent)
    rng = np.random.default_rng(8)
    Y0 = rng.normal(0, 2, (2, 2))
    Y1 = rng.normal(0, 2, (2, 2))
    src, dst = np.array([0]), np.array([1])
    ts = np.linspace(0, 1, 20001)
    work = 0.0
    for a, b in zip(ts[:-1], ts[1:]):
        t = 0.5 * (a + b)
        Yt = Y0 + t * (Y1 - Y0)
        work += (oc.pair_grad(Yt, src, dst, "localmap_nn", w=1.5, scale=2.0) * (Y1 - Y0)).sum() * (b - a)

    def f(Z):
        df = Z[0] - Z[1]
        d = df @ df + 1.0
        r = np.sqrt(d)
        # antiderivative in d of 10 s w / ((10 + d)^2 sqrt(d)) (the law is 2 (y_s - y_t) times it)
        return 1.5 * 2.0 * 10.0 * (r / (10.0 * (10.0 + d)) + np.arctan(r / np.sqrt(10.0)) / (10.0 * np.sqrt(10.0)))

    assert abs(work - (f(Y1) - f(Y0))) < 1e-6 * max(1.0, abs(work))
