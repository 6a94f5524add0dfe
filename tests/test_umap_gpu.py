"""Parity of the HIP UMAP graph, schedule, SGD step and layout with the oracle."""

import numpy as np
import pytest
import torch

from oracle import knn as oknn
from oracle import metrics as om
from oracle import umap as ou

pytestmark = pytest.mark.gpu


def _graph_inputs(n=3000, d=32, k=15, seed=0):
    rng = np.random.default_rng(seed)
    centers = rng.standard_normal((8, d)) * 3
    X = (centers[rng.integers(0, 8, n)] + rng.standard_normal((n, d))).astype(np.float32)
    idx, d2 = oknn.exact_knn(X, X, k, self_base=0)
    return X, idx.astype(np.int32), np.sqrt(d2).astype(np.float32)


@pytest.mark.parametrize("k", [15, 100, 256])
def test_fuzzy_graph_matches_oracle(k):
    """k above 64 runs the one-wavefront-per-row bisection (sums reduced in a
    different order: sigma within the same 1e-5)."""
    from mlx_vis_amd._umap import fuzzy_graph_device

    X, idx, dist = _graph_inputs(k=k)
    n_epochs = 200
    g = fuzzy_graph_device(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), n_epochs,
                           keep_sigma=True)
    w, sigma, rho = ou.fuzzy_rows(dist, k)
    np.testing.assert_array_equal(g.rho.cpu().numpy(), rho)
    np.testing.assert_allclose(g.sigma.cpu().numpy(), sigma, rtol=1e-5)
    ws = ou.symmetrize(idx, w)
    np.testing.assert_allclose(g.w_sym.cpu().numpy(), ws, rtol=2e-5, atol=1e-7)
    keep = ou.prune(ws, n_epochs)
    gk = g.eps.cpu().numpy() > 0
    # prune decisions may differ only for weights within rounding of the threshold
    thr = ws.max() / n_epochs
    diff = gk != keep
    assert np.all(np.abs(ws[diff] - thr) < 1e-5 * max(thr, 1e-6) + 1e-7)
    assert g.n_edges == int(gk.sum())
    assert abs(g.w_max - ws.max()) <= 1e-6


def test_schedule_bit_exact():
    """epochs_per_next after e epochs equals the FP32 restatement given the same weights."""
    from mlx_vis_amd._umap import fuzzy_graph_device, run_epochs

    X, idx, dist = _graph_inputs(n=1500)
    n_epochs = 60
    g = fuzzy_graph_device(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), n_epochs)
    wsym = g.w_sym.cpu().numpy().reshape(-1)
    kept = g.eps.cpu().numpy().reshape(-1) > 0
    eps_or = ou.epochs_per_sample(wsym[kept], n_epochs)
    np.testing.assert_array_equal(g.eps.cpu().numpy().reshape(-1)[kept], eps_or)
    Y = torch.zeros((1500, 2), dtype=torch.float32, device="cuda")
    dY = torch.zeros_like(Y)
    run_epochs(Y, None, g, 1, n_epochs, n_epochs, 0, 1.0, 1.577, 0.895, 7)
    epn = g.epn.cpu().numpy().reshape(-1)[kept]
    ref = eps_or.copy()
    for epoch in range(n_epochs):
        act = ref <= epoch
        ref[act] = (ref[act] + eps_or[act]).astype(np.float32)
    np.testing.assert_array_equal(epn, ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_single_sgd_step_matches_fp64_oracle(dim):
    """One layout-gradient step on fixed inputs, tolerance 1e-5."""
    from mlx_vis_amd._umap import sgd_step_device

    rng = np.random.default_rng(11)
    n, A = 5000, 4096
    Y = rng.uniform(-10, 10, size=(n, dim)).astype(np.float32)
    ef = rng.integers(0, n, A).astype(np.int32)
    et = rng.integers(0, n, A).astype(np.int32)
    nf = ef[np.arange(5 * A) % A]
    nt = rng.integers(0, n, 5 * A).astype(np.int32)
    Y[et[:10]] = Y[ef[:10]] + 1e-4  # near-coincident pairs hit the 1e-6 clamp
    a, b, alpha = 1.5769434603113077, 0.8950608779109733, 0.75
    got = sgd_step_device(torch.from_numpy(Y).cuda(), *(torch.from_numpy(v).cuda() for v in (ef, et, nf, nt)),
                          alpha, a, b).cpu().numpy()
    want = ou.sgd_step(Y, ef, et, nf, nt, np.float32(alpha), np.float32(a), np.float32(b), dtype=np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("n,k,dim,hubs,masked", [(5000, 15, 2, True, False), (3000, 10, 3, False, True),
                                                   (300_000, 15, 2, True, True), (4000, 15, 6, True, False),
                                                   (2000, 12, 8, False, True), (3000, 100, 2, True, False),
                                                   (2500, 256, 3, False, True)])
def test_epochs_without_negatives_match_oracle(n, k, dim, hubs, masked):
    """With negative_sample_rate 0 a layout epoch is the attraction step of
    the epoch's active edges from one snapshot, so several epochs can be
    checked against the oracle's sgd_step (FP64), schedule included. Covers
    tail nodes with more than 64 in-edges (summed over several 128-entry
    chunks, in fixed point beyond two), nodes without in-edges, pruned slots,
    frozen nodes, and a 300,000 x 15 graph."""
    from mlx_vis_amd import find_ab_params
    from mlx_vis_amd._umap import FuzzyGraph, run_epochs

    rng = np.random.default_rng(n + k)
    rows = np.repeat(np.arange(n), k)
    tails = ((rows + 1 + rng.integers(0, n - 1, n * k)) % n).reshape(n, k)
    if hubs:
        pick = rng.uniform(size=n) < 0.3
        tails[pick, 0] = rng.integers(0, 5, pick.sum())        # five hubs, ~0.06 n in-edges each
    tails[np.arange(n) % 7 == 0] %= max(1, n // 2)             # the upper half gets fewer in-edges
    tails = tails.astype(np.int32)
    eps = rng.uniform(1.0, 4.0, (n, k)).astype(np.float32)
    eps[rng.uniform(size=(n, k)) < 0.1] = -1.0                  # pruned slots
    kept = np.flatnonzero(eps.ravel() > 0)
    order = np.argsort(tails.ravel()[kept], kind="stable")
    in_slots = kept[order].astype(np.int32)
    in_node = tails.ravel()[in_slots].astype(np.int32)
    epn0 = np.where(eps > 0, eps, np.float32(np.inf)).astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g = FuzzyGraph(dev(tails), None, dev(eps), dev(epn0.copy()), len(in_slots), 1.0,
                   in_slots=dev(in_slots), in_node=dev(in_node))
    a, b = find_ab_params(1.0, 0.1)
    Y0 = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
    frozen = (np.arange(n) % 5 == 2) if masked else np.zeros(n, dtype=bool)
    Y = dev(Y0.copy())
    n_epochs, lr, e0, e1 = 200, 0.01, 1, 5
    run_epochs(Y, None, g, e0, e1, n_epochs, 0, lr, a, b, 7,
               movable=dev((~frozen).astype(np.uint8)) if masked else None)
    Yo = Y0.astype(np.float64)
    epn = epn0.ravel().copy()
    none = np.empty(0, dtype=np.int64)
    for ep in range(e0, e1):
        alpha = np.float32(lr * (1.0 - ep / n_epochs))
        act = np.flatnonzero(epn <= np.float32(ep))
        epn[act] = epn[act] + eps.ravel()[act]
        prev = Yo
        Yo = ou.sgd_step(Yo, rows[act], tails.ravel()[act], none, none, alpha, a, b, dtype=np.float64)
        Yo[frozen] = prev[frozen]
    np.testing.assert_array_equal(g.epn.cpu().numpy().ravel(), epn)
    got = Y.cpu().numpy()
    np.testing.assert_array_equal(got[frozen], Y0[frozen])
    # one attraction term is up to 4 alpha ~ 4e-2: a dropped or doubled term
    # shows; the hubs sum ~0.06 n terms in FP32
    hub = np.zeros(n, dtype=bool)
    hub[:5] = hubs
    np.testing.assert_allclose(got[~hub], Yo[~hub], rtol=0, atol=2e-5)
    np.testing.assert_allclose(got[hub], Yo[hub], rtol=1e-4, atol=1e-3)


def test_epochs_repeat_bit_for_bit_with_hubs():
    """Two runs of 30 epochs with negatives from the same init give the same
    positions to the bit, on a graph whose hubs have over 1,000 in-edges
    (their entries span ~20 wavefronts of the tail sums) and whose other
    nodes cross wavefront boundaries."""
    from mlx_vis_amd import find_ab_params
    from mlx_vis_amd._umap import FuzzyGraph, run_epochs

    n, k = 20_000, 15
    rng = np.random.default_rng(5)
    tails = ((np.repeat(np.arange(n), k) + 1 + rng.integers(0, n - 1, n * k)) % n).reshape(n, k)
    pick = rng.uniform(size=n) < 0.3
    tails[pick, 0] = rng.integers(0, 5, pick.sum())
    tails = tails.astype(np.int32)
    eps = rng.uniform(1.0, 3.0, (n, k)).astype(np.float32)
    kept = np.flatnonzero(eps.ravel() > 0)
    in_slots = kept[np.argsort(tails.ravel()[kept], kind="stable")].astype(np.int32)
    in_node = tails.ravel()[in_slots].astype(np.int32)
    assert np.bincount(in_node, minlength=n).max() > 1000
    a, b = find_ab_params(1.0, 0.1)
    Y0 = rng.uniform(-10, 10, (n, 2)).astype(np.float32)
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    outs = []
    for _ in range(2):
        epn0 = eps.copy()
        g = FuzzyGraph(dev(tails), None, dev(eps), dev(epn0), len(in_slots), 1.0,
                       in_slots=dev(in_slots), in_node=dev(in_node))
        Y = dev(Y0.copy())
        run_epochs(Y, None, g, 1, 31, 200, 5, 1.0, a, b, 11)
        outs.append(Y.cpu().numpy())
    assert np.isfinite(outs[0]).all()
    np.testing.assert_array_equal(outs[0], outs[1])


def test_c1_single_step_on_its_graph():
    """The full-shape single-step check: the 70,000 x 784 mixture (seed 1), its
    exact 15-NN graph and schedule, Y ~ U[-10, 10] (seed 0), the first 4,096
    slots (the reference's row-major forward-edge order) active at epoch 5,
    five fixed negatives per edge (heads tiled as umap.py:434-436), one
    _sgd_step (umap.py:359-387) against the FP64 oracle, tolerance 1e-5."""
    from mlx_vis_amd import find_ab_params, knn_brute_device
    from mlx_vis_amd._umap import fuzzy_graph_device, sgd_step_device
    from mlx_vis_amd.datasets import gaussian_mixture_784

    n, k, n_epochs, epoch = 70_000, 15, 500, 5
    X = gaussian_mixture_784(n, seed=1)
    idx, dist = knn_brute_device(torch.from_numpy(X).cuda(), k)
    g = fuzzy_graph_device(idx, dist, n_epochs)
    eps = g.eps.cpu().numpy().ravel()
    active = np.flatnonzero((eps > 0) & (eps <= epoch))[:4096]   # epochs_per_next starts at eps
    assert len(active) == 4096
    ef = (active // k).astype(np.int32)
    et = g.knn_idx.cpu().numpy().ravel()[active].astype(np.int32)
    A = len(active)
    rng = np.random.default_rng(0)
    Y = rng.uniform(-10, 10, (n, 2)).astype(np.float32)
    nf = ef[np.arange(5 * A) % A]
    nt = rng.integers(0, n, 5 * A).astype(np.int32)
    a, b = find_ab_params(1.0, 0.1)
    alpha = 1.0 * (1.0 - epoch / n_epochs)
    got = sgd_step_device(torch.from_numpy(Y).cuda(), *(torch.from_numpy(v).cuda() for v in (ef, et, nf, nt)),
                          alpha, a, b).cpu().numpy()
    want = ou.sgd_step(Y, ef, et, nf, nt, np.float32(alpha), np.float32(a), np.float32(b), dtype=np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)


def test_layout_quality_matches_oracle_run():
    """Same graph, a/b and init: 15-NN preservation and trustworthiness of the
    GPU layout within 0.01 of the oracle's epoch loop."""
    from mlx_vis_amd._umap import FuzzyGraph, fuzzy_graph_device, run_epochs

    n, n_epochs = 3000, 200
    X, idx, dist = _graph_inputs(n=n, d=32, seed=3)
    a, b = ou.find_ab_params(1.0, 0.1)
    Y0 = np.random.RandomState(0).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    g = fuzzy_graph_device(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), n_epochs)
    Y = torch.from_numpy(Y0).cuda()
    run_epochs(Y, None, g, 1, n_epochs, n_epochs, 5, 1.0, a, b, 1234)
    Yg = Y.cpu().numpy()
    rows, cols, vals = ou.fuzzy_simplicial_set(idx, dist, n_epochs)
    Yo = ou.optimize(rows, cols, vals, Y0, a, b, n_epochs, seed=0)
    sample = np.random.default_rng(0).choice(n, 600, replace=False)
    pg = om.knn_preservation(X, Yg, 15)
    po = om.knn_preservation(X, Yo, 15)
    tg = om.trustworthiness(X, Yg, 15, rows=sample)
    to = om.trustworthiness(X, Yo, 15, rows=sample)
    print(f"preservation gpu={pg:.4f} oracle={po:.4f}; trust gpu={tg:.4f} oracle={to:.4f}")
    assert np.isfinite(Yg).all()
    assert abs(pg - po) < 0.01
    assert abs(tg - to) < 0.01


@pytest.mark.parametrize("nc", [3, 6, 9])
def test_spectral_init_subspace_matches_oracle(nc):
    from mlx_vis_amd._umap import fuzzy_graph_device
    from mlx_vis_amd import _capi, _device
    import ctypes

    n = 2000
    X, idx, dist = _graph_inputs(n=n, d=16, seed=5)
    g = fuzzy_graph_device(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), 200)
    V0 = np.random.RandomState(1).standard_normal((n, nc)).astype(np.float32)
    V = torch.from_numpy(V0).cuda()
    lib = _capi.lib()
    nb = ctypes.c_size_t(0)
    _capi.check(lib.mv_umap_spectral_ws_bytes(n, 15, nc, ctypes.byref(nb)))
    ws = _device.workspace(nb.value)
    _capi.check(lib.mv_umap_spectral(g.w_sym.data_ptr(), g.eps.data_ptr(), g.knn_idx.data_ptr(), n, 15, nc, 100,
                                     V.data_ptr(), ws.data_ptr(), ws.numel(), None))
    Vg = V.cpu().numpy().astype(np.float64)
    rows, cols, vals = ou.fuzzy_simplicial_set(idx, dist, 200)
    Vo = ou.spectral_power(rows, cols, vals.astype(np.float64), n, V0, iters=100)
    # orthonormal columns, and the same leading vector
    np.testing.assert_allclose(Vg.T @ Vg, np.eye(nc), atol=1e-4)
    assert abs(abs(Vg[:, 0] @ Vo[:, 0]) - 1.0) < 1e-4
    # the same leading 3-D subspace (the trailing vectors converge slowly)
    s = np.linalg.svd(Vg[:, :3].T @ Vo[:, :3], compute_uv=False)
    print(f"nc={nc} principal cosines of the leading 3 {s}")
    assert s.min() > 0.99


@pytest.mark.parametrize("n,nc", [(1, 2), (7, 3), (100_003, 3), (4099, 5), (5001, 9)])
def test_spectral_embed_matches_oracle(n, nc):
    """mv_umap_spectral_embed against the oracle's umap.py:339-348 scaling, fed
    the same N(0, 1) noise (mv_random_normal with the kernel's seed and count);
    a non-finite V raises the flag."""
    import ctypes

    from mlx_vis_amd import _capi, _device

    lib = _capi.lib()
    rng = np.random.default_rng(n)
    V0 = rng.standard_normal((n, nc)).astype(np.float32) * np.float32(0.01)
    V = torch.from_numpy(V0).cuda()
    Y = torch.empty((n, nc - 1), dtype=torch.float32, device="cuda")
    noise = torch.empty(n * (nc - 1), dtype=torch.float32, device="cuda")
    flag = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    ws = _device.workspace(64)
    seed = 0x1234567 + n
    _capi.check(lib.mv_random_normal(noise.data_ptr(), noise.numel(), seed, None))
    _capi.check(lib.mv_umap_spectral_embed(V.data_ptr(), n, nc, seed, Y.data_ptr(), flag.data_ptr(),
                                           ws.data_ptr(), ws.numel(), None))
    assert int(flag.item()) == 0
    want = ou.spectral_scale(V0, noise.cpu().numpy().reshape(n, nc - 1))
    got = Y.cpu().numpy()
    # the device may contract v * expansion + 1e-4 * noise into one FMA
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
    if n > 1:
        np.testing.assert_allclose(got.min(axis=0), 0.0, atol=1e-6)
        np.testing.assert_allclose(got.max(axis=0), 10.0, atol=1e-5)
    V0[n // 2, nc - 1] = np.nan
    V.copy_(torch.from_numpy(V0))
    _capi.check(lib.mv_umap_spectral_embed(V.data_ptr(), n, nc, seed, Y.data_ptr(), flag.data_ptr(),
                                           ws.data_ptr(), ws.numel(), None))
    assert int(flag.item()) == 1
    small = ctypes.c_size_t(8)
    assert lib.mv_umap_spectral_embed(V.data_ptr(), n, nc, seed, Y.data_ptr(), None, ws.data_ptr(), small, None) != 0


def test_fit_transform_end_to_end():
    from mlx_vis_amd import UMAP

    X, _, _ = _graph_inputs(n=2500, d=24, seed=9)
    calls = []
    um = UMAP(n_epochs=50, random_state=0, knn_method="brute")
    Y = um.fit_transform(X, epoch_callback=lambda e, y: calls.append((e, y.shape)))
    assert Y.shape == (2500, 2) and Y.dtype == np.float32 and np.isfinite(Y).all()
    assert calls[0][0] == 0 and calls[1][0] == 2 and calls[-1][0] == 50 and len(calls) == 50
    Y2 = UMAP(n_epochs=50, random_state=0, knn_method="brute", init="random").fit_transform(X)
    assert np.isfinite(Y2).all()
    # isotropic clusters: neighbourhoods inside a cluster are noise, so check
    # trustworthiness rather than 15-NN overlap
    assert om.trustworthiness(X, Y, 15) > 0.85


@pytest.mark.parametrize("n_neighbors", [100, 200])
def test_fit_transform_with_many_neighbours(n_neighbors):
    """n_neighbors above 64: kNN from the symmetric screen's matrix, the
    wavefront fuzzy-row bisection, spectral init and epochs of 100 / 200 slots
    per row."""
    from mlx_vis_amd import UMAP

    X, _, _ = _graph_inputs(n=3000, d=24, seed=4)
    um = UMAP(n_neighbors=n_neighbors, n_epochs=100, random_state=0, knn_method="brute")
    Y = um.fit_transform(X)
    assert Y.shape == (3000, 2) and np.isfinite(Y).all()
    assert um.knn_[0].shape == (3000, n_neighbors)
    assert om.trustworthiness(X, Y, 15) > 0.85


def test_fit_transform_nndescent_with_100_neighbours():
    """Above 20,000 rows the auto rule (_knn.py:32-33) takes NNDescent; at
    n_neighbors = 100 its kernels hold two of a row's neighbours per lane."""
    from mlx_vis_amd import UMAP, trustworthiness

    rng = np.random.default_rng(21)
    n, d = 24000, 24
    centers = rng.standard_normal((8, d)) * 3
    X = (centers[rng.integers(0, 8, n)] + rng.standard_normal((n, d))).astype(np.float32)
    print(f"case UMAP n={n} d={d} n_neighbors=100 auto kNN", flush=True)
    um = UMAP(n_neighbors=100, n_epochs=100, random_state=0)
    Y = um.fit_transform(X)
    assert Y.shape == (n, 2) and np.isfinite(Y).all()
    assert um.knn_[0].shape == (n, 100)
    rows = np.random.default_rng(1).choice(n, 2000, replace=False)
    t = trustworthiness(X, Y, 15, rows=rows)
    print(f"trustworthiness {t:.4f}")
    assert t > 0.85


def test_random_normal_draws():
    """mv_random_normal (spectral start vectors): standard normal moments,
    reproducible by seed, different across seeds."""
    import torch

    from mlx_vis_amd import _capi

    lib = _capi.lib()
    a = torch.empty(1_000_003, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    for t, seed in ((a, 7), (b, 7), (c, 8)):
        _capi.check(lib.mv_random_normal(t.data_ptr(), t.numel(), seed, None))
    x = a.cpu().numpy().astype(np.float64)
    assert np.isfinite(x).all()
    assert abs(x.mean()) < 5e-3 and abs(x.std() - 1.0) < 5e-3
    assert abs((x ** 4).mean() - 3.0) < 0.05          # normal kurtosis
    assert abs((np.abs(x) < 1.0).mean() - 0.6827) < 3e-3
    assert torch.equal(a, b) and not torch.equal(a, c)


@pytest.mark.parametrize("n_components", [1, 3, 4, 6, 8])
def test_fit_transform_other_dimensions_and_init_array(n_components):
    from mlx_vis_amd import UMAP

    X, _, _ = _graph_inputs(n=1500, d=16, seed=12)
    Y = UMAP(n_components=n_components, n_epochs=60, random_state=1, knn_method="brute").fit_transform(X)
    assert Y.shape == (1500, n_components) and np.isfinite(Y).all()
    init = np.random.default_rng(3).uniform(-5, 5, (1500, n_components)).astype(np.float32)
    Y2 = UMAP(n_components=n_components, n_epochs=60, random_state=1, knn_method="brute",
              init=init).fit_transform(X)
    assert np.isfinite(Y2).all() and not np.allclose(Y2, init)
    with pytest.raises(ValueError):
        UMAP(n_components=n_components, init=init[:10]).fit_transform(X)
    if n_components > 4:
        # spectral vectors beyond five: orthonormal columns of the power iteration
        from mlx_vis_amd._umap import fuzzy_graph_device
        from mlx_vis_amd._spectral import spectral_init

        um = UMAP(n_components=n_components, n_epochs=60, random_state=1, knn_method="brute")
        um.fit_transform(X)
        Ys = spectral_init(um.graph_, 1500, n_components, 1).cpu().numpy()
        assert Ys.shape == (1500, n_components)
        np.testing.assert_allclose(Ys.min(axis=0), 0.0, atol=1e-4)
        np.testing.assert_allclose(Ys.max(axis=0), 10.0, rtol=1e-4)


def test_fit_transform_nndescent_and_precision_screens_agree():
    from mlx_vis_amd import UMAP

    X, _, _ = _graph_inputs(n=3000, d=32, seed=13)
    a = UMAP(n_epochs=30, random_state=2, knn_method="brute")
    a.fit_transform(X)
    b = UMAP(n_epochs=30, random_state=2, knn_method="brute", knn_dtype="bf16x3")
    b.fit_transform(X)
    assert torch.equal(a.knn_[0], b.knn_[0])      # certified: the screen precision does not change the graph
    assert torch.equal(a.knn_[1], b.knn_[1])
    assert torch.equal(a.graph_.w_sym, b.graph_.w_sym)
    assert np.isfinite(b.embedding_).all()
    c = UMAP(n_epochs=30, random_state=2, knn_method="nndescent")
    Yc = c.fit_transform(X)
    assert np.isfinite(Yc).all()


def test_c0_end_to_end_layout_quality_matches_oracle():
    """The small plain-method workload through the public API
    (UMAP(n_epochs=200, init="random", knn_method="brute", random_state=0) on
    10,000 x 128 N(0,1), seed 0) against the oracle's layout of the same
    graph, a/b and init (tests/golden/umap_c0_oracle.json, written by
    make_golden.py umap_c0): the kept-edge count up to threshold rounding,
    15-NN preservation and trustworthiness within 0.01 (negatives differ:
    Philox here, NumPy there)."""
    import json
    from pathlib import Path

    from mlx_vis_amd import UMAP

    fx = json.loads((Path(__file__).resolve().parent / "golden" / "umap_c0_oracle.json").read_text())
    X = np.random.default_rng(0).standard_normal((10000, 128), dtype=np.float32)
    um = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=200, negative_sample_rate=5, init="random",
              knn_method="brute", random_state=0)
    Y = um.fit_transform(X)
    assert Y.shape == (10000, 2) and Y.dtype == np.float32 and np.isfinite(Y).all()
    # prune decisions may differ only for weights within FP32 rounding of the threshold
    assert abs(um.graph_.n_edges - fx["n_edges"]) <= max(2, fx["n_edges"] // 10000)
    sample = np.sort(np.random.default_rng(fx["trust_rows_seed"]).choice(10000, fx["trust_rows"], replace=False))
    p = om.knn_preservation(X, Y, 15)
    t = om.trustworthiness(X, Y, 15, rows=sample)
    print(f"preservation gpu={p:.4f} oracle={fx['knn15_preservation']:.4f}; "
          f"trust gpu={t:.4f} oracle={fx['trustworthiness_15']:.4f}")
    assert abs(p - fx["knn15_preservation"]) < 0.01
    assert abs(t - fx["trustworthiness_15"]) < 0.01


@pytest.mark.parametrize("init", ["random", "spectral"])
def test_c1_end_to_end_layout_quality_matches_oracle(init):
    """The full-shape workload through the public API: UMAP(n_epochs=500,
    knn_method="brute", random_state=0) with random or spectral (the API
    default) init on the 70,000 x 784 mixture (seed 1), against the oracle's
    layouts of the same graph, a/b and inits (tests/golden/umap_c1_oracle.json,
    written by make_golden.py umap_c1): the kept-edge count up to threshold
    rounding, and 15-NN preservation and trustworthiness over the 2,000 rows of
    knn_c1_sample2000.npz within 0.01 (negatives and spectral start vectors
    differ: Philox here, NumPy there)."""
    import json
    from pathlib import Path

    from mlx_vis_amd import UMAP
    from mlx_vis_amd.datasets import gaussian_mixture_784

    golden = Path(__file__).resolve().parent / "golden"
    fx = json.loads((golden / "umap_c1_oracle.json").read_text())
    sample = np.load(golden / "knn_c1_sample2000.npz")["rows"].astype(np.int64)
    X = gaussian_mixture_784(70_000, seed=1)
    um = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=500, negative_sample_rate=5, init=init,
              knn_method="brute", random_state=0)
    Y = um.fit_transform(X)
    assert Y.shape == (70_000, 2) and Y.dtype == np.float32 and np.isfinite(Y).all()
    assert abs(um.graph_.n_edges - fx["n_edges"]) <= max(2, fx["n_edges"] // 10000)
    p = om.knn_preservation(X, Y, 15, rows=sample)
    t = om.trustworthiness(X, Y, 15, rows=sample)
    want = fx[init]
    print(f"{init}: preservation gpu={p:.4f} oracle={want['knn15_preservation']:.4f}; "
          f"trust gpu={t:.4f} oracle={want['trustworthiness_15']:.4f}")
    assert abs(p - want["knn15_preservation"]) < 0.01
    assert abs(t - want["trustworthiness_15"]) < 0.01


def test_umap_normalize_minmax_is_applied_before_the_knn():
    """UMAP(normalize="minmax") gives the embedding of the min-max scaled input
    (mlx_vis/_normalize.py:28-31 applied in fit_transform, umap.py:86-88)."""
    from mlx_vis_amd import UMAP

    rng = np.random.default_rng(4)
    X = (rng.standard_normal((3000, 20)) * rng.uniform(0.5, 30, 20) + 5).astype(np.float32)
    kw = dict(n_neighbors=10, n_epochs=60, init="random", knn_method="brute", random_state=3)
    ua, ub = UMAP(normalize="minmax", **kw), UMAP(**kw)
    Ya = ua.fit_transform(X)
    Xs = ((X - X.min(0)) / (X.max(0) - X.min(0) + 1e-8)).astype(np.float32)
    Yb = ub.fit_transform(Xs)
    # the same graph (the kNN of the scaled input, bit for bit) ...
    assert torch.equal(ua.knn_[0], ub.knn_[0]) and torch.equal(ua.knn_[1], ub.knn_[1])
    oi, _ = oknn.exact_knn(Xs, Xs, 10, self_base=0)
    np.testing.assert_array_equal(ua.knn_[0].cpu().numpy(), oi)
    # ... and the same layout, bit for bit: the same graph, init and Philox
    # negatives, and tail sums formed in an order that does not depend on
    # scheduling (this input has nodes with up to 99 in-edges, whose entries
    # span three wavefronts)
    assert Ya.shape == Yb.shape == (3000, 2) and np.isfinite(Ya).all()
    assert np.bincount(ua.knn_[0].cpu().numpy().ravel(), minlength=3000).max() > 64
    np.testing.assert_array_equal(Ya, Yb)
