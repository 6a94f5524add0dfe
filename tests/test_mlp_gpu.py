"""Learned form on the GPU: forward, gradients and Adam against the CPU oracle
(oracle/mlp.py), edge sampling, the preservation check, masked refinement, and
what the trained network gains."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _lib():
    from mlx_vis_amd import _capi

    return _capi


def _blob(dims, params):
    capi = _lib()
    nb, npar = ctypes.c_size_t(0), ctypes.c_int64(0)
    capi.check(capi.lib().mv_mlp_param_bytes(*dims, ctypes.byref(nb), ctypes.byref(npar)))
    blob = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    t = {k: torch.from_numpy(np.ascontiguousarray(params[k], dtype=np.float32)).cuda() for k in om.PARAM_NAMES}
    capi.check(capi.lib().mv_mlp_init(blob.data_ptr(), *dims, t["W1"].data_ptr(), t["b1"].data_ptr(),
                                      t["W2"].data_ptr(), t["b2"].data_ptr(), t["W3"].data_ptr(),
                                      t["b3"].data_ptr(), None))
    return blob, int(npar.value)


def _params(dims, seed):
    din, h1, h2, dout = dims
    p = om.init_params(din, h1, h2, dout, seed=seed)
    rng = np.random.default_rng(seed + 1)
    for k in ("b1", "b2"):
        p[k] = (p[k] + rng.normal(0, 0.2, p[k].shape)).astype(np.float32)
    p["W3"] = (p["W3"] * 8).astype(np.float32)   # outputs of embedding scale
    return p


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("dims,m", [((100, 128, 192, 2), 700), ((784, 512, 512, 2), 1000), ((33, 64, 64, 3), 130)])
def test_forward_matches_oracle(dims, m):
    capi = _lib()
    print(f"forward dims={dims} m={m}", flush=True)
    p = _params(dims, 0)
    blob, _ = _blob(dims, p)
    X = np.random.default_rng(1).uniform(0, 1, (m, dims[0])).astype(np.float32)
    nb = ctypes.c_size_t(0)
    chunk = 256
    capi.check(capi.lib().mv_mlp_batch_bytes(*dims, chunk, ctypes.byref(nb)))
    ws = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    Z = torch.empty((m, dims[3]), dtype=torch.float32, device="cuda")
    Xd = torch.from_numpy(X).cuda()   # kept alive: a freed temporary's memory may be reused at once
    capi.check(capi.lib().mv_mlp_forward(blob.data_ptr(), *dims, Xd.data_ptr(), m, Z.data_ptr(), chunk,
                                         ws.data_ptr(), ws.numel(), None))
    z = Z.cpu().numpy()
    zo, _ = om.forward(p, X, bf16=True)
    z64, _ = om.forward(p, X, bf16=False)
    # Same BF16 rounding points as the oracle: only FP32-vs-FP64 accumulation
    # differs (it can flip a BF16 rounding, 2^-8 relative, of single elements).
    assert _rel(z, zo) < 2e-3, _rel(z, zo)
    # Against exact FP64 arithmetic the BF16 storage dominates.
    assert _rel(z, z64) < 3e-2, _rel(z, z64)


def test_forward_with_large_preactivations_matches_oracle():
    """Weights scaled so that pre-activations reach |x| >= 16 (GELU returns x
    or -0 exactly there) and beyond the GELU table's range on both sides."""
    capi = _lib()
    dims, m = (100, 128, 192, 2), 3000
    p = _params(dims, 2)
    p["W1"] = (p["W1"] * 40).astype(np.float32)
    p["W2"] = (p["W2"] * 6).astype(np.float32)
    blob, _ = _blob(dims, p)
    X = np.random.default_rng(9).uniform(0, 1, (m, dims[0])).astype(np.float32)
    a1 = X @ p["W1"].T + p["b1"]
    assert (np.abs(a1) >= 16).mean() > 0.05 and (a1 <= -16).any() and (a1 >= 16).any()
    Xd = torch.from_numpy(X).cuda()
    z = _forward(blob, dims, Xd, m, 1024)
    zo, _ = om.forward(p, X, bf16=True)
    assert np.isfinite(z).all()
    assert _rel(z, zo) < 2e-3, _rel(z, zo)


def _forward(blob, dims, Xd, m, chunk):
    capi = _lib()
    nb = ctypes.c_size_t(0)
    capi.check(capi.lib().mv_mlp_batch_bytes(*dims, chunk, ctypes.byref(nb)))
    ws = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    Z = torch.empty((m, dims[3]), dtype=torch.float32, device="cuda")
    capi.check(capi.lib().mv_mlp_forward(blob.data_ptr(), *dims, Xd.data_ptr(), m, Z.data_ptr(), chunk,
                                         ws.data_ptr(), ws.numel(), None))
    return Z.cpu().numpy()


@pytest.mark.parametrize("dims,m", [((784, 512, 512, 2), 70001), ((100, 192, 320, 1), 40000)])
def test_large_chunk_forward_matches_small_chunks_and_oracle(dims, m):
    """One chunk of every row (a persistent grid over many tiles per
    workgroup; FP32 input staged as BF16 for din % 8 == 0, a BF16 copy
    otherwise; one output or two) against chunks of 16,384 rows on every row,
    and against the oracle on 2,000 sampled rows (rows past a multiple of the
    tile, hidden widths that are not multiples of it)."""
    print(f"large-chunk forward dims={dims} m={m}", flush=True)
    p = _params(dims, 3)
    blob, _ = _blob(dims, p)
    X = np.random.default_rng(4).uniform(0, 1, (m, dims[0])).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    z_big = _forward(blob, dims, Xd, m, m)
    z_small = _forward(blob, dims, Xd, m, 16384)
    assert np.isfinite(z_big).all()
    # only the order of the output layer's atomic sums differs
    assert _rel(z_big, z_small) < 1e-5, _rel(z_big, z_small)
    rows = np.sort(np.random.default_rng(5).choice(m, 2000, replace=False))
    rows[-1] = m - 1
    zo, _ = om.forward(p, X[rows], bf16=True)
    assert _rel(z_big[rows], zo) < 2e-3, _rel(z_big[rows], zo)


@pytest.mark.parametrize("dims,B,neg", [((100, 128, 192, 2), 128, 5), ((784, 512, 512, 2), 4096, 5),
                                        ((33, 64, 128, 2), 64, 0)])
def test_batch_gradient_matches_oracle(dims, B, neg):
    capi = _lib()
    print(f"grad dims={dims} B={B} neg={neg}", flush=True)
    from mlx_vis_amd import find_ab_params

    a, b = find_ab_params(1.0, 0.1)
    p = _params(dims, 2)
    blob, npar = _blob(dims, p)
    n = 3 * B
    rng = np.random.default_rng(3)
    X = rng.uniform(0, 1, (n, dims[0])).astype(np.float32)
    Y = rng.normal(0, 5, (n, dims[3])).astype(np.float32)
    idx = rng.integers(0, n, 2 * B).astype(np.int32)
    lam = 0.5
    dp = -(-dims[0] // 64) * 64
    nb = ctypes.c_size_t(0)
    capi.check(capi.lib().mv_mlp_batch_bytes(*dims, 2 * B, ctypes.byref(nb)))
    ws = torch.empty(nb.value + (-(-2 * n * dp // 256) * 256), dtype=torch.uint8, device="cuda")
    grad = torch.empty(npar, dtype=torch.float32, device="cuda")
    loss = torch.empty(2, dtype=torch.float32, device="cuda")
    Xd, Yd, idxd = (torch.from_numpy(v).cuda() for v in (X, Y, idx))
    capi.check(capi.lib().mv_mlp_grad(blob.data_ptr(), *dims, Xd.data_ptr(), n, Yd.data_ptr(), idxd.data_ptr(),
                                      B, neg, a, b, lam, grad.data_ptr(), loss.data_ptr(), ws.data_ptr(), ws.numel(),
                                      None))
    g = om.unflatten(grad.cpu().numpy(), *dims, dp=dp)
    go, (ce, reg), _, _ = om.batch_grad(p, X, Y, idx, B, neg, a, b, lam, bf16=True)
    lg = loss.cpu().numpy()
    assert abs(lg[0] - ce) <= 1e-3 * abs(ce) + 1e-5 and abs(lg[1] - reg) <= 1e-3 * abs(reg) + 1e-5, (lg, ce, reg)
    for k in om.PARAM_NAMES:
        r = _rel(g[k], go[k])
        print(f"  {k}: rel {r:.2e}")
        assert r < 1e-2, (k, r)
    # padded W1 columns carry no gradient
    flat = grad.cpu().numpy()
    W1p = flat[: dims[1] * dp].reshape(dims[1], dp)
    assert not W1p[:, dims[0]:].any()


def test_train_steps_are_adam_on_the_batch_gradient():
    capi = _lib()
    from mlx_vis_amd import find_ab_params

    dims, B, neg, lam, lr = (50, 64, 64, 2), 64, 3, 0.2, 1e-3
    a, b = find_ab_params(1.0, 0.1)
    p = _params(dims, 4)
    blob, npar = _blob(dims, p)
    n = 500
    rng = np.random.default_rng(5)
    Xd = torch.from_numpy(rng.uniform(0, 1, (n, dims[0])).astype(np.float32)).cuda()
    Yd = torch.from_numpy(rng.normal(0, 3, (n, dims[3])).astype(np.float32)).cuda()
    dp = 64
    nb = ctypes.c_size_t(0)
    capi.check(capi.lib().mv_mlp_batch_bytes(*dims, 2 * B, ctypes.byref(nb)))
    ws = torch.empty(nb.value + 2 * n * dp + 256, dtype=torch.uint8, device="cuda")
    grad = torch.empty(npar, dtype=torch.float32, device="cuda")
    master = blob[: 4 * npar].view(torch.float32)
    pv = master.cpu().numpy().astype(np.float64)
    m = np.zeros(npar)
    v = np.zeros(npar)
    for step in range(1, 4):
        idx = torch.from_numpy(rng.integers(0, n, 2 * B).astype(np.int32)).cuda()
        capi.check(capi.lib().mv_mlp_grad(blob.data_ptr(), *dims, Xd.data_ptr(), n, Yd.data_ptr(), idx.data_ptr(),
                                          B, neg, a, b, lam, grad.data_ptr(), None, ws.data_ptr(), ws.numel(), None))
        gv = grad.cpu().numpy().astype(np.float64)
        capi.check(capi.lib().mv_mlp_train(blob.data_ptr(), *dims, Xd.data_ptr(), n, Yd.data_ptr(), idx.data_ptr(),
                                           1, B, neg, a, b, lam, lr, step - 1, None, ws.data_ptr(), ws.numel(), None))
        pv, m, v = om.adam_step(pv, gv, m, v, step, lr=lr)
        got = master.cpu().numpy().astype(np.float64)
        diff = np.abs(got - pv)
        # FP32 state versus FP64; the gradients of two launches can differ in
        # the last bits (atomic summation order), which matters only where |g|
        # is near Adam's eps.
        assert (diff > 1e-6).mean() < 1e-3, (step, (diff > 1e-6).mean())
        assert diff.max() <= 2.5 * lr
        # BF16 copies follow the master weights
        pv = got
    torch.cuda.synchronize()


def _graph(n=2000, d=16, k=10, seed=0):
    from mlx_vis_amd import knn_brute_device
    from mlx_vis_amd._umap import fuzzy_graph_device

    X = np.random.default_rng(seed).normal(size=(n, d)).astype(np.float32)
    idx, dist = knn_brute_device(torch.from_numpy(X).cuda(), k)
    return fuzzy_graph_device(idx, dist, 200)


def test_edge_sampling_follows_weights():
    capi = _lib()
    g = _graph()
    n, k = g.n, g.k
    steps, B = 300, 1024
    out = torch.empty(steps * 2 * B, dtype=torch.int32, device="cuda")
    capi.check(capi.lib().mv_mlp_sample_edges(g.eps.data_ptr(), g.knn_idx.data_ptr(), n, k, 1234, 0, steps, B,
                                              out.data_ptr(), None))
    o = out.cpu().numpy().reshape(steps, 2, B)
    heads, tails = o[:, 0].ravel(), o[:, 1].ravel()
    knn = g.knn_idx.cpu().numpy()
    eps = g.eps.cpu().numpy().ravel()
    pos = (knn[heads] == tails[:, None])
    assert pos.any(axis=1).all(), "every draw is a graph edge"
    slot = heads * k + pos.argmax(axis=1)
    counts = np.bincount(slot, minlength=n * k)
    kept = eps > 0
    assert counts[~kept].sum() == 0, "pruned slots are never drawn"
    p = np.where(kept, 1.0 / np.where(kept, eps, 1.0), 0.0)
    expect = p / p.sum() * heads.size
    # Pearson chi-square over kept slots, normalised by its degrees of freedom
    chi = ((counts[kept] - expect[kept]) ** 2 / expect[kept]).sum() / (kept.sum() - 1)
    assert 0.85 < chi < 1.15, chi
    # different step offsets give different batches; the same offset repeats
    out2 = torch.empty(2 * B, dtype=torch.int32, device="cuda")
    capi.check(capi.lib().mv_mlp_sample_edges(g.eps.data_ptr(), g.knn_idx.data_ptr(), n, k, 1234, 5, 1, B,
                                              out2.data_ptr(), None))
    np.testing.assert_array_equal(out2.cpu().numpy(), o[5].ravel())


def test_knn_overlap_counts_and_flags():
    from mlx_vis_amd import knn_preservation_device

    rng = np.random.default_rng(7)
    n, k = 5000, 15
    a = np.stack([rng.permutation(40)[:k] for _ in range(n)]).astype(np.int32)
    b = np.stack([rng.permutation(40)[:k] for _ in range(n)]).astype(np.int32)
    b[:100] = a[:100][:, ::-1]
    counts, flags = knn_preservation_device(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), 0.8)
    want = np.array([len(set(x) & set(y)) for x, y in zip(a, b)])
    np.testing.assert_array_equal(counts.cpu().numpy(), want)
    np.testing.assert_array_equal(flags.cpu().numpy(), (want < 12).astype(np.uint8))
    assert (want[:100] == k).all()


def test_masked_epochs_move_only_flagged_points():
    from mlx_vis_amd import find_ab_params
    from mlx_vis_amd._umap import run_epochs

    g = _graph(3000, 8, 15, 1)
    a, b = find_ab_params(1.0, 0.1)
    Y0 = torch.from_numpy(np.random.default_rng(2).uniform(-10, 10, (g.n, 2)).astype(np.float32)).cuda()
    mask = torch.from_numpy((np.arange(g.n) % 3 == 0).astype(np.uint8)).cuda()
    epn0 = g.epn.clone()
    Ym = Y0.clone()
    g.epn.copy_(epn0)
    run_epochs(Ym, None, g, 1, 30, 200, 5, 1.0, a, b, 99, movable=mask)
    Yu = Y0.clone()
    g.epn.copy_(epn0)
    run_epochs(Yu, None, g, 1, 30, 200, 5, 1.0, a, b, 99)
    ones = torch.ones(g.n, dtype=torch.uint8, device="cuda")
    Y1 = Y0.clone()
    g.epn.copy_(epn0)
    run_epochs(Y1, None, g, 1, 30, 200, 5, 1.0, a, b, 99, movable=ones)
    m = mask.cpu().numpy().astype(bool)
    ym, y0 = Ym.cpu().numpy(), Y0.cpu().numpy()
    np.testing.assert_array_equal(ym[~m], y0[~m])
    assert (np.abs(ym[m] - y0[m]).sum(axis=1) > 0).mean() > 0.95
    np.testing.assert_allclose(Y1.cpu().numpy(), Yu.cpu().numpy(), rtol=0, atol=1e-4)


def _roll(n, seed, d=64):
    """Swiss roll (t cos t, h, t sin t) lifted into d dimensions by a fixed
    orthonormal map, plus N(0, 0.05^2) noise: data with local structure that a
    2-D embedding can keep."""
    rng = np.random.default_rng(seed)
    t = 1.5 * np.pi * (1 + 2 * rng.uniform(size=n))
    h = 20 * rng.uniform(size=n)
    P = np.stack([t * np.cos(t), h, t * np.sin(t)], axis=1)
    Q, _ = np.linalg.qr(np.random.default_rng(1000).normal(size=(d, 3)))
    return (P @ Q.T + rng.normal(0, 0.05, (n, d))).astype(np.float32)


def _pres(Xd, Z, k=15):
    from mlx_vis_amd import knn_brute_device, knn_preservation_device

    hi, _ = knn_brute_device(Xd, k)
    lo, _ = knn_brute_device(Z.contiguous(), k)
    c, _ = knn_preservation_device(hi, lo)
    return c.double().mean().item() / k, c


def test_trained_network_gains_over_untrained_and_refinement_gains():
    from mlx_vis_amd import UMAP, ParametricUMAP

    Xtr = torch.from_numpy(_roll(6000, 0)).cuda()
    Xnew = torch.from_numpy(_roll(3000, 1)).cuda()
    um = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=200, random_state=0, normalize=False, knn_method="brute")
    um.fit_transform(Xtr)
    pu = ParametricUMAP(hidden=(256, 256), batch_size=1024, random_state=0, edge_passes=80)
    pu.init_weights(Xtr.shape[1])
    z_untrained = pu.transform_device(Xnew)
    pu.fit(Xtr, umap=um)
    tr = pu.loss_trace_.cpu().numpy()
    assert np.isfinite(tr).all()
    assert tr[-20:].sum(axis=1).mean() < 0.5 * tr[:20].sum(axis=1).mean(), "training lowers the loss"
    z_new = pu.transform_device(Xnew)
    p_untrained, _ = _pres(Xnew, z_untrained)
    p_mlp, _ = _pres(Xnew, z_new)
    plain = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=200, random_state=0, normalize=False, knn_method="brute")
    z_plain = torch.from_numpy(plain.fit_transform(Xnew)).cuda()
    p_plain, _ = _pres(Xnew, z_plain)
    z_ref, before = pu.refine_device(Xnew, z_new, threshold=0.8, n_epochs=50)
    p_ref, after = _pres(Xnew, z_ref)
    flagged = (before < 12).cpu().numpy()
    print(f"preservation untrained={p_untrained:.3f} mlp={p_mlp:.3f} plain={p_plain:.3f} refined={p_ref:.3f} "
          f"flagged={flagged.mean():.3f}")
    assert p_mlp > p_untrained + 0.2
    assert p_mlp >= p_plain - 0.05
    b_np, a_np = before.cpu().numpy(), after.cpu().numpy()
    assert a_np[flagged].mean() > b_np[flagged].mean(), "refinement raises the flagged points' preservation"
    assert p_ref >= p_mlp - 0.01
