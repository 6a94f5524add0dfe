"""Edge-gradient family beyond UMAP on the GPU: CNE gradients and Adam loop,
PaCMAP/LocalMAP/t-SNE pair laws, against oracle/cne.py; CNE end to end."""

import numpy as np
import pytest
import torch

from oracle import cne as oc
from oracle import knn as oknn
from oracle import metrics as om

pytestmark = pytest.mark.gpu

LOSSES = {"infonce": 0, "nce": 1, "neg": 2}
LAWS = {"pacmap_nn": 0, "pacmap_mn": 1, "pacmap_fp": 2, "localmap_nn": 3, "tsne_attr": 4}


def _lib():
    from mlx_vis_amd import _capi

    return _capi


def _graph(n=1500, d=10, k=10, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 4, (6, d))
    X = (c[rng.integers(0, 6, n)] + rng.normal(size=(n, d))).astype(np.float32)
    idx, _ = oknn.exact_knn(X, X, k, self_base=0)
    return X, oc.build_edges(idx, n)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("loss", ["infonce", "nce", "neg"])
@pytest.mark.parametrize("dim", [2, 3])
def test_cne_gradient_matches_oracle(loss, dim):
    capi = _lib()
    print(f"case loss={loss} dim={dim}", flush=True)
    X, edges = _graph()
    n, E, m = X.shape[0], edges.shape[0], 5
    rng = np.random.default_rng(1)
    Y = rng.normal(0, 2, (n, dim)).astype(np.float32)
    neg = rng.integers(0, n, (E, m)).astype(np.int32)
    Yd, eid, ejd, negd = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (Y, edges[:, 0], edges[:, 1], neg))
    g = torch.zeros((n, dim), dtype=torch.float32, device="cuda")
    capi.check(capi.lib().mv_cne_grad(Yd.data_ptr(), n, dim, eid.data_ptr(), ejd.data_ptr(), E, negd.data_ptr(), m,
                                      LOSSES[loss], 1.0, 7, 1, 0, g.data_ptr(), None))
    want = oc.cne_grad(Y, edges[:, 0], edges[:, 1], neg, loss)
    got = g.cpu().numpy()
    assert _rel(got, want) < 2e-5, _rel(got, want)
    np.testing.assert_allclose(got, want, rtol=1e-3, atol=1e-5 * np.abs(want).max())


@pytest.mark.parametrize("loss", ["infonce", "nce", "neg"])
def test_cne_minibatch_gradient_matches_oracle(loss):
    """Mini-batch mode (cne.py:337-339): each lane draws its edge uniformly, so
    equal heads land in non-adjacent lanes. The host reproduces the kernel's
    Philox draws (edge: counter (b, iter, 0xFFFF, 0xC0E5), negative s:
    (b, iter, s, 0xC0E5), first word mapped by floor(r * N / 2^32)) and the
    oracle's gradient of that edge list must match, scaled by E / batch."""
    from _philox import philox4x32_10, uniform_below

    capi = _lib()
    print(f"case minibatch loss={loss}", flush=True)
    X, edges = _graph(2000, 10, 10, 8)
    n, E, m, B = X.shape[0], edges.shape[0], 5, 6400
    seed, it = 0x1234ABCD5678, 7
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    bidx = np.arange(B)
    e = uniform_below(philox4x32_10(bidx, it, 0xFFFF, 0xC0E5, k0, k1)[0], E)
    neg = np.stack([uniform_below(philox4x32_10(bidx, it, s, 0xC0E5, k0, k1)[0], n) for s in range(m)], axis=1)
    # a batch this size repeats heads in non-adjacent lanes of one wavefront
    heads = edges[e, 0].reshape(-1, 64)
    assert any(len(set(w)) < 64 for w in heads)
    Y = np.random.default_rng(9).normal(0, 2, (n, 2)).astype(np.float32)
    Yd, eid, ejd = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (Y, edges[:, 0], edges[:, 1]))
    g = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    scale = E / B
    capi.check(capi.lib().mv_cne_grad(Yd.data_ptr(), n, 2, eid.data_ptr(), ejd.data_ptr(), E, None, m, LOSSES[loss],
                                      scale, seed, it, B, g.data_ptr(), None))
    want = scale * oc.cne_grad(Y, edges[e, 0], edges[e, 1], neg, loss)
    got = g.cpu().numpy()
    assert _rel(got, want) < 2e-5, _rel(got, want)


@pytest.mark.parametrize("law", list(LAWS))
def test_pair_laws_match_oracle(law):
    capi = _lib()
    rng = np.random.default_rng(2)
    n, P = 3000, 20000
    Y = rng.normal(0, 5, (n, 2)).astype(np.float32)
    src = rng.integers(0, n, P).astype(np.int32)
    dst = ((src + 1 + rng.integers(0, n - 1, P)) % n).astype(np.int32)
    wp = rng.uniform(0.1, 2.0, P).astype(np.float32) if law == "tsne_attr" else None
    Yd, sd, dd = (torch.from_numpy(v).cuda() for v in (Y, src, dst))
    wpd = torch.from_numpy(wp).cuda() if wp is not None else None
    g = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    capi.check(capi.lib().mv_pair_grad(Yd.data_ptr(), n, 2, sd.data_ptr(), dd.data_ptr(), P, capi.ptr(wpd), 0.8, 1.3,
                                       LAWS[law], g.data_ptr(), None))
    want = oc.pair_grad(Y, src, dst, law, w=0.8, w_pair=wp, scale=1.3)
    assert _rel(g.cpu().numpy(), want) < 2e-5


def test_cne_optimize_step_is_adam_on_the_gradient():
    capi = _lib()
    X, edges = _graph(800, 8, 10, 3)
    n, E, m = X.shape[0], edges.shape[0], 5
    Y0 = np.random.default_rng(4).normal(0, 1, (n, 2)).astype(np.float32)
    Yd = torch.from_numpy(Y0).cuda()
    eid, ejd = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (edges[:, 0], edges[:, 1]))
    g = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    # the gradient of iteration 1 with the same Philox negatives
    capi.check(capi.lib().mv_cne_grad(Yd.data_ptr(), n, 2, eid.data_ptr(), ejd.data_ptr(), E, None, m, 0, 1.0, 99, 1,
                                      0, g.data_ptr(), None))
    gv = g.cpu().numpy().astype(np.float64)
    st = torch.zeros((3, n, 2), dtype=torch.float32, device="cuda")
    capi.check(capi.lib().mv_cne_optimize(Yd.data_ptr(), n, 2, eid.data_ptr(), ejd.data_ptr(), E, m, 0, 1, 2, 1.0, 0,
                                          99, st[0].data_ptr(), st[1].data_ptr(), st[2].data_ptr(), None))
    want, _, _ = oc.adam_cne(Y0.astype(np.float64), gv, 0.0, 0.0, 1, 1.0)
    got = Yd.cpu().numpy()
    # first Adam step is lr * sign(g) up to eps; compare where |g| is not tiny
    big = np.abs(gv) > 1e-4
    np.testing.assert_allclose(got[big], want[big], rtol=0, atol=1e-5)
    assert not st[2].any(), "gradient buffer is zero again"


@pytest.mark.parametrize("loss", ["infonce", "neg"])
def test_cne_embedding_quality_matches_oracle_loop(loss):
    from mlx_vis_amd import CNE

    X, edges = _graph(1200, 20, 15, 5)
    n, E, m = X.shape[0], edges.shape[0], 5
    emb = CNE(n_neighbors=15, loss=loss, n_iter=300, random_state=0, knn_method="brute", pca_dim=None).fit_transform(X)
    assert emb.shape == (n, 2) and np.isfinite(emb).all()
    # oracle loop: same edges and PCA init, NumPy negatives, FP64
    Xc = X - X.mean(axis=0)
    vals, vecs = np.linalg.eigh((Xc.T @ Xc) / (n - 1))
    Y = (Xc @ vecs[:, -2:][:, ::-1]) * 0.01
    rng = np.random.default_rng(0)
    mom, var = np.zeros_like(Y), np.zeros_like(Y)
    for it in range(1, 301):
        neg = rng.integers(0, n, (E, m))
        g = oc.cne_grad(Y, edges[:, 0], edges[:, 1], neg, loss)
        Y, mom, var = oc.adam_cne(Y, g, mom, var, it, 1.0)
    t_gpu = om.trustworthiness(X, emb, 15)
    t_ora = om.trustworthiness(X, Y, 15)
    print(f"trustworthiness gpu={t_gpu:.4f} oracle={t_ora:.4f}")
    assert abs(t_gpu - t_ora) < 0.02
    assert t_gpu > 0.85


def test_cne_minibatch_and_callback():
    from mlx_vis_amd import CNE

    X, _ = _graph(1000, 12, 10, 6)
    seen = []
    emb = CNE(n_neighbors=10, loss="nce", n_iter=20, batch_size=2000, random_state=1, knn_method="brute",
              pca_dim=None).fit_transform(X, epoch_callback=lambda it, Y: seen.append((it, Y.shape)))
    assert np.isfinite(emb).all()
    assert [s[0] for s in seen] == list(range(0, 21))
    with pytest.raises(ValueError):
        CNE(loss="bogus")
