"""Pins the CNE / pair-law oracle (oracle/cne.py) by finite differences and a
brute-force edge set."""

import numpy as np
import pytest

from oracle import cne as oc


def _fd(f, Y, h=1e-6):
    g = np.zeros_like(Y)
    for i in range(Y.shape[0]):
        for q in range(Y.shape[1]):
            Yp, Ym = Y.copy(), Y.copy()
            Yp[i, q] += h
            Ym[i, q] -= h
            g[i, q] = (f(Yp) - f(Ym)) / (2 * h)
    return g


@pytest.mark.parametrize("loss", ["infonce", "nce", "neg"])
def test_cne_gradient_is_the_loss_gradient(loss):
    rng = np.random.default_rng(0)
    n, E, m = 12, 20, 4
    Y = rng.normal(0, 1.5, (n, 2))
    ei = rng.integers(0, n, E)
    ej = (ei + 1 + rng.integers(0, n - 1, E)) % n
    neg = rng.integers(0, n, (E, m))
    neg = np.where(neg == ei[:, None], (neg + 1) % n, neg)   # a self-negative has zero gradient but log(0) loss
    g = oc.cne_grad(Y, ei, ej, neg, loss)
    fd = _fd(lambda Z: oc.cne_loss(Z, ei, ej, neg, loss), Y)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("law", ["pacmap_nn", "pacmap_mn", "pacmap_fp"])
def test_pacmap_laws_are_loss_gradients(law):
    rng = np.random.default_rng(1)
    n, P = 10, 25
    Y = rng.normal(0, 3, (n, 2))
    src = rng.integers(0, n, P)
    dst = (src + 1 + rng.integers(0, n - 1, P)) % n
    g = oc.pair_grad(Y, src, dst, law, w=0.7)
    fd = _fd(lambda Z: oc.pair_loss(Z, src, dst, law, w=0.7), Y)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-9)


def test_tsne_attractive_term_moves_sources_only():
    Y = np.array([[0.0, 0.0], [1.0, 2.0], [3.0, -1.0]])
    src, dst = np.array([0, 1]), np.array([1, 2])
    g = oc.pair_grad(Y, src, dst, "tsne_attr", w_pair=np.array([0.5, 2.0]))
    np.testing.assert_allclose(g[0], 4 * 0.5 * (Y[0] - Y[1]) / (1 + 5.0))
    np.testing.assert_allclose(g[1], 4 * 2.0 * (Y[1] - Y[2]) / (1 + 13.0))
    np.testing.assert_allclose(g[2], 0.0)


def test_build_edges_is_the_symmetric_edge_set():
    rng = np.random.default_rng(2)
    n, k = 40, 5
    knn = np.stack([rng.choice(n, k, replace=False) for _ in range(n)]).astype(np.int32)
    knn[3, 0] = 3   # a self-loop is dropped
    e = oc.build_edges(knn, n)
    want = set()
    for i in range(n):
        for j in knn[i]:
            if i != j:
                want.add((i, int(j)))
                want.add((int(j), i))
    assert set(map(tuple, e.tolist())) == want
    assert len(e) == len(want)
    keys = e[:, 0].astype(np.int64) * n + e[:, 1]
    assert (np.diff(keys) > 0).all(), "sorted by (head, tail)"


def test_adam_cne_matches_torch_adam():
    """oracle adam_cne (cne.py:355-360) against torch.optim.Adam on the same
    gradient sequence (bias-corrected, eps outside the square root)."""
    import torch

    rng = np.random.default_rng(21)
    Y0 = rng.standard_normal((30, 2))
    grads = [rng.standard_normal((30, 2)) for _ in range(6)]
    p = torch.nn.Parameter(torch.tensor(Y0, dtype=torch.float64))
    opt = torch.optim.Adam([p], lr=0.05, betas=(0.9, 0.999), eps=1e-8)
    Y, m, v = Y0.copy(), 0.0, 0.0
    for it, g in enumerate(grads, start=1):
        opt.zero_grad()
        p.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        Y, m, v = oc.adam_cne(Y, g, m, v, it, 0.05)
    np.testing.assert_allclose(Y, p.detach().numpy(), rtol=0, atol=1e-12)
