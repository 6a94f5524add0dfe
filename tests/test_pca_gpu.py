"""Device PCA (mv_pca_cov, mv_pca_project; the array-sized part of
mlx_vis/_umap/umap.py:92-113) against NumPy float64."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d", [(5000, 300), (70_001, 100), (257, 5), (3, 70)])
def test_covariance_and_mean_match_float64(n, d):
    import ctypes

    from mlx_vis_amd import _capi, _device

    print(f"case pca cov n={n} d={d}", flush=True)
    rng = np.random.default_rng(n + d)
    X = (rng.standard_normal((n, d)) * rng.uniform(0.1, 5, d) + rng.uniform(-3, 3, d)).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    lib = _capi.lib()
    nb = ctypes.c_size_t(0)
    _capi.check(lib.mv_pca_ws_bytes(n, d, ctypes.byref(nb)))
    ws = _device.workspace(nb.value)
    mean = torch.empty(d, dtype=torch.float64, device="cuda")
    cov = torch.empty((d, d), dtype=torch.float64, device="cuda")
    _capi.check(lib.mv_pca_cov(Xd.data_ptr(), n, d, mean.data_ptr(), cov.data_ptr(), ws.data_ptr(), ws.numel(), None))
    X64 = X.astype(np.float64)
    m64 = X64.mean(axis=0)
    c64 = (X64 - m64).T @ (X64 - m64) / max(n - 1, 1)
    np.testing.assert_allclose(mean.cpu().numpy(), m64, rtol=1e-12, atol=1e-12)
    scale = np.sqrt(np.outer(np.diag(c64), np.diag(c64))) + 1e-30
    err = np.abs(cov.cpu().numpy() - c64) / scale
    assert err.max() < 2e-5, err.max()
    assert np.array_equal(cov.cpu().numpy(), cov.cpu().numpy().T)


def test_projection_and_umap_pca_dim():
    from mlx_vis_amd import UMAP, knn_brute_device
    from mlx_vis_amd._pca import pca_device

    rng = np.random.default_rng(7)
    n, d, p = 6000, 120, 16
    basis = rng.standard_normal((p, d))
    X = (rng.standard_normal((n, p)) * np.linspace(6, 1, p)) @ basis + 0.05 * rng.standard_normal((n, d))
    X = X.astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    Z = pca_device(Xd, p).cpu().numpy()
    X64 = X.astype(np.float64)
    Xc = X64 - X64.mean(axis=0)
    vals, vecs = np.linalg.eigh(Xc.T @ Xc / (n - 1))
    Zref = Xc @ vecs[:, -p:][:, ::-1]
    # components agree up to the sign of each eigenvector
    sign = np.sign((Z * Zref).sum(axis=0))
    np.testing.assert_allclose(Z * sign, Zref, rtol=0, atol=2e-3 * np.abs(Zref).max())
    # UMAP(pca_dim) runs its kNN on the device projection
    um = UMAP(pca_dim=p, n_epochs=30, random_state=0, init="random", knn_method="brute")
    Y = um.fit_transform(X)
    assert Y.shape == (n, 2) and np.isfinite(Y).all()
    ki, _ = knn_brute_device(pca_device(Xd, p), 15)
    assert torch.equal(um.knn_[0], ki)
