"""CNE (contrastive neighbour embedding) on MI355X: drop-in for mlx_vis.CNE
(mlx_vis/_cne/cne.py:17-372), reusing the layout's edge-gradient machinery.

Same constructor (n_components=2, n_neighbors=15, n_negatives=5,
loss="infonce"|"nce"|"neg", n_iter=500, learning_rate=1.0, batch_size=None,
pca_dim=50, random_state=None, verbose=False, knn_method="auto",
normalize=False) and fit_transform(X, epoch_callback=None).

Host steps follow the reference: normalisation, PCA to pca_dim (covariance
eigendecomposition on the CPU, cne.py:155-173), the shared compute_knn
operator, the symmetrised unique edge list (cne.py:192-217) and the PCA
initialisation scaled by 0.01 (cne.py:175-190). The optimisation
(cne.py:315-372: per-iteration negatives, optional edge mini-batches, the
InfoNCE / NCE / NEG gradient and Adam) runs in libmivis (mv_cne_optimize).
The reference draws negatives and batches from MLX's generator (cne.py:338,
346); here they come from Philox-4x32-10 keyed by random_state, so runs are
reproducible but not the same draws.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi, _device
from ._knn import compute_knn
from ._normalize import normalize_input

LOSSES = {"infonce": 0, "nce": 1, "neg": 2}


def _pca_basis(X: np.ndarray, n_keep: int):
    """Top n_keep principal directions of X (covariance eigendecomposition,
    descending), and the centred data."""
    Xc = X - X.mean(axis=0)
    cov = (Xc.T @ Xc) / (X.shape[0] - 1)
    vals, vecs = np.linalg.eigh(cov)
    return Xc, vecs[:, -n_keep:][:, ::-1], vals


def build_edges(knn_indices: np.ndarray, n: int) -> np.ndarray:
    """Both directions of every kNN edge, deduplicated, without self-loops,
    sorted by (head, tail) — cne.py:192-217."""
    k = knn_indices.shape[1]
    src = np.repeat(np.arange(n, dtype=np.int32), k)
    dst = knn_indices.ravel().astype(np.int32)
    all_src = np.concatenate([src, dst])
    all_dst = np.concatenate([dst, src])
    keys = np.unique(all_src.astype(np.int64) * n + all_dst.astype(np.int64))
    edges = np.stack([(keys // n).astype(np.int32), (keys % n).astype(np.int32)], axis=1)
    return edges[edges[:, 0] != edges[:, 1]]


class CNE:
    """Contrastive Neighbor Embedding on an MI355X GPU."""

    def __init__(self, n_components=2, n_neighbors=15, n_negatives=5, loss="infonce", n_iter=500,
                 learning_rate=1.0, batch_size=None, pca_dim=50, random_state=None, verbose=False,
                 knn_method: str = "auto", normalize: str | bool = False):
        if loss not in LOSSES:
            raise ValueError(f"Unknown loss: {loss!r}. Use 'infonce', 'nce', or 'neg'.")
        self.n_components = n_components
        self.n_neighbors = n_neighbors
        self.n_negatives = n_negatives
        self.loss = loss
        self.n_iter = n_iter
        self.learning_rate = learning_rate
        self.batch_size = batch_size
        self.pca_dim = pca_dim
        self.random_state = random_state
        self.verbose = verbose
        self.knn_method = knn_method
        self.normalize = normalize
        self.embedding_ = None

    def fit_transform(self, X, epoch_callback=None) -> np.ndarray:
        if isinstance(X, torch.Tensor):
            X = X.detach().cpu().numpy()
        X = np.asarray(X, dtype=np.float32)
        X = normalize_input(X, self.normalize)
        n, dim = X.shape
        if not 1 <= self.n_components <= 4:
            raise ValueError("n_components must be in [1, 4]")
        if self.pca_dim is not None and dim > self.pca_dim:
            Xc, proj, vals = _pca_basis(X, self.pca_dim)
            X = np.ascontiguousarray(Xc @ proj, dtype=np.float32)
            if self.verbose:
                print(f"Variance retained: {vals[-self.pca_dim:].sum() / vals.sum() * 100:.1f}%")
        knn_idx, _ = compute_knn(X, self.n_neighbors, method=self.knn_method, verbose=self.verbose,
                                 random_state=self.random_state if self.random_state is not None else 42)
        edges = build_edges(np.asarray(knn_idx), n)
        n_edges = edges.shape[0]
        Xc, proj, _ = _pca_basis(X, self.n_components)
        Y0 = np.ascontiguousarray((Xc @ proj) * 0.01, dtype=np.float32)

        dev = _device.device()
        Y = torch.from_numpy(Y0).to(dev)
        ei = torch.from_numpy(np.ascontiguousarray(edges[:, 0])).to(dev)
        ej = torch.from_numpy(np.ascontiguousarray(edges[:, 1])).to(dev)
        state = torch.zeros((3, n, self.n_components), dtype=torch.float32, device=dev)   # Adam m, v; gradient
        batch = self.batch_size if self.batch_size is not None and self.batch_size < n_edges else 0
        seed = self.random_state if self.random_state is not None else int(np.random.randint(0, 2**31 - 1))
        seed = (int(seed) * 0x9E3779B97F4A7C15 + 0x3C6EF372FE94F82B) & 0xFFFFFFFFFFFFFFFF
        lib = _capi.lib()

        def run(it0, it1):
            _capi.check(lib.mv_cne_optimize(
                Y.data_ptr(), n, self.n_components, ei.data_ptr(), ej.data_ptr(), n_edges, int(self.n_negatives),
                LOSSES[self.loss], int(it0), int(it1), float(self.learning_rate), int(batch), seed,
                state[0].data_ptr(), state[1].data_ptr(), state[2].data_ptr(), _device.stream()), "mv_cne_optimize")

        if epoch_callback is not None:
            epoch_callback(0, Y.cpu().numpy())
            for it in range(1, self.n_iter + 1):
                run(it, it + 1)
                epoch_callback(it, Y.cpu().numpy())
        else:
            run(1, self.n_iter + 1)
        self.embedding_ = Y.cpu().numpy()
        return self.embedding_
