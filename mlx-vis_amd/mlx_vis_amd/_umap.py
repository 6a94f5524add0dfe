"""UMAP on MI355X: drop-in for ``mlx_vis.UMAP`` (mlx_vis/_umap/umap.py:25-449).

Same constructor arguments and defaults (umap.py:46-60), same
``fit_transform(X, epoch_callback=None)`` contract (umap.py:75-144): float32
output of shape (n, n_components), ``embedding_`` set, ``n_epochs`` defaulted
(500 if n <= 10,000 else 200) and stored on the instance (umap.py:123-124),
``epoch_callback(epoch, Y)`` called at epoch 0 and after every epoch with an
active edge (umap.py:420-442).

Additions (documented extensions, defaults keep the reference behaviour):
``init`` ("spectral" as the reference, "random" for uniform [-10, 10], or an
(n, n_components) array) and ``knn_dtype`` (brute-force screen precision).

Every array-sized step runs in libmivis kernels: kNN, the fuzzy simplicial
set with prune and schedule, spectral initialisation, and the SGD epochs.
Host work is O(1) per epoch plus the 300-point a/b curve fit (umap.py:272-296,
host NumPy in the reference too).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi, _device
from ._normalize import normalize_input


def find_ab_params(spread: float, min_dist: float) -> tuple[float, float]:
    """Fit 1 / (1 + a x^(2b)) to the UMAP membership curve by Gauss-Newton.

    Same procedure as UMAP._find_ab_params (umap.py:272-296): 300 points on
    [0, 3*spread], target 1 below min_dist and exp(-(x - min_dist)/spread)
    above, start (1, 1), at most 100 least-squares steps, stop when the squared
    step is below 1e-12.
    """
    xv = np.linspace(0, spread * 3, 300)
    yv = np.where(xv < min_dist, 1.0, np.exp(-(xv - min_dist) / spread))
    a, b = 1.0, 1.0
    for _ in range(100):
        x2b = np.power(xv, 2 * b)
        den = 1.0 + a * x2b
        resid = 1.0 / den - yv
        ja = -x2b / (den * den)
        jb = -a * 2.0 * np.log(np.maximum(xv, 1e-20)) * x2b / (den * den)
        step = np.linalg.lstsq(np.column_stack([ja, jb]), -resid, rcond=None)[0]
        a += step[0]
        b += step[1]
        if np.sum(step ** 2) < 1e-12:
            break
    return float(a), float(b)


class FuzzyGraph:
    """Device-resident UMAP graph over the dense (n, k) forward-edge slots."""

    def __init__(self, knn_idx: torch.Tensor, w_sym, eps, epn, n_edges: int, w_max: float,
                 sigma=None, rho=None, in_slots=None, in_node=None):
        self.knn_idx = knn_idx
        self.w_sym = w_sym
        self.eps = eps
        self.epn = epn
        self.n_edges = n_edges
        self.w_max = w_max
        self.sigma = sigma
        self.rho = rho
        self.in_slots = in_slots
        self.in_node = in_node

    @property
    def n(self):
        return self.knn_idx.shape[0]

    @property
    def k(self):
        return self.knn_idx.shape[1]


def fuzzy_graph_device(knn_idx: torch.Tensor, knn_dist: torch.Tensor, n_epochs: int,
                       keep_sigma: bool = False) -> FuzzyGraph:
    """Fuzzy simplicial set + prune + schedule on the device (umap.py:199-270, 399-404)."""
    _device.require_device("fuzzy_graph_device", knn_idx, knn_dist)
    n, k = knn_idx.shape
    dev = knn_idx.device
    knn_idx = knn_idx.to(torch.int32).contiguous()
    knn_dist = knn_dist.to(torch.float32).contiguous()
    w_sym = torch.empty((n, k), dtype=torch.float32, device=dev)
    eps = torch.empty((n, k), dtype=torch.float32, device=dev)
    epn = torch.empty((n, k), dtype=torch.float32, device=dev)
    counters = torch.empty(2, dtype=torch.int32, device=dev)
    sigma = torch.empty(n, dtype=torch.float32, device=dev) if keep_sigma else None
    rho = torch.empty(n, dtype=torch.float32, device=dev) if keep_sigma else None
    lib = _capi.lib()
    nbytes = ctypes.c_size_t(0)
    _capi.check(lib.mv_umap_graph_ws_bytes(n, k, ctypes.byref(nbytes)), "mv_umap_graph_ws_bytes")
    ws = _device.workspace(nbytes.value)
    _capi.check(lib.mv_umap_fuzzy_graph(
        knn_idx.data_ptr(), knn_dist.data_ptr(), n, k, int(n_epochs), w_sym.data_ptr(), eps.data_ptr(),
        epn.data_ptr(), counters.data_ptr(), _capi.ptr(sigma), _capi.ptr(rho), ws.data_ptr(), ws.numel(),
        _device.stream()), "mv_umap_fuzzy_graph")
    in_slots = torch.empty(n * k, dtype=torch.int32, device=dev)
    in_node = torch.empty(n * k, dtype=torch.int32, device=dev)
    nb = -(-n // 1024)
    ws2 = _device.workspace(3 * (-(-4 * n // 256) * 256) + (-(-4 * nb // 256) * 256) + 256)
    _capi.check(lib.mv_umap_in_edges(knn_idx.data_ptr(), eps.data_ptr(), n, k, in_slots.data_ptr(),
                                     in_node.data_ptr(), ws2.data_ptr(), ws2.numel(), _device.stream()),
                "mv_umap_in_edges")
    c = counters.cpu()
    n_edges = int(c[0].item()) & 0xFFFFFFFF
    w_max = float(np.array([int(c[1].item()) & 0xFFFFFFFF], dtype=np.uint32).view(np.float32)[0])
    return FuzzyGraph(knn_idx, w_sym, eps, epn, n_edges, w_max, sigma, rho, in_slots, in_node)


def epochs_workspace(n: int, k: int, dim: int) -> torch.Tensor:
    nbytes = ctypes.c_size_t(0)
    _capi.check(_capi.lib().mv_umap_epochs_ws_bytes(n, k, dim, ctypes.byref(nbytes)), "mv_umap_epochs_ws_bytes")
    return _device.workspace(nbytes.value)


def run_epochs(Y: torch.Tensor, ws: torch.Tensor | None, g: FuzzyGraph, e0: int, e1: int, n_epochs: int,
               neg_rate: int, lr: float, a: float, b: float, seed: int,
               movable: torch.Tensor | None = None) -> None:
    """Runs layout epochs [e0, e1) in place on Y (ws from epochs_workspace, or None).

    movable: optional (n,) uint8 device mask; nodes with 0 keep their position.
    """
    n, dim = Y.shape
    if ws is None:
        ws = epochs_workspace(n, g.k, dim)
    args = (Y.data_ptr(), n, dim, g.knn_idx.data_ptr(), g.k, g.eps.data_ptr(), g.epn.data_ptr(),
            g.in_slots.data_ptr(), g.in_node.data_ptr(), int(g.n_edges), int(e0), int(e1), int(n_epochs),
            int(neg_rate), float(lr), float(a), float(b), int(seed) & 0xFFFFFFFFFFFFFFFF)
    lib = _capi.lib()
    if movable is None:
        _capi.check(lib.mv_umap_epochs(*args, ws.data_ptr(), ws.numel(), _device.stream()), "mv_umap_epochs")
    else:
        if movable.dtype != torch.uint8 or movable.numel() != n or not movable.is_cuda:
            raise ValueError("movable must be a (n,) uint8 device tensor")
        _capi.check(lib.mv_umap_epochs_masked(*args, movable.contiguous().data_ptr(), ws.data_ptr(), ws.numel(),
                                              _device.stream()), "mv_umap_epochs_masked")


def sgd_step_device(Y: torch.Tensor, ef, et, neg_from, neg_to, alpha: float, a: float, b: float) -> torch.Tensor:
    """One exact _sgd_step (umap.py:359-387) with explicit lists; returns the new Y."""
    _device.require_device("sgd_step_device", Y)
    n, dim = Y.shape
    Y = Y.float().contiguous()
    out = torch.empty_like(Y)
    # index lists may come from the host: the kernel must only ever see device pointers
    ef = ef.to(device=Y.device, dtype=torch.int32).contiguous()
    et = et.to(device=Y.device, dtype=torch.int32).contiguous()
    neg_from = neg_from.to(device=Y.device, dtype=torch.int32).contiguous()
    neg_to = neg_to.to(device=Y.device, dtype=torch.int32).contiguous()
    _capi.check(_capi.lib().mv_umap_sgd_step(
        out.data_ptr(), Y.data_ptr(), n, dim, ef.data_ptr(), et.data_ptr(), ef.numel(),
        neg_from.data_ptr(), neg_to.data_ptr(), neg_from.numel(), float(alpha), float(a), float(b),
        _device.stream()), "mv_umap_sgd_step")
    return out


class UMAP:
    """UMAP dimensionality reduction on an MI355X GPU.

    Parameters follow mlx_vis.UMAP (umap.py:28-44); see the module docstring
    for the two extensions ``init`` and ``knn_dtype``.
    """

    def __init__(
        self,
        n_components: int = 2,
        n_neighbors: int = 15,
        min_dist: float = 0.1,
        spread: float = 1.0,
        n_epochs: int | None = None,
        learning_rate: float = 1.0,
        negative_sample_rate: int = 5,
        random_state: int | None = None,
        verbose: bool = False,
        pca_dim: int | None = None,
        knn_method: str = "auto",
        normalize: str | bool = False,
        init="spectral",
        knn_dtype: str = "fp32",
    ):
        self.n_components = n_components
        self.n_neighbors = n_neighbors
        self.min_dist = min_dist
        self.spread = spread
        self.n_epochs = n_epochs
        self.learning_rate = learning_rate
        self.negative_sample_rate = negative_sample_rate
        self.random_state = random_state
        self.verbose = verbose
        self.pca_dim = pca_dim
        self.knn_method = knn_method
        self.normalize = normalize
        self.init = init
        self.knn_dtype = knn_dtype
        self.embedding_ = None
        self.graph_ = None
        self.knn_ = None

    # -- pieces ------------------------------------------------------------
    def _pca(self, Xd: torch.Tensor) -> torch.Tensor:
        # umap.py:92-113: covariance and projection on the device, eigh on the host
        from ._pca import pca_device

        return pca_device(Xd, int(self.pca_dim), self.verbose)

    def _knn(self, Xd: torch.Tensor):
        n = Xd.shape[0]
        method = self.knn_method
        if method == "auto":
            from ._nndescent import nndescent_fits

            method = "brute" if n <= 20000 or not nndescent_fits(min(self.n_neighbors, n - 1)) else "nndescent"
        if method == "brute":
            from ._knn import knn_brute_device

            return knn_brute_device(Xd, self.n_neighbors, dtype=self.knn_dtype, euclid=True)
        if method == "nndescent":
            from ._nndescent import NNDescent

            nn = NNDescent(k=self.n_neighbors, verbose=self.verbose,
                           random_state=self.random_state if self.random_state is not None else 42)
            return nn.build_device(Xd)
        raise ValueError(f"Unknown knn method: {method}")

    def _init_embedding(self, g: FuzzyGraph, n: int) -> torch.Tensor:
        init = self.init
        dev = g.knn_idx.device
        if isinstance(init, str) and init == "random":
            rs = np.random.RandomState(self.random_state)
            Y0 = rs.uniform(-10.0, 10.0, size=(n, self.n_components)).astype(np.float32)
            return torch.from_numpy(Y0).to(dev)
        if isinstance(init, str) and init == "spectral":
            from ._spectral import spectral_init

            return spectral_init(g, n, self.n_components, self.random_state, self.verbose)
        if isinstance(init, str):
            raise ValueError(f"Unknown init: {init}")
        Y0 = np.asarray(init, dtype=np.float32)
        if Y0.shape != (n, self.n_components):
            raise ValueError(f"init array has shape {Y0.shape}, expected {(n, self.n_components)}")
        return torch.from_numpy(np.ascontiguousarray(Y0)).to(dev)

    # -- API ---------------------------------------------------------------
    def fit_transform(self, X, epoch_callback=None) -> np.ndarray:
        """Fit UMAP and return the embedding (n, n_components) float32.

        Args:
            X: input data (n_samples, n_features): numpy, torch (CPU or GPU).
            epoch_callback: optional callable(epoch, Y_numpy) for snapshots.
        """
        if isinstance(X, torch.Tensor) and X.is_cuda and not self.normalize:
            Xd = _device.f32_rows(X.detach())
        else:
            if isinstance(X, torch.Tensor):
                X = X.detach().cpu().numpy()
            X = np.asarray(X, dtype=np.float32)
            X = normalize_input(X, self.normalize)
            Xd = _device.as_device_f32(X)
        if self.pca_dim is not None and Xd.shape[1] > self.pca_dim:
            if self.verbose:
                print(f"Applying PCA: {Xd.shape[1]} -> {self.pca_dim} dims...")
            Xd = self._pca(Xd)
        n = Xd.shape[0]
        if not 1 <= self.n_components <= 8:
            raise ValueError("n_components must be in [1, 8]")

        if self.verbose:
            print("Computing nearest neighbors...")
        knn_idx, knn_dist = self._knn(Xd)
        self.knn_ = (knn_idx, knn_dist)

        if self.n_epochs is None:
            self.n_epochs = 500 if n <= 10000 else 200

        if self.verbose:
            print("Building fuzzy simplicial set...")
        g = fuzzy_graph_device(knn_idx, knn_dist, self.n_epochs)
        self.graph_ = g
        a, b = find_ab_params(self.spread, self.min_dist)
        self._a, self._b = a, b

        if self.verbose:
            print("Initializing embedding...")
        Y = self._init_embedding(g, n).contiguous()
        Y = self._optimize(g, Y, a, b, epoch_callback)
        self.embedding_ = Y.cpu().numpy()
        return self.embedding_

    def _optimize(self, g: FuzzyGraph, Y: torch.Tensor, a: float, b: float, epoch_callback=None):
        n_epochs = int(self.n_epochs)
        ws = epochs_workspace(Y.shape[0], g.k, Y.shape[1])
        seed = self.random_state if self.random_state is not None else int(np.random.randint(0, 2**31 - 1))
        seed = (int(seed) * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF
        if epoch_callback is not None:
            epoch_callback(0, Y.cpu().numpy())
        if g.n_edges == 0:
            return Y
        # Epoch 0 has no active edge (every epochs_per_sample >= 1); every later
        # epoch has the maximum-weight edge (epochs_per_sample == 1).
        if epoch_callback is None:
            step = 50 if self.verbose else n_epochs
            for e0 in range(1, n_epochs, step):
                e1 = min(n_epochs, e0 + step)
                run_epochs(Y, ws, g, e0, e1, n_epochs, self.negative_sample_rate, self.learning_rate, a, b, seed)
                if self.verbose:
                    torch.cuda.synchronize()
                    print(f"Epoch {e1}/{n_epochs}")
        else:
            for ep in range(1, n_epochs):
                run_epochs(Y, ws, g, ep, ep + 1, n_epochs, self.negative_sample_rate, self.learning_rate, a, b,
                           seed)
                epoch_callback(ep + 1, Y.cpu().numpy())
        return Y
