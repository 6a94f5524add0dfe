"""Learned form: a parametric MLP embedding trained on the UMAP graph.

The reference has no parametric UMAP. Its one MLP is MMAE
(mlx_vis/_mmae/mmae.py:22-60), an autoencoder fitted to a reference embedding
(``reference=``, mmae.py:99-112) and trained with AdamW (mmae.py:258-281):
the precedent for fitting a network to an embedding and encoding new points.

``ParametricUMAP`` fits the plain ``UMAP`` (or takes a fitted one), then
trains an MLP din -> 512 -> 512 -> n_components (GELU, BF16 weights with FP32
master copy, FP32 accumulation, Adam) on batches of graph edges drawn in
proportion to their weights: the UMAP edge cross-entropy (the layout's
gradient law, umap.py:367-380, without the clip; in-batch tails as
negatives) plus a regression of both ends to the plain embedding.
``transform`` maps new points in one pass. ``refine`` is the plain layout as
check and fallback: points whose 15-NN preservation against the exact high-D
kNN falls below a threshold are moved by plain SGD epochs (others fixed).

All arithmetic runs in libmivis (mlp.hip, umap.hip, knn.hip); torch only
holds device memory.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _capi, _device
from ._knn import knn_brute_device
from ._umap import UMAP, fuzzy_graph_device, find_ab_params, run_epochs


def _align(x: int, a: int = 256) -> int:
    return (x + a - 1) // a * a


def knn_preservation_device(knn_hi: torch.Tensor, knn_lo: torch.Tensor, threshold: float | None = None):
    """Per-row overlap counts |kNN_hi(i) & kNN_lo(i)| (int32) and, with a
    threshold, the uint8 flags count / k < threshold."""
    _device.require_device("knn_preservation_device", knn_hi, knn_lo)
    n, k = knn_hi.shape
    if knn_lo.shape[0] != n:
        raise ValueError("neighbour lists must have the same number of rows")
    a = knn_hi.to(torch.int32).contiguous()
    b = knn_lo.to(torch.int32).contiguous()
    counts = torch.empty(n, dtype=torch.int32, device=a.device)
    flags = torch.empty(n, dtype=torch.uint8, device=a.device) if threshold is not None else None
    min_count = 0 if threshold is None else int(math.ceil(threshold * k - 1e-9))
    _capi.check(_capi.lib().mv_knn_overlap(a.data_ptr(), a.shape[1], k, b.data_ptr(), b.shape[1], b.shape[1], n,
                                           min_count, counts.data_ptr(), _capi.ptr(flags), _device.stream()),
                "mv_knn_overlap")
    return counts, flags


class ParametricUMAP:
    """UMAP with a learned encoder for out-of-sample points.

    Args:
        n_components: embedding dimension (<= 4).
        hidden: the two hidden widths (multiples of 64, <= 1024).
        n_neighbors, min_dist, spread, negative_sample_rate, n_epochs,
        knn_method, random_state: passed to the plain ``UMAP`` fit.
        batch_size: edges per training step (multiple of 32).
        learning_rate: Adam step size.
        edge_passes: training length in passes over the kept edges (used when
            n_steps is None).
        n_steps: explicit number of Adam steps.
        regression_weight: weight of ||z - y||^2 to the plain embedding at
            the first step.
        regression_decay: "linear" lowers the regression weight linearly to
            zero over training (in `regression_chunks` steps), so that the
            edge loss alone shapes the final local structure; "none" keeps it.
        train_negatives: in-batch negatives per edge in the training loss.
    """

    def __init__(self, n_components: int = 2, hidden=(512, 512), n_neighbors: int = 15, min_dist: float = 0.1,
                 spread: float = 1.0, negative_sample_rate: int = 5, n_epochs: int | None = None,
                 knn_method: str = "brute", batch_size: int = 4096, learning_rate: float = 1e-3,
                 edge_passes: float = 20.0, n_steps: int | None = None, regression_weight: float = 1e-3,
                 regression_decay: str = "linear", regression_chunks: int = 16, train_negatives: int | None = None,
                 random_state: int | None = None, verbose: bool = False):
        self.n_components = n_components
        self.hidden = tuple(int(h) for h in hidden)
        self.n_neighbors = n_neighbors
        self.min_dist = min_dist
        self.spread = spread
        self.negative_sample_rate = negative_sample_rate
        self.n_epochs = n_epochs
        self.knn_method = knn_method
        self.batch_size = int(batch_size)
        self.learning_rate = learning_rate
        self.edge_passes = edge_passes
        self.n_steps = n_steps
        self.regression_weight = regression_weight
        if regression_decay not in ("linear", "none"):
            raise ValueError("regression_decay must be 'linear' or 'none'")
        self.regression_decay = regression_decay
        self.regression_chunks = int(regression_chunks)
        self.train_negatives = negative_sample_rate if train_negatives is None else train_negatives
        self.random_state = random_state
        self.verbose = verbose
        self.umap_ = None
        self.blob_ = None
        self.din_ = None
        self.steps_ = 0
        self.loss_trace_ = None
        self.refine_stats_ = None

    # -- parameters ------------------------------------------------------------
    def _dims(self):
        return (int(self.din_), self.hidden[0], self.hidden[1], int(self.n_components))

    def init_weights(self, din: int, weights: dict | None = None) -> None:
        """Allocates the parameter blob; weights (FP32 arrays W1 (h1, din), b1,
        W2, b2, W3, b3) default to the uniform(+-1/sqrt(fan_in)) init."""
        self.din_ = int(din)
        h1, h2 = self.hidden
        dout = int(self.n_components)
        if weights is None:
            rs = np.random.RandomState(self.random_state)
            weights = {}
            for name, fo, fi in (("1", h1, din), ("2", h2, h1), ("3", dout, h2)):
                bound = 1.0 / math.sqrt(fi)
                weights["W" + name] = rs.uniform(-bound, bound, (fo, fi)).astype(np.float32)
                weights["b" + name] = rs.uniform(-bound, bound, fo).astype(np.float32)
        lib = _capi.lib()
        nb = ctypes.c_size_t(0)
        npar = ctypes.c_int64(0)
        _capi.check(lib.mv_mlp_param_bytes(*self._dims(), ctypes.byref(nb), ctypes.byref(npar)), "mv_mlp_param_bytes")
        self.n_params_ = int(npar.value)
        self.blob_ = _device.workspace(nb.value)
        dev = self.blob_.device
        t = {k: torch.from_numpy(np.ascontiguousarray(weights[k], dtype=np.float32)).to(dev)
             for k in ("W1", "b1", "W2", "b2", "W3", "b3")}
        _capi.check(lib.mv_mlp_init(self.blob_.data_ptr(), *self._dims(), t["W1"].data_ptr(), t["b1"].data_ptr(),
                                    t["W2"].data_ptr(), t["b2"].data_ptr(), t["W3"].data_ptr(), t["b3"].data_ptr(),
                                    _device.stream()), "mv_mlp_init")
        torch.cuda.synchronize()
        self.steps_ = 0

    def master_params(self) -> torch.Tensor:
        """FP32 master parameters in blob order (W1 rows padded to a multiple of 64)."""
        return self.blob_[: 4 * self.n_params_].view(torch.float32)

    def train_workspace(self, n: int) -> torch.Tensor:
        nb = ctypes.c_size_t(0)
        _capi.check(_capi.lib().mv_mlp_batch_bytes(*self._dims(), 2 * self.batch_size, ctypes.byref(nb)),
                    "mv_mlp_batch_bytes")
        dp = _align(self.din_, 64)
        return _device.workspace(nb.value + _align(2 * n * dp))

    # -- training ----------------------------------------------------------------
    def fit(self, X, umap: UMAP | None = None):
        """Fits the plain UMAP (unless a fitted one is given) and trains the MLP on its
        graph and embedding. X: (n, din) float32 (the same data the UMAP saw)."""
        Xd = _device.as_device_f32(X)
        n, din = Xd.shape
        if umap is None:
            umap = UMAP(n_components=self.n_components, n_neighbors=self.n_neighbors, min_dist=self.min_dist,
                        spread=self.spread, n_epochs=self.n_epochs, negative_sample_rate=self.negative_sample_rate,
                        knn_method=self.knn_method, random_state=self.random_state, normalize=False)
            umap.fit_transform(Xd)
        if umap.graph_ is None or umap.embedding_ is None:
            raise ValueError("umap must be fitted")
        self.umap_ = umap
        if self.blob_ is None or self.din_ != din:
            self.init_weights(din)
        g = umap.graph_
        Y = torch.from_numpy(np.ascontiguousarray(umap.embedding_, dtype=np.float32)).to(Xd.device)
        a, b = find_ab_params(self.spread, self.min_dist)
        B = self.batch_size
        steps = self.n_steps or int(math.ceil(self.edge_passes * max(g.n_edges, 1) / B))
        self.train(Xd, Y, g, steps, a, b)
        return self

    def train(self, Xd: torch.Tensor, Y: torch.Tensor, g, steps: int, a: float, b: float,
              ws: torch.Tensor | None = None) -> None:
        """`steps` Adam steps on edges of graph g (continues the step count)."""
        n = Xd.shape[0]
        B = self.batch_size
        lib = _capi.lib()
        idx = torch.empty(steps * 2 * B, dtype=torch.int32, device=Xd.device)
        seed = 0x2545F491 if self.random_state is None else (int(self.random_state) * 2654435761 + 12345)
        _capi.check(lib.mv_mlp_sample_edges(g.eps.data_ptr(), g.knn_idx.data_ptr(), n, g.k, seed & 0xFFFFFFFF,
                                            self.steps_, steps, B, idx.data_ptr(), _device.stream()),
                    "mv_mlp_sample_edges")
        if ws is None:
            ws = self.train_workspace(n)
        trace = torch.empty(2 * max(steps, 1), dtype=torch.float32, device=Xd.device)
        chunks = max(1, min(self.regression_chunks, steps)) if self.regression_decay == "linear" else 1
        done = 0
        for c in range(chunks):
            cnt = (steps * (c + 1)) // chunks - done
            if cnt <= 0:
                continue
            lam = self.regression_weight * (1.0 - c / chunks) if chunks > 1 else self.regression_weight
            _capi.check(lib.mv_mlp_train(self.blob_.data_ptr(), *self._dims(), Xd.data_ptr(), n, Y.data_ptr(),
                                         idx[done * 2 * B:].data_ptr(), cnt, B, int(self.train_negatives), float(a),
                                         float(b), float(lam), float(self.learning_rate), self.steps_,
                                         trace[2 * done:].data_ptr(), ws.data_ptr(), ws.numel(), _device.stream()),
                        "mv_mlp_train")
            self.steps_ += cnt
            done += cnt
        self.loss_trace_ = trace[: 2 * steps].view(steps, 2)

    # -- inference ---------------------------------------------------------------
    def transform_device(self, Xd: torch.Tensor, chunk: int = 65536, ws: torch.Tensor | None = None) -> torch.Tensor:
        if self.blob_ is None:
            raise RuntimeError("ParametricUMAP is not fitted")
        _device.require_device("ParametricUMAP.transform_device", Xd, ws)
        if Xd.dim() != 2:
            raise ValueError(f"X must be (m, din), got shape {tuple(Xd.shape)}")
        Xd = _device.f32_rows(Xd)
        m, din = Xd.shape
        if din != self.din_:
            raise ValueError(f"expected {self.din_} features, got {din}")
        chunk = int(max(1, min(chunk, m)))
        if ws is None:
            ws = self.forward_workspace(chunk)
        Z = torch.empty((m, self.n_components), dtype=torch.float32, device=Xd.device)
        _capi.check(_capi.lib().mv_mlp_forward(self.blob_.data_ptr(), *self._dims(), Xd.data_ptr(), m, Z.data_ptr(),
                                               chunk, ws.data_ptr(), ws.numel(), _device.stream()), "mv_mlp_forward")
        return Z

    def forward_workspace(self, chunk: int) -> torch.Tensor:
        nb = ctypes.c_size_t(0)
        _capi.check(_capi.lib().mv_mlp_batch_bytes(*self._dims(), int(chunk), ctypes.byref(nb)), "mv_mlp_batch_bytes")
        return _device.workspace(nb.value)

    def transform(self, X) -> np.ndarray:
        return self.transform_device(_device.as_device_f32(X)).cpu().numpy()

    def fit_transform(self, X, umap: UMAP | None = None) -> np.ndarray:
        self.fit(X, umap)
        return self.transform(X)

    # -- check and fallback ------------------------------------------------------
    def refine_device(self, Xd: torch.Tensor, Zd: torch.Tensor, threshold: float = 0.8, n_epochs: int = 50,
                      knn=None, learning_rate: float = 1.0, knn_dtype: str = "fp32"):
        """Scores each point of the embedding Zd of Xd by its k-NN preservation
        (k = n_neighbors) against the exact high-D kNN and moves the points below
        `threshold` by `n_epochs` plain layout epochs on the exact-kNN graph; the
        others stay fixed. Returns (refined Z, per-point counts before)."""
        _device.require_device("ParametricUMAP.refine_device", Xd, Zd)
        Xd, Zd = _device.f32_rows(Xd), _device.f32_rows(Zd)
        k = self.n_neighbors
        if knn is None:
            knn = knn_brute_device(Xd, k, dtype=knn_dtype, euclid=True)
        hi_idx, hi_dist = knn
        lo_idx, _ = knn_brute_device(Zd.contiguous(), k, euclid=True)
        counts, flags = knn_preservation_device(hi_idx, lo_idx, threshold)
        g = fuzzy_graph_device(hi_idx, hi_dist, n_epochs)
        a, b = find_ab_params(self.spread, self.min_dist)
        Y = Zd.clone().contiguous()
        seed = 0x51ED270B if self.random_state is None else int(self.random_state) + 7
        run_epochs(Y, None, g, 1, n_epochs, n_epochs, self.negative_sample_rate, learning_rate, a, b, seed,
                   movable=flags)
        self.refine_stats_ = {"flagged": int(flags.sum().item()), "n": int(Xd.shape[0]),
                              "preservation_before": float(counts.double().mean().item()) / k}
        return Y, counts

    def refine(self, X, Z, threshold: float = 0.8, n_epochs: int = 50) -> np.ndarray:
        Y, _ = self.refine_device(_device.as_device_f32(X), _device.as_device_f32(Z), threshold, n_epochs)
        return Y.cpu().numpy()
