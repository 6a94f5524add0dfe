"""PaCMAP and LocalMAP on MI355X: drop-ins for ``mlx_vis.PaCMAP`` and
``mlx_vis.LocalMAP`` (mlx_vis/_pacmap/pacmap.py:362-726).

Same constructors (n_components=2, n_neighbors=10, MN_ratio=0.5,
FP_ratio=2.0, lr=1.0, num_iters=(100, 100, 250), random_state=None,
verbose=False, apply_pca=True, knn_method="auto", normalize=False; LocalMAP
adds low_dist_thres=10.0) and ``fit_transform(X, init="pca",
epoch_callback=None)``.

Host steps follow the reference: normalisation, PCA to 100 dimensions or the
min/max/mean normalisation (pacmap.py:416-453), pair counts (455-495), the
shared ``compute_knn`` operator with k = n_neighbors + 50, and the three pair
samplers (65-177), which draw from ``np.random.default_rng(random_state)``
in the reference's order, so that the same seed gives the same pairs. The
further-pair sampler draws each chunk of rows in one call and keeps, per
row, the first valid distinct candidates in draw order, which is what the
reference's row loop keeps from the same stream; a chunk in which some row
needs a second round is redone row by row from the saved generator state.

The optimisation (pacmap.py:583-667) runs on the device: per iteration the
neighbour, mid-near and further-pair force laws (mv_pair_grad) accumulate one
gradient buffer, and mv_adam_update applies the Adam step in the reference's
form (bias corrections folded into lr_t, eps 1e-7 on sqrt(v)). LocalMAP's
further-pair resampling (180-227) runs on the host every ten iterations of
the last phase; the reference draws its candidates from MLX's generator, here
they come from the same NumPy generator as the other pairs.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _capi, _device
from ._knn import compute_knn
from ._normalize import normalize_input

LAW_NN, LAW_MN, LAW_FP, LAW_LOCAL_NN = 0, 1, 2, 3
_CHUNK = 4096


def sample_neighbors(knn_dist: np.ndarray, knn_idx: np.ndarray, n_neighbors: int) -> np.ndarray:
    """Neighbour pairs: per row the n_neighbors smallest scaled distances
    d^2 / (sigma_i sigma_j), sigma = mean of the 4th-6th neighbour distances
    (pacmap.py:65-81, FP32 as there)."""
    n = knn_dist.shape[0]
    knn_dist = np.asarray(knn_dist, dtype=np.float32)
    sig = np.maximum(np.mean(knn_dist[:, 3:6], axis=1), np.float32(1e-10))
    scaled = knn_dist ** 2 / (sig[:, None] * sig[knn_idx])
    order = np.argsort(scaled, axis=1, kind="stable")[:, :n_neighbors]
    picked = np.take_along_axis(knn_idx, order, axis=1)
    return np.stack([np.repeat(np.arange(n, dtype=np.int32), n_neighbors), picked.ravel().astype(np.int32)], axis=1)


def sample_mn_pairs(X: np.ndarray, n_mn: int, rng: np.random.Generator) -> np.ndarray:
    """Mid-near pairs: per round and row, six candidates (i + offset) % n with
    offsets in [1, n); the second closest is the pair (pacmap.py:84-122)."""
    n = X.shape[0]
    if n_mn == 0:
        return np.empty((0, 2), dtype=np.int32)
    X = np.asarray(X, dtype=np.float32)
    src = np.arange(n, dtype=np.int32)
    out = np.empty((n * n_mn, 2), dtype=np.int32)
    for j in range(n_mn):
        offsets = rng.integers(1, n, size=(n, 6)).astype(np.int32)
        cand = (src[:, None] + offsets) % n
        picked = np.empty(n, dtype=np.int32)
        for r0 in range(0, n, _CHUNK):
            r1 = min(n, r0 + _CHUNK)
            diff = X[cand[r0:r1]] - X[r0:r1, None, :]
            d2 = np.einsum("ijk,ijk->ij", diff, diff)
            second = np.argsort(d2, axis=1, kind="stable")[:, 1]
            picked[r0:r1] = cand[r0:r1][np.arange(r1 - r0), second]
        out[j * n:(j + 1) * n, 0] = src
        out[j * n:(j + 1) * n, 1] = picked
    return out


def _fp_rows_one_by_one(rows, n, nbr, n_fp, rng):
    """The reference's row loop (pacmap.py:133-175) for the given rows."""
    factor = max(16, 4 * n_fp)
    out = np.empty((len(rows), n_fp), dtype=np.int32)
    for r, i in enumerate(rows):
        chosen, seen = [], set()
        reject = {int(i)}
        if nbr.shape[1]:
            reject.update(int(x) for x in nbr[i])
        rounds = 0
        while len(chosen) < n_fp and rounds < 24:
            remaining = n_fp - len(chosen)
            for c in rng.integers(0, n, size=max(factor, remaining * factor), dtype=np.int32).tolist():
                if c in reject or c in seen:
                    continue
                chosen.append(c)
                seen.add(c)
                if len(chosen) == n_fp:
                    break
            rounds += 1
        if len(chosen) < n_fp:
            unavailable = np.fromiter(reject.union(seen), dtype=np.int32)
            pool = np.setdiff1d(np.arange(n, dtype=np.int32), unavailable)
            if pool.size == 0:
                pool = np.setdiff1d(np.arange(n, dtype=np.int32), np.array([i], dtype=np.int32))
            need = n_fp - len(chosen)
            chosen.extend(rng.choice(pool, size=need, replace=pool.size < need).astype(np.int32).tolist())
        out[r] = chosen[:n_fp]
    return out


def sample_fp_pairs(n: int, pair_neighbors: np.ndarray, n_neighbors: int, n_fp: int,
                    rng: np.random.Generator, chunk: int = _CHUNK) -> np.ndarray:
    """Further pairs: per row the first n_fp uniform draws that are neither
    the row, its neighbours, nor repeats (pacmap.py:125-177)."""
    if n_fp == 0:
        return np.empty((0, 2), dtype=np.int32)
    nbr = (pair_neighbors[:, 1].reshape(n, n_neighbors) if n_neighbors
           else np.empty((n, 0), dtype=np.int32))
    factor = max(16, 4 * n_fp)
    width = max(factor, n_fp * factor)   # the first round's draws per row
    out = np.empty((n * n_fp, 2), dtype=np.int32)
    out[:, 0] = np.repeat(np.arange(n, dtype=np.int32), n_fp)
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        c = r1 - r0
        state = rng.bit_generator.state
        vals = rng.integers(0, n, size=c * width, dtype=np.int32).reshape(c, width)
        ok = vals != np.arange(r0, r1, dtype=np.int32)[:, None]
        for t in range(nbr.shape[1]):
            ok &= vals != nbr[r0:r1, t:t + 1]
        # first occurrence of each accepted value, in draw order
        key = np.where(ok, vals.astype(np.int64), -1)
        order = np.argsort(key, axis=1, kind="stable")
        sk = np.take_along_axis(key, order, axis=1)
        first = np.ones_like(ok)
        first[:, 1:] = sk[:, 1:] != sk[:, :-1]
        keep = np.zeros_like(ok)
        np.put_along_axis(keep, order, first & (sk >= 0), axis=1)
        csum = np.cumsum(keep, axis=1)
        if (csum[:, -1] < n_fp).any():
            rng.bit_generator.state = state
            out[r0 * n_fp:r1 * n_fp, 1] = _fp_rows_one_by_one(range(r0, r1), n, nbr, n_fp, rng).ravel()
            continue
        out[r0 * n_fp:r1 * n_fp, 1] = vals[keep & (csum <= n_fp)]
    return out


def resample_local_fp_pairs(pair_neighbors: np.ndarray, pair_fp: np.ndarray, Y: np.ndarray, low_dist_thres: float,
                            rng: np.random.Generator) -> np.ndarray:
    """LocalMAP's further-pair resampling (pacmap.py:180-227): per row, the
    first of max(8 n_fp, 64) uniform candidates that are not the row or its
    neighbours and lie within low_dist_thres in the embedding; rows without
    enough keep their previous pairs."""
    if pair_fp.size == 0 or low_dist_thres <= 0:
        return pair_fp
    n = Y.shape[0]
    n_fp = pair_fp.shape[0] // n
    width = max(n_fp * 8, 64)
    thr2 = float(low_dist_thres) ** 2
    nn = pair_neighbors.shape[0] // n if pair_neighbors.size else 0
    nbr = pair_neighbors[:, 1].reshape(n, nn) if nn else np.empty((n, 0), dtype=np.int32)
    cand = rng.integers(0, n, size=(n, width)).astype(np.int32)
    out = pair_fp.copy()
    existing = pair_fp[:, 1].reshape(n, n_fp)
    cols = np.arange(width)
    for r0 in range(0, n, _CHUNK):
        r1 = min(n, r0 + _CHUNK)
        cc = cand[r0:r1]
        mask = cc != np.arange(r0, r1, dtype=np.int32)[:, None]
        for t in range(nn):
            mask &= cc != nbr[r0:r1, t:t + 1]
        diff = Y[cc] - Y[r0:r1, None, :]
        mask &= np.einsum("ijk,ijk->ij", diff, diff) <= thr2
        order = np.argsort(np.where(mask, cols[None, :], width), axis=1, kind="stable")[:, :n_fp]
        got = np.take_along_axis(cc, order, axis=1)
        valid = np.take_along_axis(mask, order, axis=1)
        out[r0 * n_fp:r1 * n_fp, 1] = np.where(valid, got, existing[r0:r1]).ravel()
    return out


def decide_num_pairs(n: int, n_neighbors, mn_ratio: float, fp_ratio: float, log=None):
    """Pair counts (pacmap.py:455-495)."""
    log = log or (lambda _m: None)
    if n_neighbors is None:
        n_neighbors = 10 if n <= 10_000 else int(round(10 + 15 * (np.log10(n) - 4)))
    else:
        n_neighbors = int(n_neighbors)
    n_mn = int(round(n_neighbors * mn_ratio))
    n_fp = int(round(n_neighbors * fp_ratio))
    if n - 1 < n_neighbors:
        log("Sample size is smaller than requested neighbours; reducing n_neighbors")
    n_neighbors = min(n_neighbors, n - 1)
    if n - 1 < n_mn:
        log("Sample size is smaller than requested mid-near pairs; reducing n_MN")
    n_mn = min(n_mn, n - 1)
    if n - 1 < n_fp:
        log("Sample size is smaller than requested further pairs; reducing n_FP")
    n_fp = min(n_fp, max(0, n - 1 - n_neighbors))
    if n_neighbors + n_mn + n_fp >= n:
        log("Sample size is smaller than total assigned points; reorganising n_neighbors, n_MN, n_FP")
        denom = 1.0 + mn_ratio + fp_ratio
        n_neighbors = max(1, int(n / denom))
        n_mn = int(n_neighbors * mn_ratio)
        n_fp = int(n_neighbors * fp_ratio)
        n_neighbors = min(n_neighbors, n - 1)
        n_mn = min(n_mn, n - 1)
        n_fp = min(n_fp, max(0, n - 1 - n_neighbors))
    if n_neighbors < 1:
        raise ValueError("The number of nearest neighbours can't be less than 1")
    return n_neighbors, n_mn, n_fp


class _Pairs:
    """One pair set on the device (contiguous src and dst)."""

    def __init__(self, pairs: np.ndarray, dev):
        self.count = int(pairs.shape[0])
        self.src = torch.from_numpy(np.ascontiguousarray(pairs[:, 0], dtype=np.int32)).to(dev)
        self.dst = torch.from_numpy(np.ascontiguousarray(pairs[:, 1], dtype=np.int32)).to(dev)


class PaCMAP:
    """PaCMAP dimensionality reduction on an MI355X GPU."""

    def __init__(self, n_components=2, n_neighbors=10, MN_ratio=0.5, FP_ratio=2.0, lr=1.0,
                 num_iters=(100, 100, 250), random_state=None, verbose=False, apply_pca=True,
                 knn_method: str = "auto", normalize: str | bool = False):
        self.n_components = n_components
        self.n_neighbors = n_neighbors
        self.MN_ratio = MN_ratio
        self.FP_ratio = FP_ratio
        self.lr = lr
        self.num_iters = (100, 100, num_iters) if isinstance(num_iters, int) else tuple(num_iters)
        self.random_state = random_state
        self.verbose = verbose
        self.apply_pca = apply_pca
        self.knn_method = knn_method
        self.normalize = normalize
        self.embedding_ = None

    def _log(self, msg):
        if self.verbose:
            print(msg)

    def _preprocess(self, X: np.ndarray) -> np.ndarray:
        """PCA to 100 dimensions when dim > 100 (and apply_pca), else the
        min/max/mean normalisation (pacmap.py:416-453)."""
        n, dim = X.shape
        if dim > 100 and self.apply_pca:
            xmean = np.mean(X, axis=0)
            Xc = X - xmean
            cov = (Xc.T.astype(np.float64) @ Xc.astype(np.float64)) / (n - 1)
            vals, vecs = np.linalg.eigh(cov)
            proj = vecs[:, np.argsort(vals)[-100:][::-1]].astype(np.float32)
            self._pca_solution = True
            self._tsvd_proj = proj
            self._xmean = xmean
            self._log("Applied PCA, dimensionality reduced to 100")
            return np.ascontiguousarray(Xc @ proj, dtype=np.float32)
        xmin = np.min(X)
        X = X - xmin
        xmax = np.max(X)
        X = X / xmax
        xmean = np.mean(X, axis=0)
        X = X - xmean
        self._pca_solution = False
        self._xmin, self._xmax, self._xmean = xmin, xmax, xmean
        self._log("X normalised")
        return np.ascontiguousarray(X, dtype=np.float32)

    def _decide_num_pairs(self, n):
        return decide_num_pairs(n, self.n_neighbors, self.MN_ratio, self.FP_ratio, self._log)

    def _init(self, Xp: np.ndarray, init: str, rng) -> np.ndarray:
        n = Xp.shape[0]
        if init == "pca":
            if self._pca_solution:
                return np.ascontiguousarray(Xp[:, :self.n_components] * np.float32(0.01), dtype=np.float32)
            Xc = Xp - Xp.mean(axis=0)
            if n > Xp.shape[1]:
                cov = (Xc.T.astype(np.float64) @ Xc.astype(np.float64)) / (n - 1)
                vals, vecs = np.linalg.eigh(cov)
                vecs = vecs[:, np.argsort(vals)[-self.n_components:]][:, ::-1]
                Y = Xc @ vecs.astype(np.float32)
            else:
                U, S, _ = np.linalg.svd(Xc.astype(np.float64), full_matrices=False)
                Y = U[:, :self.n_components] * S[:self.n_components]
            return np.ascontiguousarray(Y * 0.01, dtype=np.float32)
        if init == "random":
            return np.ascontiguousarray(rng.normal(size=(n, self.n_components)).astype(np.float32) * 0.0001,
                                        dtype=np.float32)
        raise ValueError(f"Unknown init: {init}")

    def fit_transform(self, X, init="pca", epoch_callback=None) -> np.ndarray:
        """Fit and return the embedding (n, n_components) float32."""
        return self._fit_transform_impl(X, init=init, epoch_callback=epoch_callback)

    def _fit_transform_impl(self, X, init="pca", epoch_callback=None, low_dist_thres=None) -> np.ndarray:
        if isinstance(X, torch.Tensor):
            X = X.detach().cpu().numpy()
        X = normalize_input(np.array(X, dtype=np.float32), self.normalize)
        n = X.shape[0]
        if n <= 0:
            raise ValueError("The sample size must be larger than 0")
        if not 1 <= self.n_components <= 4:
            raise ValueError("n_components must be in [1, 4]")
        rng = np.random.default_rng(self.random_state)
        Xp = self._preprocess(X.copy())
        n_neighbors, n_mn, n_fp = self._decide_num_pairs(n)
        self._log(f"n={n}, dim={Xp.shape[1]}, n_neighbors={n_neighbors}, n_MN={n_mn}, n_FP={n_fp}")
        knn_idx, knn_dist = compute_knn(Xp, min(n - 1, n_neighbors + 50), method=self.knn_method,
                                        verbose=self.verbose, random_state=self.random_state)
        pair_nn = sample_neighbors(np.asarray(knn_dist), np.asarray(knn_idx), n_neighbors)
        pair_mn = sample_mn_pairs(Xp, n_mn, rng)
        pair_fp = sample_fp_pairs(n, pair_nn, n_neighbors, n_fp, rng)
        self._log(f"Pairs: NN={pair_nn.shape}, MN={pair_mn.shape}, FP={pair_fp.shape}")
        self.pairs_ = (pair_nn, pair_mn, pair_fp)
        Y0 = self._init(Xp, init, rng)
        self.embedding_ = self._optimize(Y0, pair_nn, pair_mn, pair_fp, rng, low_dist_thres, epoch_callback)
        return self.embedding_

    def _optimize(self, Y0, pair_nn, pair_mn, pair_fp, rng, low_dist_thres=None, epoch_callback=None) -> np.ndarray:
        """The three-phase schedule of pacmap.py:583-667 on the device."""
        dev = _device.device()
        n, dim = Y0.shape
        Y = torch.from_numpy(np.ascontiguousarray(Y0, dtype=np.float32)).to(dev)
        state = torch.zeros((3, n, dim), dtype=torch.float32, device=dev)   # gradient, Adam m, v
        g, m, v = state[0], state[1], state[2]
        nn, mn, fp = _Pairs(pair_nn, dev), _Pairs(pair_mn, dev), _Pairs(pair_fp, dev)
        lib = _capi.lib()
        stream = _device.stream()
        localmap = low_dist_thres is not None
        local_scale = float(low_dist_thres) / 2.0 if localmap else 1.0
        p1, p2, _ = self.num_iters
        total = sum(self.num_iters)

        def law(pairs: _Pairs, w: float, kind: int, scale: float = 1.0):
            if pairs.count:
                _capi.check(lib.mv_pair_grad(Y.data_ptr(), n, dim, pairs.src.data_ptr(), pairs.dst.data_ptr(),
                                             pairs.count, None, float(w), float(scale), kind, g.data_ptr(), stream),
                            "mv_pair_grad")

        if epoch_callback is not None:
            epoch_callback(0, Y.cpu().numpy())
        for itr in range(total):
            step = self.lr * math.sqrt(1.0 - 0.999 ** (itr + 1)) / (1.0 - 0.9 ** (itr + 1))
            if itr < p1:
                w_mn = (1.0 - itr / p1) * 1000.0 + (itr / p1) * 3.0
                law(nn, 2.0, LAW_NN)
                law(mn, w_mn, LAW_MN)
            elif itr < p1 + p2:
                law(nn, 3.0, LAW_NN)
                law(mn, 3.0, LAW_MN)
            elif localmap:
                law(nn, 1.0, LAW_LOCAL_NN, local_scale)
            else:
                law(nn, 1.0, LAW_NN)
            law(fp, 1.0, LAW_FP)
            _capi.check(lib.mv_adam_update(Y.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), n * dim, 0.9,
                                           0.999, float(step), 1e-7, stream), "mv_adam_update")
            if localmap and itr > p1 + p2 and itr % 10 == 0 and pair_fp.size:
                pair_fp = resample_local_fp_pairs(pair_nn, pair_fp, Y.cpu().numpy(), low_dist_thres, rng)
                fp = _Pairs(pair_fp, dev)
            if epoch_callback is not None:
                epoch_callback(itr + 1, Y.cpu().numpy())
        return Y.cpu().numpy()


class LocalMAP(PaCMAP):
    """LocalMAP: PaCMAP with local graph adjustment in the final phase."""

    def __init__(self, n_components=2, n_neighbors=10, MN_ratio=0.5, FP_ratio=2.0, lr=1.0,
                 num_iters=(100, 100, 250), random_state=None, verbose=False, apply_pca=True,
                 knn_method: str = "auto", low_dist_thres=10.0, normalize: str | bool = False):
        super().__init__(n_components=n_components, n_neighbors=n_neighbors, MN_ratio=MN_ratio, FP_ratio=FP_ratio,
                         lr=lr, num_iters=num_iters, random_state=random_state, verbose=verbose,
                         apply_pca=apply_pca, knn_method=knn_method, normalize=normalize)
        self.low_dist_thres = float(low_dist_thres)

    def fit_transform(self, X, init="pca", epoch_callback=None) -> np.ndarray:
        return self._fit_transform_impl(X, init=init, epoch_callback=epoch_callback,
                                        low_dist_thres=self.low_dist_thres)
