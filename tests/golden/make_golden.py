"""Writes the golden fixtures of tests/golden (run from the repository root:
``python tests/golden/make_golden.py``). The expected values come from
libraries independent of this repository's oracle and kernels:

- knn_c0_head1000.npz: exact k=15 neighbours (indices, squared Euclidean
  distances) of the first 1,000 rows of the small plain-method input
  (10,000 x 128 N(0, 1) float32, ``np.random.default_rng(0)``) against all
  10,000 rows, self included as the reference computes it (self masked out,
  umap/_knn.py:60-69 semantics: the k nearest other points), from
  scikit-learn's brute-force NearestNeighbors in float64.
- umap_ab.json: a, b of the UMAP curve 1 / (1 + a x^(2b)) for (spread,
  min_dist) pairs, fitted by scipy.optimize.curve_fit on the 300 points of
  umap.py:272-296.
- knn_c1_sample2000.npz: exact k=15 neighbours (indices, squared Euclidean
  distances, self excluded) of 2,000 sampled rows of the 70,000 x 784
  10-Gaussian mixture (mlx_vis_amd.datasets.gaussian_mixture_784, seed 1;
  rows = sorted default_rng(123).choice(70000, 2000)) against all 70,000
  rows, from scikit-learn's brute-force NearestNeighbors in float64.
- nnd_c1.npz, nnd_c3.npz: the NNDescent oracle (oracle/nndescent.py, the
  NumPy restatement of mlx_vis/_nndescent/nndescent.py with the reference's
  NumPy random stream) on the 70,000 x 784 mixture (k=15, n_iters=12,
  max_candidates=30, random_state=0) and on the 1,000,000 x 128 mixture
  (gaussian_mixture_128, seed 2; k=15, n_iters=12, random_state=42): its
  per-iteration changed counts, its neighbours on 10,000 sampled rows (sorted
  default_rng(0).choice(n, 10000)), the exact 15-NN of those rows
  (oracle.knn.exact_knn) and its recall on them (plus, at 70,000 rows, its
  recall over all rows). These are the "reference recall" figures an
  NNDescent on the device must reach.

- umap_c0_oracle.json: the oracle's layout of the 10,000 x 128 workload
  (make_umap_c0): kept edges, 15-NN preservation, trustworthiness.
- umap_c1_oracle.json: the oracle's layouts of the 70,000 x 784 mixture
  (make_umap_c1, 500 epochs, random and spectral init): kept edges, 15-NN
  preservation and trustworthiness on the 2,000 rows of knn_c1_sample2000.npz.

Usage: ``python tests/golden/make_golden.py [knn_c0 ab knn_c1 nnd_c1 nnd_c3 umap_c0 umap_c1]``
(no argument: the fast fixtures knn_c0 and ab).
"""

from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for _p in (ROOT, ROOT / "mlx-vis_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))


def c0_input() -> np.ndarray:
    return np.random.default_rng(0).standard_normal((10000, 128), dtype=np.float32)


def make_knn():
    from sklearn.neighbors import NearestNeighbors

    X = c0_input().astype(np.float64)
    nn = NearestNeighbors(n_neighbors=16, algorithm="brute", metric="euclidean").fit(X)
    dist, idx = nn.kneighbors(X[:1000])
    # drop self (distance 0, always first for distinct rows)
    assert (idx[:, 0] == np.arange(1000)).all()
    np.savez_compressed(HERE / "knn_c0_head1000.npz", idx=idx[:, 1:].astype(np.int32),
                        dist2=(dist[:, 1:] ** 2).astype(np.float64))


def make_ab():
    from scipy.optimize import curve_fit

    out = []
    for spread, min_dist in [(1.0, 0.1), (1.0, 0.5), (2.0, 0.25), (1.0, 0.001)]:
        x = np.linspace(0, spread * 3, 300)
        y = np.where(x < min_dist, 1.0, np.exp(-(x - min_dist) / spread))
        (a, b), _ = curve_fit(lambda x, a, b: 1.0 / (1.0 + a * x ** (2 * b)), x, y, p0=(1.0, 1.0), maxfev=20000)
        out.append({"spread": spread, "min_dist": min_dist, "a": float(a), "b": float(b)})
    (HERE / "umap_ab.json").write_text(json.dumps(out, indent=1) + "\n")


C1_SAMPLE_SEED = 123


def c1_sample_rows(n: int = 70_000, m: int = 2000) -> np.ndarray:
    return np.sort(np.random.default_rng(C1_SAMPLE_SEED).choice(n, m, replace=False)).astype(np.int64)


def make_knn_c1():
    from sklearn.neighbors import NearestNeighbors

    from mlx_vis_amd.datasets import gaussian_mixture_784
    from oracle import knn as oknn

    X = gaussian_mixture_784(70_000, seed=1)
    rows = c1_sample_rows()
    t = time.time()
    nn = NearestNeighbors(n_neighbors=16, algorithm="brute", metric="euclidean").fit(X.astype(np.float64))
    dist, idx = nn.kneighbors(X[rows].astype(np.float64))
    assert (idx[:, 0] == rows).all() and (dist[:, 0] < 1e-5).all()   # self (zero up to expansion rounding)
    idx, dist2 = idx[:, 1:].astype(np.int32), (dist[:, 1:] ** 2).astype(np.float64)
    print(f"knn_c1: scikit-learn {time.time() - t:.1f} s", flush=True)
    # cross-check: the canonical order of the oracle gives the same neighbours
    oi, od = oknn.exact_knn(X[rows], X, 15, q_self=rows)
    assert np.array_equal(oi, idx), int((oi != idx).sum())
    np.testing.assert_allclose(od, dist2, rtol=1e-9)
    np.savez_compressed(HERE / "knn_c1_sample2000.npz", rows=rows.astype(np.int32), idx=idx, dist2=dist2)


def _nnd_fixture(name, X, seed, n_iters, max_candidates, full_recall):
    from oracle import knn as oknn
    from oracle import metrics as om
    from oracle import nndescent as onn

    n = X.shape[0]
    rows = np.sort(np.random.default_rng(0).choice(n, 10_000, replace=False)).astype(np.int64)
    t = time.time()
    idx, _, trace = onn.nndescent(X, k=15, n_iters=n_iters, max_candidates=max_candidates, random_state=seed,
                                  return_trace=True)
    print(f"{name}: oracle NNDescent {time.time() - t:.1f} s, trace {trace}", flush=True)
    t = time.time()
    ex_rows, _ = oknn.exact_knn(X[rows], X, 15, q_self=rows)
    print(f"{name}: exact kNN of the sampled rows {time.time() - t:.1f} s", flush=True)
    out = dict(rows=rows.astype(np.int32), oracle_idx=idx[rows].astype(np.int32), exact_idx=ex_rows.astype(np.int32),
               trace=np.asarray(trace, dtype=np.int64), seed=np.int64(seed),
               recall_rows=np.float64(om.recall(idx[rows], ex_rows)))
    if full_recall:
        t = time.time()
        ex_all, _ = oknn.exact_knn(X, X, 15, self_base=0)
        out["recall_all"] = np.float64(om.recall(idx, ex_all))
        print(f"{name}: exact kNN of all rows {time.time() - t:.1f} s", flush=True)
    print(f"{name}: recall on the sampled rows {float(out['recall_rows']):.5f}"
          + (f", all rows {float(out['recall_all']):.5f}" if full_recall else ""), flush=True)
    np.savez_compressed(HERE / f"{name}.npz", **out)


def make_nnd_c1():
    from mlx_vis_amd.datasets import gaussian_mixture_784

    _nnd_fixture("nnd_c1", gaussian_mixture_784(70_000, seed=1), 0, 12, 30, True)


def make_nnd_c3():
    from mlx_vis_amd.datasets import gaussian_mixture_128

    _nnd_fixture("nnd_c3", gaussian_mixture_128(1_000_000, seed=2), 42, 12, None, False)


def make_umap_c0():
    """The oracle's layout of the small plain-method workload: exact kNN
    (k = 15), fuzzy set pruned for 200 epochs, a/b for (spread 1, min_dist
    0.1), init RandomState(0).uniform(-10, 10, (n, 2)) (what UMAP(init="random",
    random_state=0) draws), 200 epochs with 5 NumPy-drawn negatives per edge
    (seed 0); its kept-edge count, 15-NN preservation over all rows and
    trustworthiness (k = 15) over 2,000 sampled rows."""
    from oracle import knn as oknn
    from oracle import metrics as om
    from oracle import umap as ou

    X = c0_input()
    n, n_epochs = X.shape[0], 200
    t = time.time()
    idx, d2 = oknn.exact_knn(X, X, 15, self_base=0)
    dist = np.sqrt(d2).astype(np.float32)
    rows, cols, vals = ou.fuzzy_simplicial_set(idx.astype(np.int32), dist, n_epochs)
    a, b = ou.find_ab_params(1.0, 0.1)
    Y0 = np.random.RandomState(0).uniform(-10.0, 10.0, size=(n, 2)).astype(np.float32)
    Y = ou.optimize(rows, cols, vals, Y0, a, b, n_epochs, neg_rate=5, lr=1.0, seed=0)
    print(f"umap_c0: oracle layout {time.time() - t:.1f} s", flush=True)
    sample = np.sort(np.random.default_rng(7).choice(n, 2000, replace=False))
    out = {
        "workload": "10,000 x 128 N(0,1) seed 0, k=15, 200 epochs, 5 negatives, random init seed 0",
        "n_edges": int(rows.size),
        "knn15_preservation": om.knn_preservation(X, Y, 15),
        "trustworthiness_15": om.trustworthiness(X, Y, 15, rows=sample),
        "trust_rows_seed": 7,
        "trust_rows": 2000,
    }
    print("umap_c0:", out, flush=True)
    (HERE / "umap_c0_oracle.json").write_text(json.dumps(out, indent=1) + "\n")


def make_umap_c1():
    """The oracle's layouts of the 70,000 x 784 mixture (seed 1): exact kNN
    (k = 15), fuzzy set pruned for 500 epochs, a/b for (spread 1, min_dist
    0.1), then 500 epochs of the reference's loop (umap.py:423-447: the
    schedule of umap.schedule, 5 NumPy-drawn negatives per active edge, seed
    0) with the torch-CPU scatters of umap.sgd_step_index_add (the same FP32
    arithmetic as umap.sgd_step, pinned to it by test_oracle_umap), from two
    inits: "random", RandomState(0).uniform(-10, 10, (n, 2)) (what
    UMAP(init="random", random_state=0) draws), and "spectral", the reference's
    _spectral_init (umap.py:298-357) restated by umap.spectral_power (100
    power iterations from default_rng(0) normal start vectors) and
    umap.spectral_scale (noise RandomState(random_state or 42), i.e. 42).
    Records the kept-edge count, and per init 15-NN preservation and
    trustworthiness (k = 15) over the 2,000 rows of knn_c1_sample2000.npz."""
    import torch

    from mlx_vis_amd.datasets import gaussian_mixture_784
    from oracle import knn as oknn
    from oracle import metrics as om
    from oracle import umap as ou

    X = gaussian_mixture_784(70_000, seed=1)
    n, n_epochs, neg_rate = X.shape[0], 500, 5
    t = time.time()
    idx, d2 = oknn.exact_knn(X, X, 15, self_base=0)
    print(f"umap_c1: oracle exact kNN {time.time() - t:.1f} s", flush=True)
    dist = np.sqrt(d2).astype(np.float32)
    rows, cols, vals = ou.fuzzy_simplicial_set(idx.astype(np.int32), dist, n_epochs)
    a, b = ou.find_ab_params(1.0, 0.1)
    sample = c1_sample_rows()
    out = {
        "workload": "70,000 x 784 10-Gaussian mixture seed 1 (min-max scaled), k=15, 500 epochs, 5 negatives",
        "n_edges": int(rows.size),
        "rows": "knn_c1_sample2000.npz",
    }
    V = ou.spectral_power(rows, cols, vals.astype(np.float64), n, np.random.default_rng(0).standard_normal((n, 3)))
    inits = {
        "random": np.random.RandomState(0).uniform(-10.0, 10.0, size=(n, 2)).astype(np.float32),
        "spectral": ou.spectral_scale(V, np.random.RandomState(42).standard_normal((n, 2))),
    }
    for name, Y0 in inits.items():
        Y = torch.from_numpy(np.ascontiguousarray(Y0, dtype=np.float32))
        rng = np.random.default_rng(0)
        sets = ou.schedule(vals, n_epochs)
        t = time.time()
        for epoch in range(n_epochs):
            act = sets[epoch]
            if act.size == 0:
                continue
            ef, et = rows[act].astype(np.int64), cols[act].astype(np.int64)
            A = act.size
            nf = ef[np.arange(neg_rate * A) % A]
            nt = rng.integers(0, n, size=neg_rate * A)
            alpha = 1.0 * (1.0 - epoch / n_epochs)
            Y = ou.sgd_step_index_add(Y, *(torch.from_numpy(np.ascontiguousarray(v)) for v in (ef, et, nf, nt)),
                                      np.float32(alpha), np.float32(a), np.float32(b))
        print(f"umap_c1: oracle layout ({name} init) {time.time() - t:.1f} s", flush=True)
        Y = Y.numpy()
        out[name] = {"knn15_preservation": om.knn_preservation(X, Y, 15, rows=sample),
                     "trustworthiness_15": om.trustworthiness(X, Y, 15, rows=sample)}
        print("umap_c1:", name, out[name], flush=True)
    (HERE / "umap_c1_oracle.json").write_text(json.dumps(out, indent=1) + "\n")


MAKERS = {"knn_c0": make_knn, "ab": make_ab, "knn_c1": make_knn_c1, "nnd_c1": make_nnd_c1, "nnd_c3": make_nnd_c3,
          "umap_c0": make_umap_c0, "umap_c1": make_umap_c1}


if __name__ == "__main__":
    for name in sys.argv[1:] or ["knn_c0", "ab"]:
        MAKERS[name]()
    print("fixtures:", sorted(p.name for p in HERE.iterdir() if p.suffix in (".npz", ".json")))
