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
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


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


if __name__ == "__main__":
    make_knn()
    make_ab()
    print("wrote", sorted(p.name for p in HERE.iterdir() if p.suffix in (".npz", ".json")))
