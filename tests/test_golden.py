"""The oracle and the host code against the committed golden fixtures of
tests/golden (values from scikit-learn and SciPy, written by
tests/golden/make_golden.py)."""

import json
from pathlib import Path

import numpy as np

from oracle import knn as oknn
from oracle import umap as ou

GOLDEN = Path(__file__).resolve().parent / "golden"


def c0_input():
    return np.random.default_rng(0).standard_normal((10000, 128), dtype=np.float32)


def test_oracle_knn_matches_golden_c0_head():
    g = np.load(GOLDEN / "knn_c0_head1000.npz")
    X = c0_input()
    oi, od = oknn.exact_knn(X[:1000], X, 15, self_base=0)
    np.testing.assert_array_equal(oi, g["idx"])
    np.testing.assert_allclose(od, g["dist2"], rtol=1e-9)


def test_ab_params_match_golden():
    from mlx_vis_amd import find_ab_params

    for row in json.loads((GOLDEN / "umap_ab.json").read_text()):
        for fn in (ou.find_ab_params, find_ab_params):
            a, b = fn(row["spread"], row["min_dist"])
            np.testing.assert_allclose([a, b], [row["a"], row["b"]], rtol=1e-5)


def test_philox_known_answers():
    """The NumPy Philox-4x32-10 the GPU tests use to reproduce device draws,
    against the Random123 known-answer vectors (philox4x32, 10 rounds)."""
    from _philox import philox4x32_10

    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = [int(v) for v in philox4x32_10(*ctr, *key)]
        assert got == list(want), (ctr, key, [hex(v) for v in got])


def test_oracle_knn_matches_golden_c1_sample():
    """The oracle's canonical order on the 70,000 x 784 mixture (200 of the
    fixture's 2,000 sampled rows, scikit-learn float64 brute force)."""
    from mlx_vis_amd.datasets import gaussian_mixture_784

    g = np.load(GOLDEN / "knn_c1_sample2000.npz")
    X = gaussian_mixture_784(70_000, seed=1)
    rows = g["rows"][::10].astype(np.int64)
    oi, od = oknn.exact_knn(X[rows], X, 15, q_self=rows)
    np.testing.assert_array_equal(oi, g["idx"][::10])
    np.testing.assert_allclose(od, g["dist2"][::10], rtol=1e-9)
