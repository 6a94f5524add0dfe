"""normalize_input (mlx_vis/_normalize.py:6-33) against scikit-learn's
scalers, its pass-through and error contract, and its use by UMAP."""

import numpy as np
import pytest

from mlx_vis_amd import UMAP
from mlx_vis_amd._normalize import normalize_input


def _data(seed=0):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((500, 12)) * rng.uniform(0.1, 50, 12) + rng.uniform(-20, 20, 12)).astype(np.float32)
    X[:, 3] = 7.0   # a constant feature
    return X


def test_passthrough():
    X = _data()
    for m in (False, None, ""):
        assert normalize_input(X, m) is X


@pytest.mark.parametrize("method", [True, "standard"])
def test_standard_matches_sklearn(method):
    from sklearn.preprocessing import StandardScaler

    X = _data(1)
    got = normalize_input(X, method)
    assert got.dtype == np.float32 and got.shape == X.shape
    want = StandardScaler().fit_transform(X.astype(np.float64))
    # the reference divides by std + 1e-8 (population std, as sklearn)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-5)
    assert np.all(got[:, 3] == 0)   # a constant feature maps to 0, not NaN


def test_minmax_matches_sklearn():
    from sklearn.preprocessing import MinMaxScaler

    X = _data(2)
    got = normalize_input(X, "minmax")
    assert got.dtype == np.float32
    want = MinMaxScaler().fit_transform(X.astype(np.float64))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
    assert got.min() == 0.0 and got.max() <= 1.0
    assert np.all(got[:, 3] == 0)


def test_minmax_formula_on_a_hand_case():
    X = np.array([[0.0, 10.0], [2.0, 30.0], [1.0, 20.0]], dtype=np.float32)
    got = normalize_input(X, "minmax")
    np.testing.assert_allclose(got, [[0.0, 0.0], [1.0, 1.0], [0.5, 0.5]], atol=1e-7)


@pytest.mark.parametrize("bad", ["zscore", "MinMax", "l2", 3])
def test_unknown_method_raises_value_error(bad):
    with pytest.raises(ValueError, match="Unknown normalize method"):
        normalize_input(_data(), bad)


def test_umap_rejects_unknown_normalize_before_device_work():
    with pytest.raises(ValueError, match="Unknown normalize method"):
        UMAP(normalize="bogus").fit_transform(_data())
