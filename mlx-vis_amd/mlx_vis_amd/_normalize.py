"""Input normalization (host NumPy, as in the reference).

Same contract as ``normalize_input`` — mlx_vis/_normalize.py:6-33:
False/None passes through, True/"standard" z-scores each feature
(std + 1e-8), "minmax" maps each feature to [0, 1] with
(X - min) / (max - min + 1e-8); anything else raises ValueError.
"""

from __future__ import annotations

import numpy as np


def normalize_input(X: np.ndarray, method) -> np.ndarray:
    if not method:
        return X
    if method is True or method == "standard":
        mean = X.mean(axis=0)
        std = X.std(axis=0) + 1e-8
        return (X - mean) / std
    if method == "minmax":
        lo = X.min(axis=0)
        hi = X.max(axis=0)
        return (X - lo) / (hi - lo + 1e-8)
    raise ValueError(f"Unknown normalize method: {method}")
