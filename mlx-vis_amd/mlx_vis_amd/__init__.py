"""mlx_vis_amd: the mlx-vis hot path (exact kNN, NNDescent, UMAP layout) on
MI355X (gfx950) through hand-written HIP kernels in libmivis.so.

Drop-in names follow mlx_vis/__init__.py:5-17 for this path:
``UMAP``, ``NNDescent`` and the shared ``compute_knn`` operator, plus ``CNE``,
``PaCMAP`` and ``LocalMAP`` (layouts on the same edge-gradient kernels) and
the learned form ``ParametricUMAP``, and the metrics ``trustworthiness`` and
``knn_preservation`` (exact neighbour lists and ranks on the GPU).
"""

__version__ = "0.1.0"

from ._knn import compute_knn, knn_brute_device
from ._normalize import normalize_input
from ._nndescent import NNDescent
from ._umap import UMAP, find_ab_params
from ._parametric import ParametricUMAP, knn_preservation_device
from ._cne import CNE
from ._pacmap import LocalMAP, PaCMAP
from ._metrics import knn_preservation, trust_excess_device, trustworthiness

__all__ = ["UMAP", "NNDescent", "ParametricUMAP", "CNE", "PaCMAP", "LocalMAP", "compute_knn", "knn_brute_device", "knn_preservation_device",
           "normalize_input", "find_ab_params", "trustworthiness", "knn_preservation", "trust_excess_device"]
