"""ctypes binding of libmivis.so (declared in include/mivis.h).

Every entry point returns a status; a non-zero status becomes ``ValueError``
(invalid argument) or :class:`MivisError` carrying ``mv_last_error()``.
There is no fallback: if the library is missing, :func:`lib` raises.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmivis.so"
ABI_VERSION = 1

MV_OK = 0
MV_EINVAL = -1
MV_EHIP = -2
MV_EWORKSPACE = -3
MV_EUNSUPPORTED = -4
MV_EINJECTED = -5

DTYPES = {"fp32": 0, "bf16": 1, "fp8": 2, "bf16x3": 3}


class MivisError(RuntimeError):
    """A libmivis call failed; ``status`` is the negative mv_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"libmivis error {status}: {message}")
        self.status = status


_lock = threading.Lock()
_lib = None

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_f32 = C.c_float
_u64 = C.c_uint64
_u32 = C.c_uint32
_sz = C.c_size_t

# name -> argtypes (restype is always int unless listed in _RESTYPES)
_SIGNATURES: dict[str, list] = {
    "mv_abi_version": [],
    "mv_last_error": [],
    "mv_device_info": [_p, _p, _p],
    "mv_knn_ws_bytes": [_i64, _i64, _i64, _i32, _i32, _i32, _i32, C.POINTER(_sz)],
    "mv_knn_brute": [_p, _i64, _p, _i64, _i64, _i32, _i32, _i32, _p, _i64, _i32, _p, _p, _p, _p, _sz, _p],
    "mv_knn_brute_noalloc": [_p, _i64, _p, _i64, _i64, _i32, _i32, _i32, _p, _i64, _i32, _p, _p, _p, _p],
    "mv_row_sqnorms": [_p, _i64, _i64, _p, _p],
    "mv_knn_set_timing": [_i32],
    "mv_knn_set_symmetric": [_i32],
    "mv_knn_last_screen_ms": [_p, _p],
    "mv_knn_last_phase_ms": [_p],
    "mv_umap_graph_ws_bytes": [_i64, _i32, C.POINTER(_sz)],
    "mv_umap_fuzzy_graph": [_p, _p, _i64, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _sz, _p],
    "mv_umap_in_edges": [_p, _p, _i64, _i32, _p, _p, _p, _sz, _p],
    "mv_umap_epochs_ws_bytes": [_i64, _i32, _i32, C.POINTER(_sz)],
    "mv_umap_epochs": [_p, _i64, _i32, _p, _i32, _p, _p, _p, _p, _i64, _i32, _i32, _i32, _i32, _f32, _f32, _f32,
                       _u64, _p, _sz, _p],
    "mv_umap_sgd_step": [_p, _p, _i64, _i32, _p, _p, _i64, _p, _p, _i64, _f32, _f32, _f32, _p],
    "mv_zero_f32": [_p, _i64, _p],
    "mv_nnd_ws_bytes": [_i64, _i64, _i32, _i64, _i64, C.POINTER(_sz)],
    "mv_nnd_build": [_p, _i64, _i64, _i32, _i32, _i32, _f32, _p, _p, _i64, _p, _i64, _i32, _p, _p, _p, _p, _p,
                     _sz, _p],
    "mv_nnd_reverse": [_p, _i64, _i32, _p, _p, _sz, _p],
    "mv_nnd_gather_dists": [_p, _i64, _i64, _p, _i32, _i32, _p, _p, _sz, _p],
    "mv_umap_spectral_ws_bytes": [_i64, _i32, _i32, C.POINTER(_sz)],
    "mv_umap_spectral": [_p, _p, _p, _i64, _i32, _i32, _i32, _p, _p, _sz, _p],
    "mv_umap_epochs_masked": [_p, _i64, _i32, _p, _i32, _p, _p, _p, _p, _i64, _i32, _i32, _i32, _i32, _f32, _f32,
                              _f32, _u64, _p, _p, _sz, _p],
    "mv_mlp_param_bytes": [_i32, _i32, _i32, _i32, C.POINTER(_sz), C.POINTER(_i64)],
    "mv_mlp_batch_bytes": [_i32, _i32, _i32, _i32, _i32, C.POINTER(_sz)],
    "mv_mlp_init": [_p, _i32, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _p],
    "mv_mlp_grad": [_p, _i32, _i32, _i32, _i32, _p, _i64, _p, _p, _i32, _i32, _f32, _f32, _f32, _p, _p, _p, _sz, _p],
    "mv_mlp_train": [_p, _i32, _i32, _i32, _i32, _p, _i64, _p, _p, _i32, _i32, _i32, _f32, _f32, _f32, _f32, _i32,
                     _p, _p, _sz, _p],
    "mv_mlp_forward": [_p, _i32, _i32, _i32, _i32, _p, _i64, _p, _i32, _p, _sz, _p],
    "mv_mlp_sample_edges": [_p, _p, _i64, _i32, _u32, _i32, _i32, _i32, _p, _p],
    "mv_knn_overlap": [_p, _i64, _i32, _p, _i64, _i32, _i64, _i32, _p, _p, _p],
    "mv_trust_ws_bytes": [_i64, _i32, C.POINTER(_sz)],
    "mv_trust_excess": [_p, _i64, _i64, _p, _i64, _p, _i64, _i32, _p, _p, _sz, _p],
    "mv_random_normal": [_p, _i64, _u64, _p],
    "mv_umap_spectral_embed": [_p, _i64, _i32, _u64, _p, _p, _p, _sz, _p],
    "mv_cne_grad": [_p, _i64, _i32, _p, _p, _i64, _p, _i32, _i32, _f32, _u64, _i32, _i64, _p, _p],
    "mv_cne_optimize": [_p, _i64, _i32, _p, _p, _i64, _i32, _i32, _i32, _i32, _f32, _i64, _u64, _p, _p, _p, _p],
    "mv_pair_grad": [_p, _i64, _i32, _p, _p, _i64, _p, _f32, _f32, _i32, _p, _p],
    "mv_adam_update": [_p, _p, _p, _p, _i64, _f32, _f32, _f32, _f32, _p],
    "mv_pca_ws_bytes": [_i64, _i64, C.POINTER(_sz)],
    "mv_pca_cov": [_p, _i64, _i64, _p, _p, _p, _sz, _p],
    "mv_pca_project": [_p, _i64, _i64, _p, _p, _i64, _p, _p],
}
_RESTYPES = {"mv_last_error": C.c_char_p}


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def lib():
    """Loads libmivis.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"libmivis.so not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, C.c_int)
        ver = handle.mv_abi_version()
        if ver != ABI_VERSION:
            raise RuntimeError(f"libmivis ABI version {ver} != expected {ABI_VERSION}")
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().mv_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(status: int, what: str = "") -> None:
    if status == MV_OK:
        return
    msg = last_error()
    if what:
        msg = f"{what}: {msg}"
    if status == MV_EINVAL:
        raise ValueError(msg)
    raise MivisError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    return None if t is None else t.data_ptr()


def stream_handle(torch_stream=None) -> int | None:
    import torch

    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return s.cuda_stream or None
