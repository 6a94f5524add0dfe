"""The C-ABI library loads, exports every declared symbol, and maps errors
(CPU only: no call here reaches the GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "mivis.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mv_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from mlx_vis_amd import _capi

    lib = _capi.lib()
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in mivis.h but not exported"
    assert set(declared) == set(_capi.exported_symbols())
    assert lib.mv_abi_version() == _capi.ABI_VERSION


def test_ws_query_and_invalid_arguments_map_to_valueerror():
    from mlx_vis_amd import _capi

    lib = _capi.lib()
    nbytes = ctypes.c_size_t(0)
    _capi.check(lib.mv_knn_ws_bytes(1000, 1000, 128, 15, 0, 2, 1, ctypes.byref(nbytes)))
    assert nbytes.value > 1000 * 128 * 4
    with pytest.raises(ValueError, match="k=0"):
        _capi.check(lib.mv_knn_ws_bytes(10, 10, 4, 0, 0, 2, 1, ctypes.byref(nbytes)))
    # Invalid arguments are rejected before any device work.
    with pytest.raises(ValueError, match="null pointer"):
        _capi.check(lib.mv_knn_brute(None, 10, None, 10, 4, 3, 0, 2, None, 0, 1,
                                     None, None, None, None, 0, None))
    with pytest.raises(_capi.MivisError) as ei:
        _capi.check(lib.mv_knn_ws_bytes(10, 10, 4, 3, 7, 2, 1, ctypes.byref(nbytes)))
    assert ei.value.status == _capi.MV_EUNSUPPORTED


def test_symmetric_screen_workspace_and_toggle():
    """The symmetric self-kNN screen adds the n_pad^2 FP32 matrix (and block
    minima) to the workspace, only for self kNN within its size range; the
    toggle returns its previous setting."""
    from mlx_vis_amd import _capi

    lib = _capi.lib()

    def ws(nq, n, same):
        nb = ctypes.c_size_t(0)
        _capi.check(lib.mv_knn_ws_bytes(nq, n, 784, 15, 0, 2, same, ctypes.byref(nb)))
        return nb.value

    n_pad = 70_016
    on = ws(70_000, 70_000, 1)
    assert lib.mv_knn_set_symmetric(0) == 1
    try:
        off = ws(70_000, 70_000, 1)
        assert lib.mv_knn_set_symmetric(0) == 0
    finally:
        assert lib.mv_knn_set_symmetric(1) == 0
    assert on - off >= n_pad * n_pad * 4 + n_pad * (n_pad // 128) * 4
    assert ws(70_000, 70_000, 0) < n_pad * n_pad * 4           # queries are not the data
    assert ws(4_000, 4_000, 1) < 4_096 * 4_096 * 4              # below 4,096 rows: fused screen
    assert ws(200_000, 200_000, 1) < 200_064 * 200_064 * 4      # beyond 48 GiB: fused screen


def test_fault_injection_maps_to_mivis_error(monkeypatch):
    from mlx_vis_amd import _capi

    monkeypatch.setenv("MIVIS_FAULT", "launch")
    with pytest.raises(_capi.MivisError) as ei:
        _capi.check(_capi.lib().mv_row_sqnorms(ctypes.c_void_p(16), 4, 4, ctypes.c_void_p(16), None))
    assert ei.value.status == _capi.MV_EINJECTED
    assert "MIVIS_FAULT" in str(ei.value)


def test_device_entry_points_reject_host_tensors():
    """The kernels take raw device pointers: host tensors must be refused
    before any launch (a host pointer in a kernel faults the GPU)."""
    import torch

    from mlx_vis_amd import NNDescent, knn_brute_device, knn_preservation_device
    from mlx_vis_amd._umap import fuzzy_graph_device, sgd_step_device

    X = torch.zeros((16, 4))
    idx = torch.zeros((16, 3), dtype=torch.int32)
    with pytest.raises(ValueError, match="device"):
        knn_brute_device(X, 3)
    with pytest.raises(ValueError, match="device"):
        NNDescent(k=3).build_device(X)
    with pytest.raises(ValueError, match="device"):
        knn_preservation_device(idx, idx)
    with pytest.raises(ValueError, match="device"):
        fuzzy_graph_device(idx, torch.zeros((16, 3)), 10)
    with pytest.raises(ValueError, match="device"):
        sgd_step_device(torch.zeros((16, 2)), idx[:, 0], idx[:, 1], idx[:, 0], idx[:, 1], 1.0, 1.5, 0.9)
