"""mv_knn_brute_noalloc (the kNN entry point that allocates nothing: its
workspace is a static device array of the library) against mv_knn_brute with
a caller workspace, through the C ABI: two calls in a row, the second with a
smaller input, bit-identical to mv_knn_brute and to the oracle; a request
beyond the static scratch is refused with a clear error."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import knn as oknn

pytestmark = pytest.mark.gpu


def _call(fn_name, Q, X, k, dtype, self_base, q_self=None):
    from mlx_vis_amd import _capi, _device

    lib = _capi.lib()
    nq, d = Q.shape
    n = X.shape[0]
    idx = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    stats = torch.zeros(4, dtype=torch.int32, device="cuda")
    args = (Q.data_ptr(), nq, X.data_ptr(), n, d, k, _capi.DTYPES[dtype], 2, _capi.ptr(q_self), self_base, 1,
            idx.data_ptr(), dist.data_ptr(), stats.data_ptr())
    if fn_name == "mv_knn_brute":
        nb = ctypes.c_size_t(0)
        _capi.check(lib.mv_knn_ws_bytes(nq, n, d, k, _capi.DTYPES[dtype], 2, int(Q.data_ptr() == X.data_ptr()),
                                        ctypes.byref(nb)))
        ws = _device.workspace(nb.value)
        _capi.check(lib.mv_knn_brute(*args, ws.data_ptr(), ws.numel(), None), fn_name)
    else:
        _capi.check(lib.mv_knn_brute_noalloc(*args, None), fn_name)
    torch.cuda.synchronize()
    return idx.cpu().numpy(), dist.cpu().numpy(), stats.cpu().numpy()


def test_noalloc_matches_workspace_entry_point_on_two_calls():
    rng = np.random.default_rng(31)
    # call 1: self kNN 6,000 x 200, FP32 screen
    X1 = rng.standard_normal((6000, 200)).astype(np.float32)
    X1[3000:3040] = X1[:40]                                     # exact ties
    X1d = torch.from_numpy(X1).cuda()
    print("case noalloc call 1: self 6000 x 200 fp32 k=15", flush=True)
    a_i, a_d, a_s = _call("mv_knn_brute_noalloc", X1d, X1d, 15, "fp32", 0)
    b_i, b_d, _ = _call("mv_knn_brute", X1d, X1d, 15, "fp32", 0)
    np.testing.assert_array_equal(a_i, b_i)
    np.testing.assert_array_equal(a_d, b_d)
    oi, od = oknn.exact_knn(X1, X1, 15, self_base=0)
    np.testing.assert_array_equal(a_i, oi)
    np.testing.assert_allclose(a_d, np.sqrt(od), rtol=1e-5, atol=1e-6)
    # call 2, smaller: 1,000 queries against 3,000 x 64 rows, BF16 screen, k = 10
    X2 = (rng.standard_normal((3000, 64)) * 2 + 1).astype(np.float32)
    Q2 = (rng.standard_normal((1000, 64)) * 2 + 1).astype(np.float32)
    X2d, Q2d = torch.from_numpy(X2).cuda(), torch.from_numpy(Q2).cuda()
    print("case noalloc call 2: 1000 queries x 3000 x 64 bf16 k=10", flush=True)
    a_i, a_d, a_s = _call("mv_knn_brute_noalloc", Q2d, X2d, 10, "bf16", -1)
    b_i, b_d, b_s = _call("mv_knn_brute", Q2d, X2d, 10, "bf16", -1)
    np.testing.assert_array_equal(a_i, b_i)
    np.testing.assert_array_equal(a_d, b_d)
    np.testing.assert_array_equal(a_s, b_s)
    oi, od = oknn.exact_knn(Q2, X2, 10)
    np.testing.assert_array_equal(a_i, oi)
    np.testing.assert_allclose(a_d, np.sqrt(od), rtol=1e-5, atol=1e-6)


def test_noalloc_refuses_requests_beyond_its_scratch():
    from mlx_vis_amd import _capi

    X = torch.zeros((100_000, 784), dtype=torch.float32, device="cuda")
    idx = torch.empty((100_000, 15), dtype=torch.int32, device="cuda")
    dist = torch.empty((100_000, 15), dtype=torch.float32, device="cuda")
    rc = _capi.lib().mv_knn_brute_noalloc(X.data_ptr(), 100_000, X.data_ptr(), 100_000, 784, 15, 0, 2, None, 0, 1,
                                         idx.data_ptr(), dist.data_ptr(), None, None)
    assert rc == _capi.MV_EUNSUPPORTED
    msg = _capi.last_error()
    assert "static scratch" in msg and "mv_knn_brute" in msg, msg
