"""The query entry point that allocates nothing (tk_knn_query_gpu_noalloc), called through the C
ABI: its results equal tk_knn_query_gpu's bit for bit, on two calls in a row (the second with a
smaller input, so it runs on scratch the first call left behind), and a query count above its
limit fails with an error instead of running."""
import ctypes
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "torch_kdtree_amd", "libtorch_kdtree_amd.so")
MAX_Q = 2097152  # TK_NOALLOC_MAX_QUERIES


def _lib():
    lib = ctypes.CDLL(LIB)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    lib.tk_last_error.restype = ctypes.c_char_p
    lib.tk_tree_build.argtypes = [vp, i64, i32, i32, i32, i32, ctypes.POINTER(vp)]
    lib.tk_tree_upload.argtypes = [vp, i32]
    lib.tk_tree_free.argtypes = [vp]
    lib.tk_tree_free.restype = None
    lib.tk_knn_query_gpu.argtypes = [vp, vp, i64, i32, vp, vp, vp, i32, i32, vp, ctypes.c_size_t, vp]
    lib.tk_knn_query_gpu_noalloc.argtypes = [vp, vp, i64, i32, vp, vp, vp, i32, i32, vp]
    return lib


def _check(lib, rc):
    assert rc == 0, lib.tk_last_error().decode()


def _outputs(nq, k, dtype, dev):
    return (torch.full((nq, k), -1.0, dtype=dtype, device=dev), torch.full((nq, k), -1, dtype=torch.int32, device=dev),
            torch.full((nq, k), -1, dtype=torch.int64, device=dev))


@pytest.mark.parametrize("dtype,tie,k", [(torch.float32, 0, 16), (torch.float32, 1, 8), (torch.float64, 0, 40)])
def test_noalloc_query_matches_query(cuda, dtype, tie, k):
    lib = _lib()
    rng = np.random.default_rng(5 + k)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    ref = rng.random((300_000, 3)).astype(npdt)
    tree = ctypes.c_void_p()
    _check(lib, lib.tk_tree_build(ref.ctypes.data, ref.shape[0], 3, 0 if npdt == np.float32 else 1, 13, 0,
                                  ctypes.byref(tree)))
    try:
        _check(lib, lib.tk_tree_upload(tree, cuda.index or 0))
        stream = torch.cuda.current_stream(cuda).cuda_stream
        for nq in (MAX_Q, 150_000, 3000):  # full scratch, then less, then no presort at all
            q = torch.from_numpy((rng.random((nq, 3)) * 1.2 - 0.1).astype(npdt)).to(cuda)
            d0, i0, o0 = _outputs(nq, k, dtype, cuda)
            d1, i1, o1 = _outputs(nq, k, dtype, cuda)
            _check(lib, lib.tk_knn_query_gpu(tree, q.data_ptr(), nq, k, d0.data_ptr(), i0.data_ptr(), o0.data_ptr(),
                                             tie, -1, None, 0, stream))
            _check(lib, lib.tk_knn_query_gpu_noalloc(tree, q.data_ptr(), nq, k, d1.data_ptr(), i1.data_ptr(),
                                                     o1.data_ptr(), tie, -1, stream))
            torch.cuda.synchronize()
            assert torch.equal(d0, d1), "nq %d distances" % nq
            assert torch.equal(i0, i1), "nq %d structured indices" % nq
            assert torch.equal(o0, o1), "nq %d original indices" % nq
            assert int((i1 < 0).sum()) == 0
        # above the limit: an error before anything runs (buffers sized so nothing could overflow)
        nq = MAX_Q + 1
        q = torch.zeros((nq, 3), dtype=dtype, device=cuda)
        d1, i1, o1 = _outputs(nq, k, dtype, cuda)
        rc = lib.tk_knn_query_gpu_noalloc(tree, q.data_ptr(), nq, k, d1.data_ptr(), i1.data_ptr(), o1.data_ptr(),
                                          tie, -1, stream)
        torch.cuda.synchronize()
        assert rc != 0 and b"exceed the limit" in lib.tk_last_error()
        assert int((i1 != -1).sum()) == 0, "nothing may have been written"
    finally:
        lib.tk_tree_free(tree)
