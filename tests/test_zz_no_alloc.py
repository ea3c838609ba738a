"""The allocation-free entry point scl_gpu_diffuse_matrix_device_noalloc,
called through the C ABI: bit-identical to scl_gpu_diffuse_matrix_device (the
same kernels with the caller's work buffer) on two calls in a row, the second
on a smaller operator and matrix, so the shared scratch keeps nothing from
the first; sizes beyond the scratch and an operator not yet on the device
are refused with a clear error."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from scl_amd import CsrMatrix, _lib, synth

pytestmark = pytest.mark.gpu

SCRATCH_FLOATS = 16 * 1024 * 1024


def _run_both(P, X, t):
    lib = _lib.load()
    n, F = X.shape
    s = torch.cuda.current_stream().cuda_stream
    Xd = torch.from_numpy(X).cuda()
    Yr, W = torch.empty_like(Xd), torch.empty_like(Xd)
    _lib.check(lib.scl_gpu_diffuse_matrix_device(P.handle, Xd.data_ptr(), Yr.data_ptr(),
                                                 W.data_ptr(), n, F, t, s))
    Yn = torch.full_like(Xd, float("nan"))
    _lib.check(lib.scl_gpu_diffuse_matrix_device_noalloc(P.handle, Xd.data_ptr(),
                                                         Yn.data_ptr(), n, F, t, s))
    torch.cuda.synchronize()
    return Yr.cpu().numpy(), Yn.cpu().numpy()


def test_noalloc_matches_device_entry_point(gpu):
    cases = [(30_000, 12, 96, 5), (4_000, 7, 20, 4)]   # second call smaller
    for n, k, F, t in cases:
        g = synth.knn_graph(n, k, seed=n).normalized()
        P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
        P.prepare_device()
        X = synth.normal(n, F, seed=F)
        ref, got = _run_both(P, X, t)
        print(f"noalloc n={n} F={F} t={t}: bit-identical={np.array_equal(ref, got)}", flush=True)
        assert np.array_equal(ref, got), (n, F, t)


def test_noalloc_refusals(gpu):
    lib = _lib.load()
    s = torch.cuda.current_stream().cuda_stream
    n = 2000
    g = synth.uniform_knn(n, 5, seed=3).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd = torch.zeros((n, 8), device="cuda")
    Yd = torch.empty_like(Xd)
    # not on the device yet
    rc = lib.scl_gpu_diffuse_matrix_device_noalloc(P.handle, Xd.data_ptr(), Yd.data_ptr(),
                                                   n, 8, 2, s)
    assert rc == _lib.SCL_ERROR_INVALID_ARGUMENT
    assert b"scl_gpu_sparse_prepare" in lib.scl_get_last_error()
    P.prepare_device()
    # n * F beyond the scratch for t >= 2 (checked before any device work)
    F_big = SCRATCH_FLOATS // n + 1
    rc = lib.scl_gpu_diffuse_matrix_device_noalloc(P.handle, Xd.data_ptr(), Yd.data_ptr(),
                                                   n, F_big, 2, s)
    assert rc == _lib.SCL_ERROR_INVALID_ARGUMENT
    assert b"scratch" in lib.scl_get_last_error()
    # t = 1 needs no scratch
    X = synth.normal(n, 8, seed=4)
    ref, got = _run_both(P, X, 1)
    assert np.array_equal(ref, got)
