"""GPU parity of the operator-construction entry points (scl_kernel_*,
scl_transition_*) against the CPU restatement (oracle/construct.c).  The
transition functions are bit-identical, the stationary distribution
included (same gather order, the eight-partial vector_sum on the device);
kernel weights are exact for the algebraic kernels and within a few ulp
where exp or cos is evaluated (device libm vs host libm)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from scl_amd import CscMatrix, CsrMatrix, construct, synth
from scl_amd._lib import SCLError

pytestmark = pytest.mark.gpu


def distances(n, k, seed):
    g = synth.uniform_knn(n, k, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.gamma(2.0, 0.6, len(g.data)).astype(np.float32)
    return g.indptr, g.indices, d


@pytest.mark.parametrize("kind", list(construct.KERNELS))
def test_kernel_weights(gpu, kind):
    ip, ix, d = distances(20_000, 12, seed=31)
    D = CsrMatrix.from_arrays(ip, ix, d, (20_000, 20_000))
    k = construct.KERNELS[kind]
    for h in (0.8, 2.5):
        got = construct.kernel_matrix(D, h, kind)
        ref = orc.kernel_matrix_f32(d, h, k)
        sums = construct.kernel_row_sums(D, h, kind)
        rsum = orc.kernel_row_sums_f32(ip, d, h, k)
        if kind in ("epanechnikov", "cauchy", "uniform", "triangular"):
            assert np.array_equal(got, ref) and np.array_equal(sums, rsum), (kind, h)
        else:
            # exp / cos: device and host libm may differ in the last bits
            assert np.abs(got - ref).max() <= 4e-7 * max(1.0, np.abs(ref).max()), (kind, h)
            assert np.allclose(sums, rsum, rtol=2e-6, atol=1e-7), (kind, h)


def test_mid_size_host_copies(gpu):
    """Host arrays of 8 to 64 MiB cross through the pinned bounce buffers in
    chunks of a quarter of the copy (at least 4 MiB): 3,000,027 values
    (12 MB in, 12 MB out; chunks of 4, 4 and 3.4 MB) come back exactly."""
    n, k = 300_001, 10
    g = synth.uniform_knn(n, k, seed=41)
    rng = np.random.Generator(np.random.PCG64(41))
    d = rng.gamma(2.0, 0.6, len(g.data)).astype(np.float32)
    ip = g.indptr.copy()
    ix = g.indices.copy()
    # 17 more entries on the last row: the copy is not a multiple of 4 KiB
    extra = np.setdiff1d(np.arange(40), ix[ip[-2]:ip[-1]])[:17]
    ix = np.concatenate([ix, extra]).astype(np.int64)
    d = np.concatenate([d, rng.gamma(2.0, 0.6, 17).astype(np.float32)])
    ip[-1] += 17
    order = np.argsort(ix[ip[-2]:], kind="stable") + ip[-2]
    ix[ip[-2]:], d[ip[-2]:] = ix[order], d[order]
    D = CsrMatrix.from_arrays(ip, ix, d, (n, n))
    assert D.nnz == len(d) == 3_000_027
    got = construct.kernel_matrix(D, 1.3, "epanechnikov")
    assert np.array_equal(got, orc.kernel_matrix_f32(d, 1.3, 1))
    assert np.array_equal(construct.row_normalize(D), orc.row_normalize_f32(ip, d))


def test_row_normalize(gpu):
    ip, ix, d = distances(30_000, 10, seed=32)
    d[ip[5]:ip[6]] = 0.0                     # a zero row
    M = CsrMatrix.from_arrays(ip, ix, d, (30_000, 30_000))
    assert np.array_equal(construct.row_normalize(M), orc.row_normalize_f32(ip, d))
    part = construct.row_normalize(M, 1000)
    assert np.array_equal(part, orc.row_normalize_f32(ip[:1001], d[:ip[1000]]))


def test_symmetrize_and_velocity(gpu):
    n = 3000
    g = synth.uniform_knn(n, 9, seed=33)
    T = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    assert np.array_equal(construct.symmetrize(T),
                          orc.symmetrize_f32(g.indptr, g.indices, g.data))
    # rows with repeated and descending indices and self-loops: stored order
    ip = np.array([0, 4, 6, 9, 10], np.int64)
    ix = np.array([2, 2, 0, 1, 3, 0, 2, 2, 1, 1], np.int64)
    v = np.array([0.3, 0.2, 0.5, 1e-3, 1, 2, 0.7, 0.1, 0.2, 4], np.float32)
    U = CsrMatrix.from_arrays(ip, ix, v, (4, 4))
    assert np.array_equal(construct.symmetrize(U), orc.symmetrize_f32(ip, ix, v))
    # velocity graph with signed weights
    rng = np.random.Generator(np.random.PCG64(34))
    sv = rng.standard_normal(len(g.data)).astype(np.float32)
    sv[g.indptr[11]:g.indptr[12]] = -1.0
    V = CsrMatrix.from_arrays(g.indptr, g.indices, sv, (n, n))
    got = construct.matrix_from_velocity(V)
    assert np.array_equal(got, orc.from_velocity_f32(g.indptr, g.indices, sv))
    assert got[11, 11] == 1.0
    assert np.array_equal(construct.matrix_from_velocity(U), orc.from_velocity_f32(ip, ix, v))


def test_stationary_distribution(gpu):
    for g, tol in ((synth.knn_graph(50_000, 10, seed=35).normalized(), 1e-9),
                   (synth.uniform_knn(40_000, 8, seed=36).normalized(), 1e-6),
                   (synth.uniform_knn(2000, 5, seed=37).normalized(), 1e-4)):
        T = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n))
        for max_iter in (0, 1, 7, 1000):
            got = construct.stationary_distribution(T, max_iter, tol)
            ref, _ = orc.stationary_f32(g.indptr, g.indices, g.data, max_iter, tol)
            assert np.array_equal(got, ref), (g.n, tol, max_iter)


def test_construct_argument_errors(gpu):
    g = synth.uniform_knn(100, 4, seed=38)
    T = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (100, 100))
    from scl_amd import _lib
    lib = _lib.load()
    out = np.zeros(100 * 100, np.float32)
    assert lib.scl_transition_symmetrize(T.handle, 99, out.ctypes.data) == 11
    assert lib.scl_transition_stationary_distribution(T.handle, 101, out.ctypes.data, 5,
                                                      1e-6) == 11
    assert lib.scl_transition_row_normalize_to_stochastic(T.handle, 100, out.ctypes.data,
                                                          10) == 11
    assert lib.scl_kernel_row_sums(T.handle, out.ctypes.data, 99, 1.0, 0) == 11
    with pytest.raises(SCLError) as e:
        construct.row_normalize(T, 101)
    assert e.value.code == 11


def upper_distances(n):
    """The distance fixture of test_kernel.cpp:24-64: row i stores columns
    i..n-1 with distance j - i (0 on the diagonal)."""
    ip, ix, d = [0], [], []
    for i in range(n):
        for j in range(i, n):
            ix.append(j)
            d.append(float(j - i))
        ip.append(len(ix))
    return np.array(ip, np.int64), np.array(ix, np.int64), np.array(d, np.float32)


def test_kernel_reference_cases(gpu):
    # test_kernel.cpp:347-366 (basic: values >= 0), 369-385 (K(0) = 1 for the
    # Gaussian), 416-432 (row sums > 0), 435-453 (four kernel kinds succeed)
    for n in (5, 3, 4):
        ip, ix, d = upper_distances(n)
        D = CsrMatrix.from_arrays(ip, ix, d, (n, n))
        vals = construct.kernel_matrix(D, 1.0, "gaussian")
        assert np.all(vals >= 0.0)
        assert vals[0] == 1.0 and np.all(vals[ip[:-1]] == 1.0)   # every diagonal entry
        assert np.allclose(vals, np.exp(-0.5 * d.astype(np.float64) ** 2), rtol=1e-6)
        assert np.abs(vals - orc.kernel_matrix_f32(d, 1.0, 0)).max() <= 1e-7
        sums = construct.kernel_row_sums(D, 1.0, "gaussian")
        assert np.all(sums > 0.0)
        for kind in ("gaussian", "epanechnikov", "cosine", "linear"):
            k = construct.KERNELS[kind]
            got = construct.kernel_row_sums(D, 1.0, kind)
            ref = orc.kernel_row_sums_f32(ip, d, 1.0, k)
            assert np.allclose(got, ref, rtol=2e-6, atol=1e-7), kind
    # closed forms at distance 0 and past the bandwidth
    ip, ix, d = upper_distances(4)
    D = CsrMatrix.from_arrays(ip, ix, d, (4, 4))
    assert np.all(construct.kernel_matrix(D, 1.0, "epanechnikov")[d == 0] == 0.75)
    assert np.all(construct.kernel_matrix(D, 1.0, "epanechnikov")[d > 1] == 0.0)
    assert np.all(construct.kernel_matrix(D, 1.0, "uniform")[d == 0] == 0.5)
    assert np.all(construct.kernel_matrix(D, 1.0, "triangular")[d >= 1] == 0.0)
    assert np.allclose(construct.kernel_matrix(D, 1.0, "cosine")[d == 0], np.pi / 4, rtol=1e-6)
    assert np.array_equal(construct.kernel_matrix(D, 2.0, "cauchy"),
                          orc.kernel_matrix_f32(d, 2.0, 6))
