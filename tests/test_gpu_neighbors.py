"""GPU parity of the exact k-nearest-neighbour search (scl_knn,
scl_knn_simple, scl_neighbors_compute_norms) against the CPU restatement of
the reference's heap search (oracle/neighbors.c).  Norms, distances and
index lists are bit-identical as returned: the GPU offers candidates to each
query's heap in index order with the reference's pruning test, heap steps
and final heap sort, so equal distances come out in the reference's order
and ties at the k-th distance keep the reference's members."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from scl_amd import CscMatrix, CsrMatrix, construct
from scl_amd._lib import SCLError

pytestmark = pytest.mark.gpu


def samples(n, f, density, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    ip, ix = [0], []
    for _ in range(n):
        cols = np.flatnonzero(rng.random(f) < density)
        ix.extend(cols)
        ip.append(len(ix))
    v = rng.gamma(1.0, 1.0, len(ix)).astype(np.float32)
    return np.array(ip, np.int64), np.array(ix, np.int64), v


@pytest.mark.parametrize("n,f,k", [(3000, 150, 15), (1000, 37, 1), (777, 64, 70),
                                   (130, 20, 200)])
def test_knn_matches_restatement(gpu, n, f, k):
    ip, ix, v = samples(n, f, 0.15, seed=n + f)
    M = CsrMatrix.from_arrays(ip, ix, v, (n, f))
    norms = construct.neighbor_norms(M)
    assert np.array_equal(norms, orc.norms_f32(ip, v))
    ri, rd = orc.knn_f32(ip, ix, v, k)
    for gi, gd in (construct.knn(M, k), construct.knn(M, k, norms)):
        assert np.array_equal(gd, rd), (n, f, k)
        assert np.array_equal(gi, ri), (n, f, k)
    if k >= n:
        assert np.all(gi[:, n - 1:] == -1) and np.all(np.isinf(gd[:, n - 1:]))


def test_knn_samples_of_a_csc_handle_and_ties(gpu):
    n, f = 500, 40
    ip, ix, v = samples(n, f, 0.2, seed=5)
    C = CscMatrix.from_arrays(ip, ix, v, (f, n))     # samples are the columns
    gi, gd = construct.knn(C, 12)
    ri, rd = orc.knn_f32(ip, ix, v, 12)
    assert np.array_equal(gd, rd) and np.array_equal(gi, ri)
    # four identical samples: zero distances among them, in the reference's
    # heap order
    ip = np.array([0, 2, 4, 6, 8, 11], np.int64)
    ix = np.array([0, 2, 0, 2, 0, 2, 0, 2, 0, 1, 2], np.int64)
    v = np.array([1, 2, 1, 2, 1, 2, 1, 2, 9, 9, 9], np.float32)
    gi, gd = construct.knn(CsrMatrix.from_arrays(ip, ix, v, (5, 3)), 3)
    ri, rd = orc.knn_f32(ip, ix, v, 3)
    assert np.array_equal(gi, ri) and np.array_equal(gd, rd)
    assert sorted(gi[0]) == [1, 2, 3] and np.all(gd[:4] == 0)
    assert np.allclose(gd[4], np.sqrt(64 + 81 + 49), rtol=1e-6)


@pytest.mark.parametrize("copies,k", [(10, 7), (10, 15), (37, 20), (3, 4), (200, 130)])
def test_knn_ties_at_the_kth_distance_keep_reference_members(gpu, copies, k):
    """Samples that are copies of a few distinct vectors (and integer-valued
    ones whose distances repeat), so that many candidates share the k-th
    distance: the members kept and their order are the reference heap's, as
    returned, index for index, across several candidate tiles."""
    rng = np.random.Generator(np.random.PCG64(copies + k))
    base = (rng.integers(0, 3, (6, 12)) * (rng.random((6, 12)) < 0.5)).astype(np.float32)
    dense = np.repeat(base, copies, axis=0)
    dense = dense[rng.permutation(len(dense))]
    n = len(dense)
    ip = np.concatenate([[0], np.cumsum((dense != 0).sum(1))]).astype(np.int64)
    ix = np.concatenate([np.flatnonzero(r) for r in dense]).astype(np.int64)
    v = dense[dense != 0].astype(np.float32)
    M = CsrMatrix.from_arrays(ip, ix, v, (n, 12))
    ri, rd = orc.knn_f32(ip, ix, v, k)
    gi, gd = construct.knn(M, k)
    assert np.array_equal(gd, rd)
    assert np.array_equal(gi, ri)
    # the case is what it claims: ties straddle the k-th place
    _, rd1 = orc.knn_f32(ip, ix, v, k + 1)
    assert np.any(rd1[:, k - 1] == rd1[:, k])


def test_knn_argument_errors(gpu):
    ip, ix, v = samples(100, 10, 0.3, seed=6)
    M = CsrMatrix.from_arrays(ip, ix, v, (100, 10))
    from scl_amd import _lib
    lib = _lib.load()
    out_i = np.zeros(100 * 5, np.int64)
    out_d = np.zeros(100 * 5, np.float32)
    nr = np.ones(100, np.float32)
    assert lib.scl_knn(M.handle, nr.ctypes.data, 99, 5, out_i.ctypes.data,
                       out_d.ctypes.data) == 11
    assert lib.scl_knn_simple(M.handle, 100, 0, out_i.ctypes.data, out_d.ctypes.data) == 10
    with pytest.raises(SCLError) as e:
        construct.knn(M, 0)
    assert e.value.code == 10


def test_neighbors_reference_cases(gpu):
    # test_neighbors.cpp:71-90: norms of [1, 2], [3, 4], [5, 6]
    M = CsrMatrix.from_arrays(np.array([0, 2, 4, 6], np.int64),
                              np.array([0, 1, 0, 1, 0, 1], np.int64),
                              np.array([1, 2, 3, 4, 5, 6], np.float32), (3, 2))
    assert list(construct.neighbor_norms(M)) == [5.0, 25.0, 61.0]
    # 142-158: an empty matrix; as in the reference (core/sparse.cpp:64-66)
    # a handle with no rows is refused at creation
    with pytest.raises(SCLError) as e:
        CsrMatrix.from_arrays(np.zeros(1, np.int64), np.zeros(0, np.int64),
                              np.zeros(0, np.float32), (0, 10))
    assert e.value.code == 10
    # 166-205 and 348-385: the unit square with explicit zeros stored, k = 2:
    # the two axis neighbours at distance 1, in the reference heap's order
    # (by hand for point 1: j = 0 enters, j = 2 (d^2 = 2) enters and rises to
    # the top, j = 3 (d^2 = 1) replaces it; the heap sort of [(1, 3), (1, 0)]
    # then returns 3 before 0)
    sq = CsrMatrix.from_arrays(np.array([0, 2, 4, 6, 8], np.int64),
                               np.array([0, 1, 0, 1, 0, 1, 0, 1], np.int64),
                               np.array([0, 0, 1, 0, 0, 1, 1, 1], np.float32), (4, 2))
    for gi, gd in (construct.knn(sq, 2), construct.knn(sq, 2, construct.neighbor_norms(sq))):
        assert gi.tolist() == [[1, 2], [3, 0], [3, 0], [2, 1]]
        assert np.all(gd == 1.0)
    # 208-244, 246-268 and 387-425 (seeds 42, 123, 456): 10-30 samples of 5-20
    # features, density 0.05-0.3, k = 1..5; knn with norms equals knn_simple
    for seed in (42, 123, 456):
        rng = np.random.Generator(np.random.PCG64(seed))
        n, f = int(rng.integers(10, 31)), int(rng.integers(5, 21))
        ip, ix, v = samples(n, f, float(rng.uniform(0.05, 0.3)), seed)
        A = CsrMatrix.from_arrays(ip, ix, v, (n, f))
        norms = construct.neighbor_norms(A)
        assert np.array_equal(norms, orc.norms_f32(ip, v))
        ref64 = np.array([float((v[ip[i]:ip[i + 1]].astype(np.float64) ** 2).sum())
                          for i in range(n)])
        assert np.allclose(norms, ref64, rtol=1e-6)
        for k in (1, int(rng.integers(2, min(n - 1, 5) + 1))):
            si, sd = construct.knn(A, k)
            wi, wd = construct.knn(A, k, norms)
            assert np.array_equal(si, wi) and np.array_equal(sd, wd)
            assert np.all(sd >= 0) and np.all((si >= 0) & (si < n))
            assert np.all(np.diff(sd, axis=1) >= 0)
            ri, rd = orc.knn_f32(ip, ix, v, k)
            assert np.array_equal(si, ri) and np.array_equal(sd, rd), (seed, k)
    # 320-346: k larger than the other samples: the rest is -1 / inf
    gi, gd = construct.knn(A, n + 5)
    assert np.all(gi[:, n - 1:] == -1) and np.all(np.isinf(gd[:, n - 1:]))
