"""GPU parity of the other SpMM callers (scl_impute_diffusion,
scl_impute_magic, scl_pseudotime_diffusion, scl_algebra_spmm) against the CPU
restatement (oracle/callers.c).  Imputation, MAGIC and the algebra product
keep every element's FMA chain and are bit-identical; pseudotime sums its
Gram-Schmidt dot products in another order and agrees to rounding."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from scl_amd import CscMatrix, CsrMatrix, callers, synth
from scl_amd._lib import SCLError

pytestmark = pytest.mark.gpu

X_IP, X_IX, X_V = [0, 2, 3, 5], [0, 1, 0, 1, 2], np.array([1, 2, 3, 4, 5], np.float32)
T_IP, T_IX = [0, 2, 4, 6], [0, 1, 0, 2, 1, 2]


def csr(ip, ix, v, shape):
    return CsrMatrix.from_arrays(np.asarray(ip, np.int64), np.asarray(ix, np.int64),
                                 np.asarray(v, np.float32), shape)


def expression(n_cells, n_genes, density, seed, csc=False):
    """Random sparse expression matrix as (indptr, indices, values) with
    sorted slots; primary slots are genes when csc."""
    rng = np.random.Generator(np.random.PCG64(seed))
    P, S = (n_genes, n_cells) if csc else (n_cells, n_genes)
    ip, ix = [0], []
    for _ in range(P):
        cols = np.flatnonzero(rng.random(S) < density)
        ix.extend(cols)
        ip.append(len(ix))
    v = rng.gamma(1.0, 1.0, len(ix)).astype(np.float32)
    return np.array(ip, np.int64), np.array(ix, np.int64), v


def test_impute_diffusion_known_answer(gpu):
    X = csr(X_IP, X_IX, X_V, (3, 3))
    P = csr(T_IP, T_IX, np.full(6, 0.5), (3, 3))
    got = callers.impute_diffusion(X, P, 3, 3, 2)
    assert np.array_equal(got, [[1.25, 2, 1.25], [1.75, 1.5, 1.25], [1, 2.5, 2.5]])


@pytest.mark.parametrize("csc,G", [(False, 96), (True, 96), (False, 160)])
def test_impute_diffusion_matches_restatement(gpu, csc, G):
    # 160 genes on 20,000 cells: a wide operator, so the row-grouped form runs
    n = 20_000
    g = synth.uniform_knn(n, 12, seed=21).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    ip, ix, v = expression(n, G, 0.1, seed=22, csc=csc)
    M = (CscMatrix if csc else CsrMatrix).from_arrays(ip, ix, v, (n, G))
    for steps in (1, 4):
        got = callers.impute_diffusion(M, P, n, G, steps)
        X0 = orc.densify_f32(ip, ix, v, n, G, is_csr=not csc)
        assert np.array_equal(got, orc.impute_diffusion_f32(g.indptr, g.indices, g.data, X0,
                                                            steps))


def test_impute_diffusion_unordered_slots_keep_last_value(gpu):
    # repeated and descending indices, genes past n_genes, a cell past n_cells
    ip, ix = [0, 4, 6, 7, 9], [2, 0, 2, 9, 1, 7, 0, 2, 2]
    v = np.arange(1, 10, dtype=np.float32)
    M = csr(ip, ix, v, (4, 10))
    P = csr(T_IP, T_IX, np.full(6, 0.5), (3, 3))
    got = callers.impute_diffusion(M, P, 3, 3, 1)
    X0 = orc.densify_f32(ip, ix, v, 3, 3)
    assert X0[0, 2] == 3.0
    assert np.array_equal(got, orc.impute_diffusion_f32(T_IP, T_IX, np.full(6, 0.5, np.float32),
                                                        X0, 1))


def test_magic_matches_restatement(gpu):
    X = csr(X_IP, X_IX, X_V, (3, 3))
    av = np.array([0.8, 0.2, 0.7, 0.3, 0.6, 0.4], np.float32)
    A = csr(T_IP, T_IX, av, (3, 3))
    X0 = orc.densify_f32(X_IP, X_IX, X_V, 3, 3)
    assert np.array_equal(callers.impute_magic(X, A, 3, 3, 2),
                          orc.magic_f32(T_IP, T_IX, av, X0, 2))
    # raw kNN weights with a few rows emptied (sum 0: copied) or negative
    n, G = 20_000, 64
    g = synth.knn_graph(n, 10, seed=23)
    vals = g.data.copy()
    for r in (0, 17, 5000):
        vals[g.indptr[r]:g.indptr[r + 1]] = 0.0
    vals[g.indptr[9]:g.indptr[10]] *= -1.0
    A = CsrMatrix.from_arrays(g.indptr, g.indices, vals, (n, n))
    ip, ix, v = expression(n, G, 0.2, seed=24)
    M = CsrMatrix.from_arrays(ip, ix, v, (n, G))
    X0 = orc.densify_f32(ip, ix, v, n, G)
    for t in (1, 3):
        got = callers.impute_magic(M, A, n, G, t)
        ref = orc.magic_f32(g.indptr, g.indices, vals, X0, t)
        assert np.array_equal(got, ref), t
        assert np.array_equal(got[[0, 9, 17, 5000]], X0[[0, 9, 17, 5000]])
    # a genes-major expression matrix is not copied in: all zeros
    Mc = CscMatrix.from_arrays(*expression(n, G, 0.2, seed=24, csc=True), (n, G))
    assert not np.any(callers.impute_magic(Mc, A, n, G, 2))


def test_impute_argument_errors(gpu):
    X = csr(X_IP, X_IX, X_V, (3, 3))
    P = csr(T_IP, T_IX, np.full(6, 0.5), (3, 3))
    for fn in (callers.impute_diffusion, callers.impute_magic):
        with pytest.raises(SCLError) as e:
            fn(X, P, 4, 3, 1)                   # n_cells beyond the operator
        assert e.value.code == 11
        with pytest.raises(SCLError) as e:
            fn(X, P, 3, 3, 0)
        assert e.value.code == 10
    R = csr([0, 1, 2, 2], [0, 1], [1.0, 1.0], (3, 2))
    with pytest.raises(SCLError) as e:
        callers.impute_diffusion(X, R, 3, 3, 1)
    assert e.value.code == 11


def test_pseudotime_matches_restatement(gpu):
    from test_oracle_walks import three_blocks
    ip, ix, v, _ = three_blocks(seed=11, m=600)
    n = len(ip) - 1
    A = CsrMatrix.from_arrays(ip, ix, v, (n, n))
    for root, K in ((0, 3), (1234, 6)):
        got = callers.pseudotime_diffusion(A, root, K)
        ref = orc.pseudotime_f32(ip, ix, v, root, K, 100)
        ref64 = orc.pseudotime_f64(ip, ix, v, root, K, 100)
        # anchored to FP64: the GPU (its sums fold in another order) is at
        # most twice as far from the FP64 result as the restatement
        eg, ec = np.abs(got - ref64).max(), np.abs(ref - ref64).max()
        print(f"\npseudotime three blocks root {root} K {K}: GPU {eg:.2e}, "
              f"restatement {ec:.2e} vs FP64")
        assert eg <= 2 * ec + 1e-7, (root, K)
        assert got[root] == 0.0 and got.max() == 1.0
    # a kNN graph (raw weights, as the reference's tests pass adjacencies)
    g = synth.knn_graph(30_000, 10, seed=25)
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n))
    got = callers.pseudotime_diffusion(A, 77, 4)
    ref = orc.pseudotime_f32(g.indptr, g.indices, g.data, 77, 4, 100)
    ref64 = orc.pseudotime_f64(g.indptr, g.indices, g.data, 77, 4, 100)
    eg, ec = np.abs(got - ref64).max(), np.abs(ref - ref64).max()
    print(f"\npseudotime knn 30k: GPU {eg:.2e}, restatement {ec:.2e} vs FP64")
    assert eg <= 2 * ec + 1e-7
    assert np.all((got >= 0) & (got <= 1)) and got[77] == 0.0


def test_pseudotime_degenerate_basis_falls_back(gpu):
    """The components are orthonormalised by a Cholesky QR; a degenerate basis
    (an adjacency with no nonzeros: every product is zero) is redone with
    Gram-Schmidt, whose zero columns become the normalised FastRNG(c * 12345)
    directions, as in the reference."""
    n = 20_000
    ip, ix, v = np.zeros(n + 1, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32)
    A = CsrMatrix.from_arrays(ip, ix, v, (n, n))
    for root, K in ((0, 3), (777, 5)):
        got = callers.pseudotime_diffusion(A, root, K)
        ref = orc.pseudotime_f32(ip, ix, v, root, K, 100)
        assert np.abs(got - ref).max() < 1e-5, (root, K)
        assert got[root] == 0.0 and got.max() == 1.0


def test_pseudotime_reference_cases(gpu):
    # test_pseudotime.cpp:213-248: path and ring of 5 cells with 10
    # components (more than the cells: the extra columns are rounding noise)
    path = csr([0, 1, 3, 5, 7, 8], [1, 0, 2, 1, 3, 2, 4, 3], np.ones(8), (5, 5))
    pt = callers.pseudotime_diffusion(path, 0, 10)
    assert pt[0] == 0.0 and pt.max() == 1.0 and np.all(pt >= 0)
    ring = csr([0, 2, 4, 6, 8, 10], [1, 4, 0, 2, 1, 3, 2, 4, 0, 3], np.ones(10), (5, 5))
    assert callers.pseudotime_diffusion(ring, 0, 10)[0] == 0.0
    assert not np.any(callers.pseudotime_diffusion(ring, 2, 0))
    with pytest.raises(SCLError) as e:
        callers.pseudotime_diffusion(ring, 5, 3)
    assert e.value.code == 10
    with pytest.raises(SCLError) as e:
        callers.pseudotime_diffusion(ring, 0, 3, n=4)
    assert e.value.code in (10, 11)


def test_algebra_spmm_known_answers(gpu):
    A = csr([0, 2, 3, 6], [0, 2, 1, 0, 1, 2], [1, 2, 3, 4, 5, 6], (3, 3))
    Y = np.full((3, 2), np.nan, np.float32)
    callers.algebra_spmm(A, np.array([[1, 2], [3, 4], [5, 6]], np.float32), Y, 1.0, 0.0)
    assert np.array_equal(Y, [[11, 14], [9, 12], [49, 64]])
    Y = np.array([[10], [20], [30]], np.float32)
    callers.algebra_spmm(A, np.array([[1], [2], [3]], np.float32), Y, 2.0, 3.0)
    assert np.array_equal(Y[:, 0], [44, 72, 154])


@pytest.mark.parametrize("F", [1, 7, 64, 300])
def test_algebra_spmm_matches_restatement(gpu, F):
    rng = np.random.Generator(np.random.PCG64(F))
    m, k = 5000, 3000
    ip, ix, v = expression(m, k, 0.004, seed=F + 1)
    v[::17] = 0.0                              # explicit zeros: skipped
    A = CsrMatrix.from_arrays(ip, ix, v, (m, k))
    X = rng.standard_normal((k, F)).astype(np.float32)
    X[ix[::17][:5]] = np.inf                   # only reached through zero weights
    for alpha, beta in ((1.0, 0.0), (0.7, 1.0), (-1.5, 0.25), (0.0, 2.0)):
        Y0 = rng.standard_normal((m, F)).astype(np.float32)
        got = callers.algebra_spmm(A, X, Y0.copy(), alpha, beta)
        ref = orc.algebra_spmm_f32(ip, ix, v, X, Y0, alpha, beta)
        assert np.array_equal(got, ref, equal_nan=True), (alpha, beta)
    # a CSC handle multiplies through its columns: A^T X for the same arrays
    At = CscMatrix.from_arrays(ip, ix, v, (k, m))
    Xt = rng.standard_normal((k, F)).astype(np.float32)
    got = callers.algebra_spmm(At, Xt, np.zeros((m, F), np.float32), 1.0, 0.0)
    ref = orc.algebra_spmm_f32(ip, ix, v, Xt, np.zeros((m, F), np.float32), 1.0, 0.0)
    assert np.array_equal(got, ref)
