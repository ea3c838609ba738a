"""Pins of the CPU restatement of the other SpMM callers (oracle/callers.c):
diffusion imputation, MAGIC, diffusion-component pseudotime and
Y = alpha A X + beta Y, against the reference's own known answers, hand
computations, published generator definitions and dense FP64 algebra."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from scl_amd import synth

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1


def dense(ip, ix, v, cols=None):
    n = len(ip) - 1
    A = np.zeros((n, cols if cols is not None else n))
    for i in range(n):
        for k in range(ip[i], ip[i + 1]):
            A[i, ix[k]] += v[k]
    return A


# the expression matrix and transition of test_impute.cpp:268-282
X_IP, X_IX, X_V = [0, 2, 3, 5], [0, 1, 0, 1, 2], np.array([1, 2, 3, 4, 5], np.float32)
T_IP, T_IX = [0, 2, 4, 6], [0, 1, 0, 2, 1, 2]


def test_densify_keeps_stored_order_and_range():
    X = orc.densify_f32(X_IP, X_IX, X_V, 3, 3)
    assert np.array_equal(X, [[1, 2, 0], [3, 0, 0], [0, 4, 5]])
    # a repeated index keeps the last stored value; out-of-range entries and
    # cells past n_cells are dropped
    ip, ix = [0, 3, 4, 5], [1, 1, 7, -1, 0]
    v = np.array([9, 8, 7, 6, 5], np.float32)
    assert np.array_equal(orc.densify_f32(ip, ix, v, 2, 3), [[0, 8, 0], [0, 0, 0]])
    # CSC: primary slots are genes
    cip, cix = [0, 2, 3], [0, 2, 1]
    cv = np.array([1, 2, 3], np.float32)
    assert np.array_equal(orc.densify_f32(cip, cix, cv, 3, 2, is_csr=False),
                          [[1, 0], [0, 3], [2, 0]])


def test_impute_diffusion_known_answer():
    # test_impute.cpp:268-296 (the reference asserts finiteness only); P^2 X
    # by hand: every number is a dyadic rational, so exact
    X = orc.densify_f32(X_IP, X_IX, X_V, 3, 3)
    got = orc.impute_diffusion_f32(T_IP, T_IX, np.full(6, 0.5, np.float32), X, 2)
    assert np.array_equal(got, [[1.25, 2, 1.25], [1.75, 1.5, 1.25], [1, 2.5, 2.5]])
    # odd step counts end in the caller's buffer too
    one = orc.impute_diffusion_f32(T_IP, T_IX, np.full(6, 0.5, np.float32), X, 1)
    assert np.array_equal(one, [[2, 1, 0], [0.5, 3, 2.5], [1.5, 2, 2.5]])


def test_impute_diffusion_is_a_matrix_power():
    g = synth.uniform_knn(400, 7, seed=3).normalized()
    X = synth.normal(400, 33, seed=4)
    got = orc.impute_diffusion_f32(g.indptr, g.indices, g.data, X, 3)
    P = dense(g.indptr, g.indices, g.data)
    ref = np.linalg.matrix_power(P, 3) @ X.astype(np.float64)
    assert orc.rel_err(got, ref) < 2e-6


def test_magic_normalises_rows_and_copies_empty_ones():
    # test_impute.cpp:366-391: affinity rows sum to one (in fp32, nearly)
    X = orc.densify_f32(X_IP, X_IX, X_V, 3, 3)
    av = np.array([0.8, 0.2, 0.7, 0.3, 0.6, 0.4], np.float32)
    A = dense(T_IP, T_IX, av.astype(np.float64))
    P = A / A.sum(1, keepdims=True)
    got = orc.magic_f32(T_IP, T_IX, av, X, 2)
    assert np.allclose(got, P @ P @ X.astype(np.float64), rtol=1e-6, atol=1e-6)
    # a row summing to zero (or below 1e-12, negative included) keeps its input
    ip, ix = [0, 2, 4, 5], [0, 1, 1, 2, 0]
    v = np.array([2.0, 6.0, 1.0, -1.0, -3.0], np.float32)
    X = np.arange(6, dtype=np.float32).reshape(3, 2) - 2.0
    got = orc.magic_f32(ip, ix, v, X, 3)
    Y = X.astype(np.float64)
    for _ in range(3):
        Y = np.stack([0.25 * Y[0] + 0.75 * Y[1], Y[1], Y[2]])
    assert np.allclose(got, Y, rtol=1e-6, atol=1e-6)
    assert np.array_equal(got[1:], X[1:])


def _splitmix64_first(state):
    """First output of SplitMix64 (Steele, Lea and Flood 2014) from a state."""
    z = (state + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _fin(z):
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_fastrng_seeding_is_chained_splitmix_rounds():
    # the published first output of SplitMix64 from state 0
    assert _splitmix64_first(0) == 0xE220A8397B1DCDAF
    for seed in (0, 42, 12345, 2 ** 40 + 7):
        z, st = orc.fastrng_seed(seed)
        prev = seed
        for i in range(4):
            # each round is the first half of one SplitMix64 output whose
            # state is the previous round's mixed value
            assert _fin(int(z[i])) == _splitmix64_first(prev)
            assert int(st[i]) == int(z[i]) >> 32
            prev = int(z[i])


def _rev32(x):
    return int(f"{x:032b}"[::-1], 2)


def test_fastrng_step_is_mirrored_marsaglia_xorshift128():
    # Marsaglia, "Xorshift RNGs" (J. Stat. Softw. 8(14), 2003), xor128:
    # t = x ^ (x << 11); x, y, z = y, z, w; w = w ^ (w >> 19) ^ (t ^ (t >> 8)).
    # Reversing the bits of every word turns left shifts into right shifts:
    # the generator here must be xor128 conjugated by bit reversal, with the
    # newest word first.
    rng = np.random.Generator(np.random.PCG64(9))
    for _ in range(20):
        st = rng.integers(1, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        x, y, z, w = (_rev32(int(st[3])), _rev32(int(st[2])), _rev32(int(st[1])),
                      _rev32(int(st[0])))
        out, _ = orc.fastrng_step(st, 6)
        for o in out:
            t = (x ^ (x << 11)) & M32
            x, y, z = y, z, w
            w = (w ^ (w >> 19) ^ (t ^ (t >> 8))) & M32
            assert _rev32(int(o)) == w
    u = orc.fastrng_uniform(42, 10000)
    _, st = orc.fastrng_seed(42)
    raw, _ = orc.fastrng_step(st, 10000)
    assert np.array_equal(u, raw.astype(np.float32) * np.float32(2.3283064365386963e-10))
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def test_pseudotime_orthonormalisation():
    rng = np.random.Generator(np.random.PCG64(5))
    A = rng.standard_normal((400, 5)).astype(np.float32)
    Q = orc.pt_orthonormalise_f32(A)
    q, r = np.linalg.qr(A.astype(np.float64))
    assert np.abs(Q - q * np.sign(np.diag(r))).max() < 1e-5
    # a zero column (norm at most 1e-15) is replaced by the normalised
    # FastRNG(c * 12345) direction uniform() - 0.5 (here c = 3), not
    # orthogonalised; a dependent column leaves rounding noise above 1e-15
    A[:, 3] = 0.0
    Q = orc.pt_orthonormalise_f32(A)
    d = orc.fastrng_uniform(3 * 12345, 400).astype(np.float64) - 0.5
    assert np.abs(Q[:, 3] - d / np.linalg.norm(d)).max() < 1e-6


def test_pseudotime_depends_only_on_the_leading_subspace():
    from test_oracle_walks import three_blocks
    ip, ix, v, S = three_blocks(seed=11, m=80)
    w, V = np.linalg.eigh(S)
    order = np.argsort(-np.abs(w))
    assert abs(w[order[2]]) > 1.5 * abs(w[order[3]])
    B = V[:, order[:3]]
    for root in (0, 137):
        got = orc.pseudotime_f32(ip, ix, v, root, 3, 100)
        # distances in the embedding depend on the spanned subspace only:
        # |Q^T (e_i - e_r)| = |B^T (e_i - e_r)| once the iteration converged
        d = np.linalg.norm(B - B[root], axis=1)
        assert np.abs(got - d / d.max()).max() < 1e-4
        assert got[root] == 0.0 and got.max() == 1.0
    # no iteration: the distances of the orthonormalised FastRNG(42) start
    n = len(ip) - 1
    Q0 = orc.pt_orthonormalise_f32(
        (orc.fastrng_uniform(42, n * 4) - np.float32(0.5)).reshape(n, 4))
    d = np.sqrt(((Q0.astype(np.float64) - Q0[5]) ** 2).sum(1))
    assert np.abs(orc.pseudotime_f32(ip, ix, v, 5, 4, 0) - d / d.max()).max() < 1e-6


def test_pseudotime_fp64_anchor():
    """The FP64 anchor: converged distances equal the leading eigenspace's
    to FP64 accuracy (the fp32 restatement reaches 1e-4); with no iteration
    they are those of the QR of the fp32 FastRNG(42) start in FP64."""
    from test_oracle_walks import three_blocks
    ip, ix, v, S = three_blocks(seed=11, m=80)
    w, V = np.linalg.eigh(S.astype(np.float32).astype(np.float64))
    order = np.argsort(-np.abs(w))
    B = V[:, order[:3]]
    for root in (0, 137):
        got = orc.pseudotime_f64(ip, ix, v, root, 3, 300)
        d = np.linalg.norm(B - B[root], axis=1)
        assert np.abs(got - d / d.max()).max() < 1e-9
        assert got[root] == 0.0 and got.max() == 1.0
        e = np.abs(orc.pseudotime_f32(ip, ix, v, root, 3, 100)
                   - orc.pseudotime_f64(ip, ix, v, root, 3, 100)).max()
        assert 0 < e < 1e-5
    n = len(ip) - 1
    A = (orc.fastrng_uniform(42, n * 4) - np.float32(0.5)).reshape(n, 4).astype(np.float64)
    q, r = np.linalg.qr(A)
    Q0 = q * np.sign(np.diag(r))
    d = np.linalg.norm(Q0 - Q0[5], axis=1)
    assert np.abs(orc.pseudotime_f64(ip, ix, v, 5, 4, 0) - d / d.max()).max() < 1e-12


def test_algebra_spmm_known_answers_and_skips():
    # test_algebra.cpp:281-325 with make_test_matrix_3x3 [[1,0,2],[0,3,0],[4,5,6]]
    ip, ix = [0, 2, 3, 6], [0, 2, 1, 0, 1, 2]
    v = np.array([1, 2, 3, 4, 5, 6], np.float32)
    X = np.array([[1, 2], [3, 4], [5, 6]], np.float32)
    Y = orc.algebra_spmm_f32(ip, ix, v, X, np.full((3, 2), np.nan, np.float32), 1.0, 0.0)
    assert np.array_equal(Y, [[11, 14], [9, 12], [49, 64]])
    Y = orc.algebra_spmm_f32(ip, ix, v, [[1], [2], [3]], [[10], [20], [30]], 2.0, 3.0)
    assert np.array_equal(Y[:, 0], [44, 72, 154])
    # alpha 0: only the beta scaling; beta 1 keeps Y as it is
    Y0 = np.array([[1], [-2], [3]], np.float32)
    assert np.array_equal(orc.algebra_spmm_f32(ip, ix, v, [[1], [2], [3]], Y0, 0.0, 0.5),
                          Y0 * 0.5)
    # a zero weight skips its neighbour: an infinite X row does not leak in
    vz = v.copy()
    vz[1] = 0.0
    Xi = np.array([[1], [1], [np.inf]], np.float32)
    Y = orc.algebra_spmm_f32(ip, ix, vz, Xi, [[0], [0], [0]], 1.0, 1.0)
    assert Y[0, 0] == 1.0 and Y[1, 0] == 3.0 and np.isinf(Y[2, 0])


def test_algebra_spmm_rectangular_against_dense():
    rng = np.random.Generator(np.random.PCG64(8))
    m, k, F = 300, 170, 9
    D = rng.standard_normal((m, k)) * (rng.random((m, k)) < 0.05)
    ip = np.concatenate([[0], np.cumsum((D != 0).sum(1))])
    ix = np.concatenate([np.flatnonzero(D[i]) for i in range(m)])
    v = np.concatenate([D[i, D[i] != 0] for i in range(m)]).astype(np.float32)
    X = rng.standard_normal((k, F)).astype(np.float32)
    Y = rng.standard_normal((m, F)).astype(np.float32)
    got = orc.algebra_spmm_f32(ip, ix, v, X, Y, 0.7, -1.3)
    A = dense(ip, ix, v.astype(np.float64), k)
    ref = np.float64(np.float32(0.7)) * (A @ X) + np.float64(np.float32(-1.3)) * Y
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()
