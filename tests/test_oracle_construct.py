"""Pins of the CPU restatement of the operator-construction steps
(oracle/construct.c): kernel weights, row normalisation, symmetrisation,
the velocity transition and the stationary distribution, against closed
forms, hand computations and dense FP64 linear algebra."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from scl_amd import synth

KINDS = {"gaussian": 0, "epanechnikov": 1, "cosine": 2, "linear": 3, "polynomial": 4,
         "laplacian": 5, "cauchy": 6, "sigmoid": 7, "uniform": 8, "triangular": 9}


def test_kernel_values_against_closed_forms():
    d = np.array([0.0, 0.25, 0.5, 0.9, 1.0, 1.5, 3.0], np.float32)
    h = 1.25
    u = d.astype(np.float64) / h
    ref = {
        "gaussian": np.exp(-u ** 2 / 2),
        "epanechnikov": np.where(u <= 1, 0.75 * (1 - u ** 2), 0.0),
        "cosine": np.where(u <= 1, np.pi / 4 * np.cos(np.pi * u / 2), 0.0),
        "laplacian": np.exp(-u),
        "cauchy": 1 / (1 + u ** 2),
        "uniform": np.where(u < 1, 0.5, 0.0),
        "triangular": np.where(u < 1, 1 - u, 0.0),
    }
    for name, r in ref.items():
        got = orc.kernel_matrix_f32(d, h, KINDS[name])
        assert np.allclose(got, r, rtol=1e-6, atol=1e-7), name
    # kinds without a formula use the Gaussian; tiny bandwidths are clamped
    for name in ("linear", "polynomial", "sigmoid", "gaussian"):
        assert np.array_equal(orc.kernel_matrix_f32(d, h, KINDS[name]),
                              orc.kernel_matrix_f32(d, h, 0))
    assert np.array_equal(orc.kernel_matrix_f32(d, 0.0, 8), np.where(d == 0, 0.5, 0.0))
    # hand values: u = 0.4 gives 0.75 * 0.84 and 1 - 0.4
    assert np.isclose(orc.kernel_matrix_f32([0.5], 1.25, 1)[0], 0.63, rtol=1e-6)
    assert np.isclose(orc.kernel_matrix_f32([0.5], 1.25, 9)[0], 0.6, rtol=1e-6)
    # row sums: stored-order sums of the same values
    ip = np.array([0, 3, 3, 7], np.int64)
    s = orc.kernel_row_sums_f32(ip, d, h, 0)
    k = orc.kernel_matrix_f32(d, h, 0)
    assert s[1] == 0.0 and np.allclose(s, [k[:3].sum(), 0, k[3:].sum()], rtol=1e-6)


def test_vector_sum_keeps_eight_strided_partials():
    # 2^24 + 1 rounds back to 2^24: in index order both ones are lost; with
    # eight strided partials the ones land in lane 1 (1 + 1 = 2) and survive
    v = np.zeros(16, np.float32)
    v[0], v[1], v[8], v[9] = 2.0 ** 24, 1.0, 1.0, 1.0
    assert orc.vector_sum8_f32(v) == 2.0 ** 24 + 2
    assert np.float32(np.cumsum(v, dtype=np.float32)[-1]) == 2.0 ** 24
    # the tail past the last full group of 8 is added after the lanes
    w = np.zeros(9, np.float32)
    w[0], w[1], w[8] = 2.0 ** 24, 1.0, 1.0
    assert orc.vector_sum8_f32(w) == 2.0 ** 24
    r = np.random.Generator(np.random.PCG64(1)).integers(-50, 50, 1001).astype(np.float32)
    assert orc.vector_sum8_f32(r) == r.astype(np.int64).sum()


def test_row_normalize():
    ip = np.array([0, 3, 5, 5, 7], np.int64)
    v = np.array([1, 2, 5, 3, 3, -1, 1], np.float32)
    out = orc.row_normalize_f32(ip, v)
    assert np.allclose(out[:5], [0.125, 0.25, 0.625, 0.5, 0.5], rtol=1e-7)
    assert np.all(out[5:] == 0)              # the row summing to 0 becomes 0


def dense(ip, ix, v, n):
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(ip[i], ip[i + 1]):
            A[i, ix[k]] += v[k]
    return A


def test_symmetrize_and_velocity_transition():
    # by hand: T = [[0, 1], [0.5, 0.5]] -> (T + T^T)/2 = [[0, .75], [.75, .5]]
    S = orc.symmetrize_f32([0, 1, 3], [1, 0, 1], np.array([1, .5, .5], np.float32))
    assert np.allclose(S, [[0, 1], [0.6, 0.4]], rtol=1e-7)
    g = synth.uniform_knn(300, 6, seed=2)
    T = dense(g.indptr, g.indices, g.data, 300)
    M = (T + T.T) / 2
    ref = M / M.sum(1, keepdims=True)
    assert np.allclose(orc.symmetrize_f32(g.indptr, g.indices, g.data), ref, rtol=1e-5,
                       atol=1e-7)
    # velocity graph with signed weights: negatives dropped, all-negative
    # rows become self-loops
    rng = np.random.Generator(np.random.PCG64(3))
    v = rng.standard_normal(len(g.data)).astype(np.float32)
    v[g.indptr[7]:g.indptr[8]] = -np.abs(v[g.indptr[7]:g.indptr[8]])
    S = orc.from_velocity_f32(g.indptr, g.indices, v)
    V = np.maximum(dense(g.indptr, g.indices, v, 300), 0)
    sums = V.sum(1, keepdims=True)
    ref = np.where(sums > 0, V / np.where(sums > 0, sums, 1), np.eye(300))
    assert np.allclose(S, ref, rtol=1e-5, atol=1e-7)
    assert S[7, 7] == 1.0 and np.count_nonzero(S[7]) == 1


def test_stationary_distribution():
    # every row the same distribution p: one step reaches p, the next moves
    # it by rounding only (below tol)
    n = 50
    p = np.random.Generator(np.random.PCG64(6)).random(n)
    p = (p / p.sum()).astype(np.float32)
    ip = np.arange(0, n * n + 1, n, dtype=np.int64)
    ix = np.tile(np.arange(n, dtype=np.int64), n)
    pi, it = orc.stationary_f32(ip, ix, np.tile(p, n), 100, 1e-6)
    assert it == 2 and np.allclose(pi, p, rtol=1e-5)
    # a kNN chain: after `it` iterations pi is the normalised (P^T)^it 1/n,
    # and `it` is the first iteration whose change drops below tol or whose
    # Aitken ratio (diff - d1) / (diff - 2 d1 + d2) of three shrinking
    # changes exceeds 0.9
    for g in (synth.knn_graph(400, 8, seed=4).normalized(),
              synth.uniform_knn(300, 5, seed=5).normalized()):
        n = g.n
        Pt = dense(g.indptr, g.indices, g.data, n).T
        for tol in (1e-9, 1e-4):
            pi, it = orc.stationary_f32(g.indptr, g.indices, g.data, 5000, tol)
            x, diffs = np.full(n, 1 / n), []
            for _ in range(it):
                y = Pt @ x
                y /= y.sum()
                diffs.append(np.abs(y - x).max())
                x = y
            assert np.allclose(pi, x, rtol=2e-5, atol=1e-9)

            def stops(k):
                if diffs[k] < tol:
                    return True
                if k >= 2 and diffs[k - 1] < diffs[k - 2] and diffs[k] < diffs[k - 1]:
                    den = diffs[k] - 2 * diffs[k - 1] + diffs[k - 2]
                    return abs(den) > 1e-15 and (diffs[k] - diffs[k - 1]) / den > 0.9
                return False
            assert stops(it - 1) and not any(stops(k) for k in range(it - 1)), (n, tol, it)
    # no iteration: uniform; one iteration: one normalised step
    u = synth.uniform_knn(300, 5, seed=5).normalized()
    pi0, it0 = orc.stationary_f32(u.indptr, u.indices, u.data, 0, 1e-9)
    assert it0 == 0 and np.all(pi0 == np.float32(1 / 300))
    pi1, it1 = orc.stationary_f32(u.indptr, u.indices, u.data, 1, 1e-9)
    step = dense(u.indptr, u.indices, u.data, 300).T @ np.full(300, 1 / 300)
    assert it1 == 1 and np.allclose(pi1, step / step.sum(), rtol=1e-5)
