"""Pins of the CPU restatement of the vector-side consumers (oracle/walks.c)
against closed forms, hand-computed answers and dense FP64 linear algebra."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from scl_amd import synth


def dense(g):
    n = len(g.indptr) - 1
    P = np.zeros((n, n))
    for i in range(n):
        for k in range(g.indptr[i], g.indptr[i + 1]):
            P[i, g.indices[k]] += g.data[k]
    return P


def small(n=300, k=6, seed=1, sym=False):
    g = (synth.knn_graph(n, k, seed=seed) if sym else synth.uniform_knn(n, k, seed=seed))
    return g.normalized()


def test_lazy_walk_is_a_matrix_power():
    g = small()
    x = synth.normal(300, 1, seed=2)[:, 0]
    P = dense(g)
    for lam in (0.5, 0.2):
        got = orc.lazy_walk_f32(g.indptr, g.indices, g.data, x, 7, lam)
        M = lam * np.eye(300) + (1 - lam) * P
        ref = np.linalg.matrix_power(M, 7) @ x.astype(np.float64)
        assert orc.rel_err(got, ref) < 2e-6


def test_heat_kernel_signature_is_a_matrix_power():
    g = synth.knn_graph(300, 6, seed=1)       # raw kernel weights
    # symmetric normalisation: not stochastic, so the signature moves
    ip, ix = g.indptr, g.indices
    v = orc.transition_f32(ip, ix, g.data, symmetric=True)
    P = dense(type(g)(g.n, ip, ix, v))
    got = orc.heat_kernel_f32(ip, ix, v, 1.5, 12)
    dt = 1.5 / 12
    ref = np.linalg.matrix_power((1 - dt) * np.eye(g.n) + dt * P, 12) @ np.ones(g.n)
    assert orc.rel_err(got, ref) < 2e-6
    assert not np.allclose(got, 1.0)


def test_rwr_reaches_the_linear_fixed_point_and_keeps_the_swap():
    g = small()
    P = dense(g)
    seeds = [3, 77, 77, -5, 10_000]          # duplicates and out-of-range kept as given
    r = np.zeros(300)
    r[[3, 77]] = 1.0
    r /= 3.0                                   # three in-range entries, 77 counted twice
    alpha = 0.85
    got, it = orc.rwr_f32(g.indptr, g.indices, g.data, seeds, alpha, 500, 1e-7)
    ref = np.linalg.solve(np.eye(300) - alpha * P, (1 - alpha) * r)
    assert it < 500 and orc.rel_err(got, ref) < 1e-5
    # no convergence (negative tolerance): an odd max_iter leaves the iterate
    # before the last in the caller's buffer, an even one the last
    a3, it3 = orc.rwr_f32(g.indptr, g.indices, g.data, [5], alpha, 3, -1.0)
    a2, _ = orc.rwr_f32(g.indptr, g.indices, g.data, [5], alpha, 2, -1.0)
    a4, _ = orc.rwr_f32(g.indptr, g.indices, g.data, [5], alpha, 4, -1.0)
    assert it3 == 3 and np.array_equal(a3, a2) and not np.array_equal(a4, a2)
    # no seed in range: all zero
    z, _ = orc.rwr_f32(g.indptr, g.indices, g.data, [-1, 400], alpha, 10, 1e-6)
    assert np.all(z == 0)


def test_dpt_known_answers():
    # two nodes swapping: node 1 is reached at step 1 with probability 1
    pt = orc.dpt_f32([0, 1, 2], [1, 0], np.ones(2, np.float32), 0, 100, 1e-6)
    assert np.array_equal(pt, [0.0, 1.0])
    # path 0-1-2 (rows [0,1,0], [.5,0,.5], [0,1,0]) from node 0, by hand:
    # after 4 iterations visited = [1, .75, .75], hitting = [0, 1.25, 2]
    ip, ix = [0, 1, 3, 4], [1, 0, 2, 1]
    v = np.array([1, .5, .5, 1], np.float32)
    pt = orc.dpt_f32(ip, ix, v, 0, 4, 1e-6)
    assert np.allclose(pt, [0.0, (1.25 / .75) / (2 / .75), 1.0], atol=1e-6)
    # after 2 iterations nothing but the root passed 0.5 visited
    assert np.array_equal(orc.dpt_f32(ip, ix, v, 0, 2, 1e-6), [0.0, 1.0, 1.0])


def test_dpt_multi_root_is_the_mean_of_single_roots():
    g = small(200, 5, seed=4, sym=True)
    singles = [orc.dpt_f32(g.indptr, g.indices, g.data, r, 60, 1e-6) for r in (0, 9)]
    multi = orc.dpt_multi_f32(g.indptr, g.indices, g.data, [0, -3, 9, 999], 60)
    assert np.allclose(multi, (singles[0] + singles[1]) / 2, rtol=1e-6, atol=1e-7)
    assert np.all(orc.dpt_multi_f32(g.indptr, g.indices, g.data, [-1], 60) == 1.0)
    assert singles[0][0] == 0.0 and np.all((singles[0] >= 0) & (singles[0] <= 1))


def test_gram_schmidt_against_qr():
    rng = np.random.Generator(np.random.PCG64(3))
    A = rng.standard_normal((500, 6)).astype(np.float32)
    Q = orc.mgs_f32(A)
    q, r = np.linalg.qr(A.astype(np.float64))
    q *= np.sign(np.diag(r))                  # MGS keeps positive diagonals
    assert np.abs(Q - q).max() < 1e-5
    # a zero column becomes a unit vector
    A[:, 2] = 0.0
    Q = orc.mgs_f32(A)
    assert Q[2, 2] == 1.0 and np.count_nonzero(Q[:, 2]) == 1


def three_blocks(seed=5, m=100):
    """Three dense random blocks joined by a few edges, symmetrically
    normalised: a symmetric operator whose three leading eigenvalues sit far
    above the rest."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 3 * m
    W = np.zeros((n, n))
    for b in range(3):
        blk = rng.random((m, m)) * (rng.random((m, m)) < 0.3)
        W[b * m:(b + 1) * m, b * m:(b + 1) * m] = blk + blk.T
    for _ in range(6):
        i, j = rng.integers(0, n, 2)
        W[i, j] = W[j, i] = 0.5
    np.fill_diagonal(W, 0)
    d = W.sum(1)
    S = W / np.sqrt(d[:, None] * d[None, :])
    ip = np.concatenate([[0], np.cumsum((S != 0).sum(1))]).astype(np.int64)
    ix = np.concatenate([np.flatnonzero(S[i]) for i in range(n)]).astype(np.int64)
    v = np.concatenate([S[i, S[i] != 0] for i in range(n)]).astype(np.float32)
    return ip, ix, v, S


def test_diffusion_map_finds_the_leading_eigenvectors():
    ip, ix, v, S = three_blocks()
    Q = orc.dmap_f32(ip, ix, v, 3, 300)
    w, V = np.linalg.eigh(S)
    order = np.argsort(-np.abs(w))
    assert abs(w[order[2]]) > 1.5 * abs(w[order[3]])
    top = V[:, order[:3]]
    # the embedding spans the top-3 eigenspace: residual of projecting it out
    B, _ = np.linalg.qr(top)
    resid = Q - B @ (B.T @ Q)
    assert np.abs(resid).max() < 1e-3
    assert np.abs(Q.T.astype(np.float64) @ Q - np.eye(3)).max() < 1e-5
    init = orc.dmap_init_f32(4, 2)
    assert np.isclose(init[1, 0], np.cos(np.pi * 0.25) + 0.1 * np.sin(np.pi * 0.25 * 2.7),
                      atol=1e-6)


def test_diffusion_map_fp64_anchor():
    """The FP64 anchor: no iteration is the QR of the fp32 start (R with a
    positive diagonal); one step on a non-symmetric P is the QR of P Q0 (a
    transposed or misindexed SpMM fails it); 300 steps span the top-3
    eigenspace to FP64 accuracy; the fp32 restatement stays near it."""
    g = small(200, 5, seed=4)
    P = dense(g)
    init = orc.dmap_init_f32(200, 3).astype(np.float64)

    def qr_pos(A):
        q, r = np.linalg.qr(A)
        return q * np.sign(np.diag(r))

    Q0 = qr_pos(init)
    assert np.abs(orc.dmap_f64(g.indptr, g.indices, g.data, 3, 0) - Q0).max() < 1e-12
    Q1 = qr_pos(P @ Q0)
    assert np.abs(orc.dmap_f64(g.indptr, g.indices, g.data, 3, 1) - Q1).max() < 1e-12
    assert np.abs(qr_pos(P.T @ Q0) - Q1).max() > 1e-2
    ip, ix, v, S = three_blocks()
    Q = orc.dmap_f64(ip, ix, v, 3, 300)
    w, V = np.linalg.eigh(S.astype(np.float32).astype(np.float64))
    order = np.argsort(-np.abs(w))
    B, _ = np.linalg.qr(V[:, order[:3]])
    assert np.abs(Q - B @ (B.T @ Q)).max() < 1e-9
    assert np.abs(Q.T @ Q - np.eye(3)).max() < 1e-13
    Q32 = orc.dmap_f32(ip, ix, v, 3, 60)
    e = orc.rel_err(Q32, orc.dmap_f64(ip, ix, v, 3, 60))
    assert 0 < e < 1e-5


def test_diffusion_distance_both_branches():
    for n in (120, 1001):                     # rows of P^t; rows of (P^T)^t
        g = small(n, 5, seed=n)
        P = dense(g)
        D = orc.diffusion_distance_f32(g.indptr, g.indices, g.data, 2)
        A = np.linalg.matrix_power(P if n <= 1000 else P.T, 2)
        sq = (A * A).sum(1)
        ref = np.sqrt(np.maximum(sq[:, None] + sq[None, :] - 2 * A @ A.T, 0))
        np.fill_diagonal(ref, 0)
        assert np.allclose(D, ref, rtol=1e-4, atol=1e-6)
        assert np.array_equal(D, D.T) and np.all(np.diag(D) == 0)
