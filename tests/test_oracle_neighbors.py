"""Pins of the CPU restatement of the exact k-nearest-neighbour search
(oracle/neighbors.c) against brute force in FP64, hand-built ties and the
reference's own test shapes."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc


def sparse_rows(n, d, density, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    D = (rng.random((n, d)) < density) * rng.gamma(1.0, 1.0, (n, d))
    D = D.astype(np.float32)
    ip = np.concatenate([[0], np.cumsum((D != 0).sum(1))]).astype(np.int64)
    ix = np.concatenate([np.flatnonzero(D[i]) for i in range(n)]).astype(np.int64)
    v = np.concatenate([D[i, D[i] != 0] for i in range(n)]).astype(np.float32)
    return D, ip, ix, v


def test_sum_squared():
    rng = np.random.Generator(np.random.PCG64(2))
    for m in (0, 1, 7, 8, 31, 32, 33, 100, 1000):
        x = rng.standard_normal(m).astype(np.float32)
        ref = float((x.astype(np.float64) ** 2).sum())
        assert abs(orc.sum_squared_f32(x) - ref) <= 1e-6 * max(ref, 1e-30) + 1e-30
    # small integers: every order is exact
    x = rng.integers(-20, 20, 1003).astype(np.float32)
    assert orc.sum_squared_f32(x) == float((x.astype(np.int64) ** 2).sum())
    # 2^24 + 1 + 1: lane 0 of the first accumulator holds 2^24 and lane 0 of
    # the second a 1, which meet only in (a0 + a1) after the blocks of 32
    x = np.zeros(64, np.float32)
    x[0], x[8], x[32 + 8] = 4096.0, 1.0, 1.0      # squares 2^24, 1, 1
    assert orc.sum_squared_f32(x) == 2.0 ** 24 + 2


def test_knn_matches_brute_force():
    D, ip, ix, v = sparse_rows(400, 60, 0.2, seed=3)
    k = 10
    idx, dist = orc.knn_f32(ip, ix, v, k)
    X = D.astype(np.float64)
    sq = (X * X).sum(1)
    full = np.sqrt(np.maximum(sq[:, None] + sq[None, :] - 2 * X @ X.T, 0))
    np.fill_diagonal(full, np.inf)
    checked = 0
    for i in range(400):
        order = np.argsort(full[i], kind="stable")
        assert np.all(np.diff(dist[i]) >= 0)
        assert np.allclose(dist[i], full[i, idx[i]], rtol=1e-5, atol=1e-5)
        # the sorted distances are the k smallest; the set is fixed wherever
        # the k-th and (k+1)-th differ by more than fp32 rounding
        assert np.allclose(dist[i], full[i, order[:k]], rtol=1e-5, atol=1e-5)
        if full[i, order[k]] - full[i, order[k - 1]] > 1e-4:
            assert set(idx[i]) == set(order[:k]), i
            checked += 1
    assert checked > 380


def test_knn_ties_padding_and_duplicates():
    # rows 0..3 identical (distance 0 to each other), row 4 far away
    D = np.array([[1, 0, 2], [1, 0, 2], [1, 0, 2], [1, 0, 2], [9, 9, 9]], np.float32)
    ip = np.array([0, 2, 4, 6, 8, 11], np.int64)
    ix = np.array([0, 2, 0, 2, 0, 2, 0, 2, 0, 1, 2], np.int64)
    v = D[D != 0]
    idx, dist = orc.knn_f32(ip, ix, v, 2)
    # the first zero-distance rows visited fill the heap; later ties are not
    # strictly closer and stay out
    assert list(idx[0]) == [1, 2] and np.all(dist[0] == 0)
    assert set(idx[3]) == {0, 1} and np.all(dist[3] == 0)
    assert set(idx[4]) <= {0, 1, 2, 3}
    assert np.allclose(dist[4], np.sqrt(64 + 81 + 49), rtol=1e-6)
    # more neighbours than other rows: the rest is -1 / inf
    idx, dist = orc.knn_f32(ip, ix, v, 6)
    assert np.all(idx[:, 4:] == -1) and np.all(np.isinf(dist[:, 4:]))
    assert sorted(idx[4, :4]) == [0, 1, 2, 3]
    # an empty row: its norm is 0 and every distance is the other norm
    ip2 = np.array([0, 0, 2, 4], np.int64)
    idx, dist = orc.knn_f32(ip2, ix[:4], v[:4], 2)
    assert np.allclose(dist[0], np.sqrt(5.0), rtol=1e-6)


def test_knn_heap_order_on_the_unit_square_by_hand():
    """The reference's heap (neighbors.hpp:330-476) traced by hand on the
    unit square (reference test_neighbors.cpp:166-205), k = 2.  Point 1 =
    (1, 0): j = 0 enters (d^2 = 1); j = 2 (d^2 = 2) enters and sifts up to the
    top; j = 3 (d^2 = 1 < 2) replaces the top and stays; the heap sort of
    [(1, 3), (1, 0)] (min-heap build, extract, reverse) returns 3 before 0.
    Point 0 fills with 1 and 2 (equal keys never swap) and rejects 3."""
    ip = np.array([0, 2, 4, 6, 8], np.int64)
    ix = np.array([0, 1, 0, 1, 0, 1, 0, 1], np.int64)
    v = np.array([0, 0, 1, 0, 0, 1, 1, 1], np.float32)
    gi, gd = orc.knn_f32(ip, ix, v, 2)
    assert gi.tolist() == [[1, 2], [3, 0], [3, 0], [2, 1]]
    assert np.all(gd == 1.0)
