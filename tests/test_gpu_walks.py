"""GPU parity of the vector-side diffusion entry points against the CPU
restatement (oracle/walks.c).  Walks whose only reductions are maxima (DPT)
or that reduce nothing (lazy walk, heat kernel, distances) are bit-identical;
random walk with restart is bit-identical while it runs to max_iter and
within rounding of its convergence test otherwise; the diffusion map's
Gram-Schmidt sums in a different order and agrees to 1e-4."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from scl_amd import CsrMatrix, diffusion, synth
from scl_amd._lib import SCLError

pytestmark = pytest.mark.gpu


def op(n=20_000, k=10, seed=1, sym=True):
    g = (synth.knn_graph(n, k, seed=seed) if sym else synth.uniform_knn(n, k, seed=seed))
    g = g.normalized()
    return g, CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n))


def test_lazy_walk_and_heat_kernel(gpu):
    g, P = op()
    x = synth.normal(g.n, 1, seed=3)[:, 0]
    for lam, used in ((0.3, 0.3), (0.0, 0.5)):        # 0 selects 0.5
        got = diffusion.lazy_random_walk(P, x.copy(), 6, lam)
        assert np.array_equal(got, orc.lazy_walk_f32(g.indptr, g.indices, g.data, x, 6, used))
    # heat kernel on a symmetric (non-stochastic) normalisation
    raw = synth.knn_graph(20_000, 10, seed=2)
    A = CsrMatrix.from_arrays(raw.indptr, raw.indices, raw.data, (raw.n, raw.n))
    diffusion.compute_transition_matrix(A, symmetric=True)
    ip, ix, v = A.export()
    for t, steps, used in ((1.5, 12, 12), (0.7, 0, 10)):
        got = diffusion.heat_kernel_signature(A, t, steps)
        assert np.array_equal(got, orc.heat_kernel_f32(ip, ix, v, t, used))


def test_renumbered_products_follow_value_changes(gpu):
    """From 16,384 rows an operator has a locality schedule, and the walks'
    single-vector products run a copy renumbered in that order (x gathered
    into it, each row stored back at its place): bit-identical, also after
    device values change (set_values) and after host values change (a
    re-normalisation), on a symmetric and on a non-symmetric pattern."""
    import torch
    from scl_amd import torch_ops
    for sym in (True, False):
        g, P = op(20_000, 10, seed=5, sym=sym)
        x = synth.normal(g.n, 1, seed=6)[:, 0]
        got = diffusion.lazy_random_walk(P, x.copy(), 5, 0.3)
        assert np.array_equal(got, orc.lazy_walk_f32(g.indptr, g.indices, g.data, x, 5, 0.3))
        v2 = (g.data * 0.5).astype(np.float32)
        torch_ops.set_values(P, torch.from_numpy(v2).cuda())
        got = diffusion.lazy_random_walk(P, x.copy(), 5, 0.3)
        assert np.array_equal(got, orc.lazy_walk_f32(g.indptr, g.indices, v2, x, 5, 0.3))
    raw = synth.knn_graph(20_000, 10, seed=7)
    A = CsrMatrix.from_arrays(raw.indptr, raw.indices, raw.data, (raw.n, raw.n))
    x = synth.normal(raw.n, 1, seed=8)[:, 0]
    got = diffusion.lazy_random_walk(A, x.copy(), 3, 0.3)
    assert np.array_equal(got, orc.lazy_walk_f32(raw.indptr, raw.indices, raw.data, x, 3, 0.3))
    diffusion.compute_transition_matrix(A)
    ip, ix, v = A.export()
    assert not np.array_equal(v, raw.data)
    got = diffusion.lazy_random_walk(A, x.copy(), 3, 0.3)
    assert np.array_equal(got, orc.lazy_walk_f32(ip, ix, v, x, 3, 0.3))


def test_renumbered_products_of_a_csc_handle(gpu):
    """A CSC handle walks its primary slots (columns) as rows, as the
    reference's visit() does; its renumbered copy is built from the same
    slots, so a walk on it equals the restatement on those arrays."""
    from scl_amd import CscMatrix
    g = synth.knn_graph(20_000, 10, seed=12).normalized()
    C = CscMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n))
    x = synth.normal(g.n, 1, seed=13)[:, 0]
    got = diffusion.lazy_random_walk(C, x.copy(), 4, 0.25)
    assert np.array_equal(got, orc.lazy_walk_f32(g.indptr, g.indices, g.data, x, 4, 0.25))


def test_renumbered_dpt(gpu):
    """From 16,384 rows DPT runs in the operator's locality numbering (its
    reductions are maxima, which do not depend on the order): bit-identical
    to the restatement, one root and several."""
    g, P = op(20_000, 10, seed=14, sym=False)
    for root, iters, tol in ((3, 40, 1e-6), (19_999, 12, 1e-2)):
        got = diffusion.compute_dpt(P, root, iters, tol)
        assert np.array_equal(got, orc.dpt_f32(g.indptr, g.indices, g.data, root, iters, tol))
    multi = diffusion.compute_dpt_multi_root(P, [5, 70, -2], 20)
    assert np.array_equal(multi, orc.dpt_multi_f32(g.indptr, g.indices, g.data, [5, 70, -2],
                                                   20))


def test_random_walk_with_restart(gpu):
    g, P = op(sym=False)
    seeds = [5, 17, 17, -4, 10 ** 9]
    # runs to max_iter (tiny tol): bit-identical, odd counts keep the swap
    for iters in (7, 8):
        got = diffusion.random_walk_with_restart(P, seeds, 0.8, iters, 1e-30)
        ref, it = orc.rwr_f32(g.indptr, g.indices, g.data, seeds, 0.8, iters, -1.0)
        assert it == iters and np.array_equal(got, ref)
    # defaults (alpha 0.85, tol 1e-6, 100 iterations): converged
    got = diffusion.random_walk_with_restart(P, seeds)
    ref, it = orc.rwr_f32(g.indptr, g.indices, g.data, seeds, 0.85, 100, 1e-6)
    assert it < 100 and orc.rel_err(got, ref) < 1e-4
    ppr = diffusion.personalized_pagerank(P, 17, 0.85, 9, 1e-30)
    assert np.array_equal(ppr, orc.rwr_f32(g.indptr, g.indices, g.data, [17], 0.85, 9,
                                           -1.0)[0])
    assert np.all(diffusion.random_walk_with_restart(P, [-1]) == 0)


def test_walks_stopping_inside_a_batch_of_steps(gpu):
    """The device checks convergence every step and the host reads the flag
    every few steps: walks that converge on the first, a middle or the last
    step of such a batch give the restatement's result.  The RWR tolerance
    is placed mid-way between two consecutive step totals (the device sums
    them in another order); the DPT test takes a maximum, which is exact."""
    g, P = op(sym=False)
    seeds = [5, 17, 17]
    grid = 10.0 ** np.linspace(-1, -8, 141)
    stops = {}
    for tol in grid:
        stops.setdefault(orc.rwr_f32(g.indptr, g.indices, g.data, seeds, 0.85, 100,
                                     tol)[1], []).append(tol)
    checked = 0
    for it in (3, 8, 9, 11, 16, 17):
        if len(stops.get(it, [])) < 3:
            continue
        tol = float(np.median(stops[it]))
        ref, n_it = orc.rwr_f32(g.indptr, g.indices, g.data, seeds, 0.85, 100, tol)
        got = diffusion.random_walk_with_restart(P, seeds, 0.85, 100, tol)
        assert n_it == it and np.array_equal(got, ref), it
        checked += 1
    assert checked >= 3
    g, P = op(8000, 8, seed=4)
    for tol in (0.5, 0.2, 0.05, 1e-2, 3e-3):
        got = diffusion.compute_dpt(P, 77, 100, tol)
        assert np.array_equal(got, orc.dpt_f32(g.indptr, g.indices, g.data, 77, 100, tol)), tol


def test_diffusion_pseudotime(gpu):
    g, P = op(8000, 8, seed=4)
    for root, iters, tol in ((0, 60, 1e-6), (123, 0, 0.0), (4000, 25, 1e-3)):
        got = diffusion.compute_dpt(P, root, iters, tol)
        ref = orc.dpt_f32(g.indptr, g.indices, g.data, root, iters or 100, tol or 1e-6)
        assert np.array_equal(got, ref), root
        assert got[root] == 0.0
    multi = diffusion.compute_dpt_multi_root(P, [0, -1, 123, 10 ** 6], 40)
    assert np.array_equal(multi, orc.dpt_multi_f32(g.indptr, g.indices, g.data,
                                                   [0, -1, 123, 10 ** 6], 40))
    assert np.all(diffusion.compute_dpt_multi_root(P, [-5], 10) == 1.0)
    with pytest.raises(SCLError) as e:
        diffusion.compute_dpt(P, g.n, 10)
    assert e.value.code == 10


def test_diffusion_map_embedding(gpu):
    from test_oracle_walks import three_blocks
    ip, ix, v, _ = three_blocks(seed=7, m=400)
    n = len(ip) - 1
    P = CsrMatrix.from_arrays(ip, ix, v, (n, n))
    got = diffusion.diffusion_map_embedding(P, 3, 60)
    ref = orc.dmap_f32(ip, ix, v, 3, 60)
    ref64 = orc.dmap_f64(ip, ix, v, 3, 60)
    # anchored to FP64: the GPU (its sums fold in another order) is at most
    # twice as far from the FP64 result as the restatement
    eg, ec = orc.rel_err(got, ref64), orc.rel_err(ref, ref64)
    print(f"\ndiffusion map three blocks: GPU {eg:.2e}, restatement {ec:.2e} vs FP64")
    assert eg <= 2 * ec + 1e-7
    assert np.abs(got.T.astype(np.float64) @ got - np.eye(3)).max() < 1e-5
    # defaults (50 iterations) and more components than the operator's rank
    assert diffusion.diffusion_map_embedding(P, 5).shape == (n, 5)


def test_diffusion_map_embedding_wide(gpu):
    # from 16384 rows the basis is orthogonalised by grid-wide reductions
    # (transposed columns); sums fold in another order than the restatement's
    g, P = op(30_000, 10, seed=9)
    for K, iters in ((6, 15), (1, 3)):
        got = diffusion.diffusion_map_embedding(P, K, iters)
        ref = orc.dmap_f32(g.indptr, g.indices, g.data, K, iters)
        ref64 = orc.dmap_f64(g.indptr, g.indices, g.data, K, iters)
        eg, ec = orc.rel_err(got, ref64), orc.rel_err(ref, ref64)
        print(f"\ndiffusion map 30k K={K}: GPU {eg:.2e}, restatement {ec:.2e} vs FP64")
        assert eg <= 2 * ec + 1e-7, K
        G = got.T.astype(np.float64) @ got
        assert np.abs(G - np.eye(K)).max() < 1e-5


def test_diffusion_map_cholesky_qr_and_its_fallbacks(gpu):
    """From 16,384 rows the basis is orthonormalised by a Cholesky QR (FP64
    Gram matrix and factor); more than 22 components take the Gram-Schmidt
    path, and a degenerate basis (here an operator with no nonzeros, whose
    products are exactly zero) is redone with Gram-Schmidt, which turns its
    zero columns into the unit vectors e_c as the reference does."""
    g, P = op(30_000, 10, seed=9)
    for K, iters in ((23, 2), (22, 3)):
        got = diffusion.diffusion_map_embedding(P, K, iters)
        ref = orc.dmap_f32(g.indptr, g.indices, g.data, K, iters)
        ref64 = orc.dmap_f64(g.indptr, g.indices, g.data, K, iters)
        eg, ec = orc.rel_err(got, ref64), orc.rel_err(ref, ref64)
        print(f"\ndiffusion map 30k K={K}: GPU {eg:.2e}, restatement {ec:.2e} vs FP64")
        assert eg <= 2 * ec + 1e-7, K
    n = 20_000
    E = CsrMatrix.from_arrays(np.zeros(n + 1, np.int64), np.zeros(0, np.int64),
                              np.zeros(0, np.float32), (n, n))
    got = diffusion.diffusion_map_embedding(E, 4, 3)
    assert np.array_equal(got, np.eye(n, 4, dtype=np.float32))
    assert np.array_equal(got, orc.dmap_f32(np.zeros(n + 1, np.int64), np.zeros(0, np.int64),
                                            np.zeros(0, np.float32), 4, 3))


@pytest.mark.parametrize("n", [300, 1100])
def test_diffusion_distance(gpu, n):
    g, P = op(n, 6, seed=n, sym=False)
    got = diffusion.diffusion_distance(P, 3)
    ref = orc.diffusion_distance_f32(g.indptr, g.indices, g.data, 3)
    assert np.array_equal(got, ref)


def test_walk_argument_errors(gpu):
    g, P = op(2000, 5)
    R = CsrMatrix.from_arrays([0, 1, 2], [0, 1], np.ones(2, np.float32), (2, 3))
    from scl_amd import _lib
    lib = _lib.load()
    x = np.zeros(2000, np.float32)
    assert lib.scl_diffusion_lazy_random_walk(R.handle, x.ctypes.data, 3, 2, 0.5) == 11
    assert lib.scl_diffusion_lazy_random_walk(P.handle, x.ctypes.data, 1999, 2, 0.5) == 11
    assert lib.scl_diffusion_lazy_random_walk(P.handle, x.ctypes.data, 2000, 2, 1.5) == 10
    assert lib.scl_diffusion_heat_kernel_signature(P.handle, x.ctypes.data, 2000, 0.0, 3) == 10
    seeds = np.array([1], np.int64)
    assert lib.scl_diffusion_random_walk_with_restart(
        P.handle, seeds.ctypes.data, 1, x.ctypes.data, 2000, -0.1, 5, 0.0) == 10
