"""GPU parity of the FP32 diffusion path against the CPU restatement.

Bar: the HIP kernels accumulate every output element as the same fp32 FMA
chain in stored order as the reference (scl/kernel/diffusion.hpp:249-256 with
axpy_simd 100-134), so results must be BIT-IDENTICAL to the restatement
(oracle/oracle.c); the error against the FP64 oracle is checked to be
<= 1e-5 in the reference's relative convention and printed.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import scl_amd
from oracle import oracle as orc
from scl_amd import CscMatrix, CsrMatrix, _lib, diffusion, synth

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(GOLDEN.glob("*.npz"))


def fixture(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def gpu_diffuse(ip, ix, d, X, t, n=None):
    n = X.shape[0] if n is None else n
    P = CsrMatrix.from_arrays(ip, ix, d, (n, n))
    Y = np.ascontiguousarray(X, np.float32).copy()
    diffusion.diffuse_matrix(P, Y, t)
    return Y


@pytest.mark.parametrize("path", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_golden_fixtures(gpu, path):
    fx = fixture(path)
    t = int(fx["t"])
    Y = gpu_diffuse(fx["indptr"], fx["indices"], fx["data"], fx["X"], t)
    ref = orc.diffuse_matrix_f32(fx["indptr"], fx["indices"], fx["data"], fx["X"], t)
    assert np.array_equal(Y, ref)
    if "expected_f32" in fx:
        assert np.array_equal(Y, fx["expected_f32"])
    assert orc.rel_err(Y, fx["expected_f64"]) <= 1e-5


def test_reference_known_answers(gpu):
    fx = fixture(GOLDEN / "algebra_3x3.npz")
    P = CsrMatrix.from_arrays(fx["indptr"], fx["indices"], fx["data"], (3, 3))
    Y = fx["X"].copy()
    diffusion.diffuse_matrix(P, Y, 1)
    assert np.array_equal(Y, [[11, 14], [9, 12], [49, 64]])
    x = fx["x"].copy()
    diffusion.diffuse_vector(P, x, 1)
    assert np.array_equal(x, [7, 6, 32])


def test_uniform15_10k_f50_t3(gpu):
    """n=10,000, k=15 uniform neighbours, F=50, t=3 (the CPU-reference case)."""
    g = synth.uniform_knn(10_000, 15, seed=1).normalized()
    X = synth.normal(10_000, 50, seed=1)
    Y = gpu_diffuse(g.indptr, g.indices, g.data, X, 3)
    ref = orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X, 3)
    assert np.array_equal(Y, ref)
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X, 3)
    err, med = orc.rel_err(Y, y64), orc.median_rel_err(Y, y64)
    print(f"\nuniform15_10k: rel err vs CPU 0 (bit-identical); vs FP64 max {err:.3e} median {med:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("F", [1, 2, 3, 5, 8, 31, 50, 64, 100, 128, 256, 257, 1000, 2000])
def test_feature_widths(gpu, F):
    g = synth.uniform_knn(3000, 7, seed=F).normalized()
    X = synth.normal(3000, F, seed=F)
    Y = gpu_diffuse(g.indptr, g.indices, g.data, X, 2)
    assert np.array_equal(Y, orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X, 2))


def test_magic_impute_default_steps_and_vector(gpu):
    g = synth.uniform_knn(2000, 10, seed=3).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (2000, 2000))
    X = synth.normal(2000, 16, seed=3)
    Y = X.copy()
    diffusion.magic_impute(P, Y, 0)  # t == 0 means 3
    assert np.array_equal(Y, orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X, 3))
    Z = X.copy()
    diffusion.magic_impute(P, Z, -2)  # negative: unchanged
    assert np.array_equal(Z, X)
    x = X[:, 0].copy()
    diffusion.diffuse_vector(P, x, 4)
    assert np.array_equal(x, orc.diffuse_vector_f32(g.indptr, g.indices, g.data, X[:, 0], 4))


def test_ragged_rows_empty_long_unsorted(gpu):
    rng = np.random.Generator(np.random.PCG64(8))
    n = 5000
    lens = rng.integers(0, 4, n)
    lens[0] = n            # one row touching every node
    lens[1:50] = 0         # a run of empty rows
    ip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ix = np.concatenate([rng.permutation(n)[:l] for l in lens]).astype(np.int64)
    d = rng.random(ip[-1]).astype(np.float32)  # unsorted slots, unnormalised
    X = synth.normal(n, 24, seed=8)
    Y = gpu_diffuse(ip, ix, d, X, 3)
    assert np.array_equal(Y, orc.diffuse_matrix_f32(ip, ix, d, X, 3))
    assert np.all(Y[1:50] == 0)


@pytest.mark.parametrize("n", [5000, 20_000])
def test_single_feature_rows_of_every_length(gpu, n):
    """F = 1 runs one lane per row, 64 rows per wave, the rows' entries
    staged through LDS in windows: rows of 0 to 40 nonzeros at every offset
    within a wave's window, with and without the locality schedule (from
    16,384 rows)."""
    rng = np.random.Generator(np.random.PCG64(n))
    lens = np.arange(n) % 41
    rng.shuffle(lens)
    ip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ix = np.concatenate([np.sort(rng.integers(0, n, l)) for l in lens]).astype(np.int64)
    d = rng.random(ip[-1]).astype(np.float32)
    x = synth.normal(n, 1, seed=n)
    assert np.array_equal(gpu_diffuse(ip, ix, d, x, 3), orc.diffuse_matrix_f32(ip, ix, d, x, 3))
    P = CsrMatrix.from_arrays(ip, ix, d, (n, n))
    got = diffusion.diffuse_vector(P, x[:, 0].copy(), 4)
    assert np.array_equal(got, orc.diffuse_vector_f32(ip, ix, d, x[:, 0], 4))


@pytest.mark.parametrize("unroll", [4, 8, 16])
def test_single_feature_windows(gpu, unroll):
    """Windows of 1,024, 2,048 and 4,096 entries (unroll 4, 8, 16): rows
    longer than a window (up to 5,000 entries) continue their chain across
    windows; a wave of 64 rows holds several windows; a run of empty rows;
    n not a multiple of 64 leaves the last wave partial; the transposed
    operator runs the same kernel."""
    from scl_amd import torch_ops
    import torch
    n = 6_001
    rng = np.random.Generator(np.random.PCG64(unroll))
    lens = rng.integers(0, 60, n)
    lens[[3, 64, 700, 701, 6000]] = [5000, 1500, 2049, 4097, 1025]
    lens[100:200] = 0
    ip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ix = np.concatenate([rng.integers(0, n, l) for l in lens]).astype(np.int64)
    d = rng.random(ip[-1]).astype(np.float32)
    x = synth.normal(n, 1, seed=unroll)
    P = CsrMatrix.from_arrays(ip, ix, d, (n, n))
    P.prepare_device()
    for tr in (False, True):
        torch_ops.set_shape(P, 1, vec=1, group=1, unroll=unroll, transpose=tr)
        got = torch_ops.spmm(P, torch.from_numpy(x).cuda(), transpose=tr).cpu().numpy()
        if tr:
            tp, ti, td = P.transpose().export()
            ref = orc.spmm_step_f32(tp, ti, td, x)
        else:
            ref = orc.spmm_step_f32(ip, ix, d, x)
            assert np.all(got[100:200] == 0)
        assert np.array_equal(got, ref), (unroll, tr)


@pytest.mark.parametrize("F", [260, 300, 1000])
def test_slab_planar_iterates(gpu, F):
    """Wide operators keep the iterates between the first and the last step
    slab-planar when F floats per row leave rows off the 128-byte lines (the
    row-grouped default shape): every step's output is still the stored-order
    chain, so Y equals the restatement for t = 1..4, through the device entry
    point and the host-array drop-in, with a partial last slab of 4, 44 and
    232 features."""
    import torch
    from scl_amd import torch_ops
    g = synth.knn_graph(20_000, 10, seed=F).normalized()
    n = g.n
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    X = synth.normal(n, F, seed=F)
    Xd = torch.from_numpy(X).cuda()
    cols = np.unique(np.r_[0, 1, 127, 128, 255, 256, F - 1, F // 2])
    for t in (1, 2, 3, 4):
        Y = torch_ops.diffuse(P, Xd, t).cpu().numpy()
        assert torch_ops.get_shape(P, F)["rows_per_group"] == 4
        ref = orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X[:, cols], t)
        assert np.array_equal(Y[:, cols], ref), t
    Yh = X.copy()
    diffusion.diffuse_matrix(P, Yh, 3)
    assert np.array_equal(Yh, torch_ops.diffuse(P, Xd, 3).cpu().numpy())


def test_no_nonzeros_and_single_node(gpu):
    Y = gpu_diffuse(np.zeros(6, np.int64), np.zeros(0, np.int64),
                    np.zeros(0, np.float32), np.ones((5, 3), np.float32), 2)
    assert np.array_equal(Y, np.zeros((5, 3), np.float32))
    Y = gpu_diffuse(np.array([0, 1]), np.array([0]), np.array([0.5], np.float32),
                    np.full((1, 7), 4.0, np.float32), 3)
    assert np.array_equal(Y, np.full((1, 7), 0.5, np.float32))


def test_csc_handle_applies_transpose(gpu):
    """visit() on a CSC handle walks columns as rows (core/internal.hpp:
    133-151), so the diffusion applies P^T."""
    g = synth.uniform_knn(4000, 9, seed=12).normalized()
    X = synth.normal(4000, 12, seed=12)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (4000, 4000))
    T = P.transpose()
    Y = X.copy()
    diffusion.diffuse_matrix(T, Y, 2)
    tp, ti, td = T.export()
    assert np.array_equal(Y, orc.diffuse_matrix_f32(tp, ti, td, X, 2))
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X, 2, transpose=True)
    assert orc.rel_err(Y, y64) <= 1e-5
    # a CSC handle built from the same arrays is the same operator
    C = CscMatrix.from_arrays(tp, ti, td, (4000, 4000))
    Z = X.copy()
    diffusion.diffuse_matrix(C, Z, 2)
    assert np.array_equal(Z, Y)


def test_device_api_matches_drop_in(gpu):
    import torch
    from scl_amd import torch_ops
    g = synth.uniform_knn(6000, 12, seed=6).normalized()
    X = synth.normal(6000, 40, seed=6)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (6000, 6000))
    Xd = torch.from_numpy(X).cuda()
    for t in (1, 2, 3, 4):
        Yd = torch_ops.diffuse(P, Xd, t).cpu().numpy()
        Y = X.copy()
        diffusion.diffuse_matrix(P, Y, t)
        assert np.array_equal(Yd, Y), t
    Y1 = torch_ops.spmm(P, Xd).cpu().numpy()
    assert np.array_equal(Y1, orc.spmm_step_f32(g.indptr, g.indices, g.data, X))
    Yt = torch_ops.spmm(P, Xd, transpose=True).cpu().numpy()
    T = P.transpose()
    tp, ti, td = T.export()
    assert np.array_equal(Yt, orc.spmm_step_f32(tp, ti, td, X))
    assert torch.equal(Xd.cpu(), torch.from_numpy(X))


def test_drop_in_wide_host_matrix(gpu):
    """A 288 MB host X (copied through the pinned bounce buffers in chunks):
    bit-identical to the device API for even and odd t, under a tuned
    row-grouped shape and with MAGIC's default t, and to the oracle on
    sampled columns."""
    import torch
    from scl_amd import torch_ops
    n, F = 120_000, 600
    g = synth.knn_graph(n, 10, seed=31).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    X = synth.normal(n, F, seed=31)
    Xd = torch.from_numpy(X).cuda()
    for t in (2, 3):
        Y = X.copy()
        diffusion.diffuse_matrix(P, Y, t)
        assert np.array_equal(Y, torch_ops.diffuse(P, Xd, t).cpu().numpy()), t
    cols = np.array([0, 255, 256, 511, 512, 599])
    assert np.array_equal(Y[:, cols], orc.diffuse_matrix_f32(g.indptr, g.indices, g.data,
                                                             X[:, cols], 3))
    torch_ops.set_shape(P, F, 4, 64, 16, nontemporal=True, rows_per_group=4)
    Z = X.copy()
    diffusion.diffuse_matrix(P, Z, 3)
    assert np.array_equal(Z, Y)
    M = X.copy()
    diffusion.magic_impute(P, M, 0)
    assert np.array_equal(M, Y)


@pytest.mark.parametrize("symmetric", [False, True])
def test_transition_matrix_fp64(gpu, symmetric):
    g = synth.knn_graph(5000, 10, seed=21)
    exp, deg = orc.transition_f64(g.indptr, g.indices, g.data, symmetric)
    data = g.data.copy()
    A = CsrMatrix.wrap(g.indptr, g.indices, data, (5000, 5000))
    got_deg, res = diffusion.compute_transition_matrix_ex(A, symmetric)
    assert np.array_equal(data, exp)          # in place, in the caller's array
    assert np.array_equal(got_deg, deg)
    ref_res = orc.stationarity_residual_f64(g.indptr, g.indices, exp, deg)
    assert abs(res - ref_res) <= 1e-12 * max(1.0, ref_res) + 1e-15
    if not symmetric:
        print(f"\nknn 5000: FP64 stationarity residual {res:.3e}")
        assert res < 1e-5
        rs = np.add.reduceat(data.astype(np.float64), g.indptr[:-1])
        assert np.abs(rs - 1).max() < 64 * 2.0 ** -24
    # diffusion through the normalised wrapped handle sees the new values
    X = synth.normal(5000, 8, seed=2)
    Y = X.copy()
    diffusion.diffuse_matrix(A, Y, 2)
    assert np.array_equal(Y, orc.diffuse_matrix_f32(g.indptr, g.indices, exp, X, 2))
    # the reference's FP32 formula (fp32 row sums, diffusion.hpp:452-603)
    # carries its own rounding: each entry lies within the first-order bound
    # of that formula's error of the FP64 result (ulps_bound_fp32_transition)
    ref32 = orc.transition_f32(g.indptr, g.indices, g.data, symmetric)
    ulps = orc.ulp_distance(ref32, data)
    bound = orc.ulps_bound_fp32_transition(g.indptr, g.indices, symmetric)
    print(f"\nknn 5000 symmetric={symmetric}: ulp distance to the reference FP32 formula "
          f"max {ulps.max()}, mean {ulps.mean():.2f} (bound per entry {bound.min()}..{bound.max()})")
    assert np.all(ulps <= bound)


def test_transition_matrix_rectangular(gpu):
    """A rectangular adjacency normalises its rows (symmetric == 0), bit-
    identical to the FP64 restatement; the stationarity residual and the
    symmetric mode need a square matrix (SCL_ERROR_DIMENSION_MISMATCH)."""
    rng = np.random.default_rng(5)
    n, m = 300, 700
    ip = np.concatenate([[0], np.cumsum(rng.integers(0, 9, n))]).astype(np.int64)
    ix = np.concatenate([np.sort(rng.choice(m, ip[i + 1] - ip[i], replace=False))
                         for i in range(n)]).astype(np.int64)
    d = rng.random(len(ix)).astype(np.float32)
    A = CsrMatrix.from_arrays(ip, ix, d, (n, m))
    for sym, res in ((False, True), (True, False)):
        with pytest.raises(scl_amd.SCLError) as e:
            diffusion.compute_transition_matrix_ex(A, sym, residual=res)
        assert e.value.code == _lib.SCL_ERROR_DIMENSION_MISMATCH
    deg, res = diffusion.compute_transition_matrix_ex(A, False, residual=False)
    exp, exp_deg = orc.transition_f64(ip, ix, d, False)
    assert res is None
    assert np.array_equal(A.export()[2], exp) and np.array_equal(deg, exp_deg)


def test_dimension_errors_with_device(gpu):
    g = synth.uniform_knn(100, 5, seed=1).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (100, 100))
    X = np.ones((99, 4), np.float32)
    with pytest.raises(scl_amd.SCLError) as e:
        diffusion.diffuse_matrix(P, X, 1)
    assert e.value.code == _lib.SCL_ERROR_DIMENSION_MISMATCH
    assert np.all(X == 1)


def test_knn30_200k_f2000_t8_column_sample(gpu):
    """The at-scale workload: n=200,000 geometric kNN (k=30, union-max),
    F=2000 log1p(Poisson), t=8, run on the GPU at full width; columns never
    mix, so 8 sampled columns are checked bit for bit against the
    restatement run on those columns alone, plus the FP64 error on them."""
    import torch
    from scl_amd import torch_ops
    n, F, t = 200_000, 2000, 8
    g = synth.knn_graph(n, 30, seed=2)
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    diffusion.compute_transition_matrix(A)
    ip, ix, p = A.export()
    X = synth.log1p_poisson(n, F, seed=2)
    Yd = torch_ops.diffuse(A, torch.from_numpy(X).cuda(), t)
    cols = np.array([0, 1, 333, 999, 1000, 1499, 1998, 1999])
    Ys = Yd[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    Xs = synth.log1p_poisson(n, F, seed=2, cols=cols)
    assert np.array_equal(Xs, X[:, cols])
    ref = orc.diffuse_matrix_f32(ip, ix, p, Xs, t)
    assert np.array_equal(Ys, ref)
    y64 = orc.diffuse_f64(ip, ix, p, Xs, t)
    err = orc.rel_err(Ys, y64)
    print(f"\nknn30_200k: nnz {g.nnz}, sampled columns bit-identical to CPU; "
          f"vs FP64 max {err:.3e} median {orc.median_rel_err(Ys, y64):.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("F", [64, 128, 200, 2000])
def test_every_launch_shape_is_bit_identical(gpu, F):
    """All launch shapes (row-split, scalar multi-row, both XCD mappings,
    with and without the locality schedule) reproduce the restatement."""
    import torch
    from scl_amd import torch_ops
    n = 20_000
    g = synth.knn_graph(n, 12, seed=F)
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    diffusion.compute_transition_matrix(A)
    ip, ix, p = A.export()
    X = synth.normal(n, F, seed=F)
    ref = orc.spmm_step_f32(ip, ix, p, X)
    Xd = torch.from_numpy(X).cuda()
    vmax = 4 if F % 4 == 0 else (2 if F % 2 == 0 else 1)
    for vec in sorted({1, vmax}):
        for group in (8, 16, 32, 64):
            for sr in (False, True):
                for xr in (False, True):
                    for u in (4, 16):
                        torch_ops.set_shape(A, F, vec, group, u, nontemporal=u == 16,
                                            schedule=not xr, xcd_rows=xr, scalar_rows=sr,
                                            wg_per_cu=4 if u == 4 else 0)
                        assert torch_ops.get_shape(A, F)["wg_per_cu"] == (4 if u == 4 else 0)
                        Y = torch_ops.spmm(A, Xd).cpu().numpy()
                        assert np.array_equal(Y, ref), (vec, group, sr, xr, u)


@pytest.mark.parametrize("F", [1, 50, 256, 2000])
def test_row_grouped_shapes_are_bit_identical(gpu, F):
    """Grouped form (4 or 8 rows per wave over the union of their columns):
    every row keeps its sorted-order FMA chain, so the result is the
    restatement's bit for bit -- with and without the locality schedule,
    through t steps, and after the operator's values change in place."""
    import torch
    from scl_amd import torch_ops
    for n in (20_000, 3_001):              # with a schedule; without, ragged end
        g = synth.knn_graph(n, 12, seed=F + n)
        A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
        ip, ix, p = A.export()
        X = synth.normal(n, F, seed=F)
        Xd = torch.from_numpy(X).cuda()
        vmax = 4 if F % 4 == 0 else (2 if F % 2 == 0 else 1)
        # chunks of 8 or 16 entries: up to 128 values per chunk with 8 rows
        for R in (4, 8):
            for u in (8, 16):
                for vec in sorted({1, vmax}):
                    torch_ops.set_shape(A, F, vec, 64, u, rows_per_group=R,
                                        nontemporal=vec == 1, wg_per_cu=4 if R == 8 else 0)
                    assert torch_ops.get_shape(A, F)["rows_per_group"] == R
                    Y = torch_ops.spmm(A, Xd).cpu().numpy()
                    assert np.array_equal(Y, orc.spmm_step_f32(ip, ix, p, X)), (n, R, u, vec)
        # values change in place (normalisation): the grouped values follow
        diffusion.compute_transition_matrix(A)
        ip, ix, p = A.export()
        Y = torch_ops.diffuse(A, Xd, 3).cpu().numpy()
        assert torch_ops.get_shape(A, F)["rows_per_group"] == 8
        assert np.array_equal(Y, orc.diffuse_matrix_f32(ip, ix, p, X, 3))


def test_row_grouped_needs_sorted_rows(gpu):
    from scl_amd import torch_ops
    ip = np.array([0, 2, 3], np.int64)
    A = CsrMatrix.from_arrays(ip, np.array([1, 0, 1], np.int64),
                              np.ones(3, np.float32), (2, 2))   # row 0 unsorted
    with pytest.raises(Exception):
        torch_ops.set_shape(A, 8, 4, 64, 8, rows_per_group=4)
    B = CsrMatrix.from_arrays(ip, np.array([0, 1, 1], np.int64),
                              np.ones(3, np.float32), (2, 2))
    torch_ops.set_shape(B, 8, 4, 64, 8, rows_per_group=4)
    with pytest.raises(Exception):   # forward direction only
        torch_ops.set_shape(B, 8, 4, 64, 8, rows_per_group=4, transpose=True)


@pytest.mark.parametrize("F", [4, 260, 2000])
def test_staged_groups_are_bit_identical(gpu, F):
    """LDS-staged groups of 64 rows: the union of a group's columns is
    staged through LDS and each row reads its neighbours there in sorted
    order, so the result is the restatement's bit for bit -- with and
    without the locality schedule, for rows longer than one 64-entry window,
    through t steps, and after the values change in place."""
    import torch
    from scl_amd import torch_ops
    cases = [(20_000, 12), (3_001, 12), (1_500, 150)]   # schedule; none; long rows
    for n, k in cases:
        g = synth.knn_graph(n, k, seed=F + n) if k < 100 else synth.uniform_knn(n, k, seed=F)
        A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
        ip, ix, p = A.export()
        X = synth.normal(n, F, seed=F)
        Xd = torch.from_numpy(X).cuda()
        ref = orc.spmm_step_f32(ip, ix, p, X)
        for nt, xr, u in ((False, False, 4), (True, True, 8), (True, False, 16)):
            torch_ops.set_shape(A, F, 4, 64, u, rows_per_group=64, nontemporal=nt,
                                xcd_rows=xr)
            assert torch_ops.get_shape(A, F)["rows_per_group"] == 64
            Y = torch_ops.spmm(A, Xd).cpu().numpy()
            assert np.array_equal(Y, ref), (n, k, nt, xr)
        # values change in place: the staged form reads the operator's values
        diffusion.compute_transition_matrix(A)
        ip, ix, p = A.export()
        Y = torch_ops.diffuse(A, Xd, 3).cpu().numpy()
        assert torch_ops.get_shape(A, F)["rows_per_group"] == 64
        assert np.array_equal(Y, orc.diffuse_matrix_f32(ip, ix, p, X, 3)), (n, k)


def test_staged_groups_with_other_values_and_errors(gpu):
    """MAGIC runs the operator's pattern with its own weights: under a staged
    shape the kernel reads those weights, bit-identical to the restatement.
    The staged form needs sorted rows, vec 4 and the forward direction."""
    from scl_amd import callers, torch_ops
    n, G = 20_000, 64
    g = synth.knn_graph(n, 10, seed=23)
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    torch_ops.set_shape(A, G, 4, 64, 8, rows_per_group=64)
    X = synth.normal(n, G, seed=24)
    ip = np.arange(0, n * G + 1, G, dtype=np.int64)
    ix = np.tile(np.arange(G, dtype=np.int64), n)
    M = CsrMatrix.from_arrays(ip, ix, X.ravel(), (n, G))
    got = callers.impute_magic(M, A, n, G, 2)
    assert np.array_equal(got, orc.magic_f32(g.indptr, g.indices, g.data, X, 2))
    bad = CsrMatrix.from_arrays(np.array([0, 2, 3], np.int64), np.array([1, 0, 1], np.int64),
                                np.ones(3, np.float32), (2, 2))
    with pytest.raises(Exception):
        torch_ops.set_shape(bad, 8, 4, 64, 8, rows_per_group=64)
    with pytest.raises(Exception):
        torch_ops.set_shape(A, 8, 2, 64, 8, rows_per_group=64)
    with pytest.raises(Exception):
        torch_ops.set_shape(A, 8, 4, 64, 8, rows_per_group=64, transpose=True)
