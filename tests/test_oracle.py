"""Pins the CPU restatement / FP64 oracle (oracle/oracle.c) on the CPU.

Each check is chosen so that a plausible mistake in the oracle fails it:
reference known answers (test_algebra.cpp, test_impute.cpp), closed forms
(ring, complete graph), dense FP64 matrix powers, invariants (constant
columns, pi-weighted means, adjoint identity), finite differences, and
constructed inputs where the accumulation order or a fused multiply-add
changes the FP32 result.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from conftest import csr_from_dense, dense_from_csr
from oracle import oracle as orc
from scl_amd import synth

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(GOLDEN.glob("*.npz"))


def load_fixture(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    fx = load_fixture(path)
    t = int(fx["t"])
    y64 = orc.diffuse_f64(fx["indptr"], fx["indices"], fx["data"],
                          fx["X"].astype(np.float64), t)
    np.testing.assert_allclose(y64, fx["expected_f64"], rtol=0, atol=1e-12 *
                               max(1.0, np.abs(fx["expected_f64"]).max()))
    y32 = orc.diffuse_matrix_f32(fx["indptr"], fx["indices"], fx["data"], fx["X"], t)
    assert orc.rel_err(y32, fx["expected_f64"]) <= 1e-5
    if "expected_f32" in fx:
        assert np.array_equal(y32, fx["expected_f32"])


def test_reference_spmm_known_answers():
    fx = load_fixture(GOLDEN / "algebra_3x3.npz")
    Y = orc.spmm_step_f32(fx["indptr"], fx["indices"], fx["data"], fx["X"])
    assert np.array_equal(Y, np.array([[11, 14], [9, 12], [49, 64]], np.float32))
    y = orc.diffuse_vector_f32(fx["indptr"], fx["indices"], fx["data"], fx["x"], 1)
    assert np.array_equal(y, np.array([7, 6, 32], np.float32))
    # and the alpha/beta combination of test_algebra.cpp:309-325
    assert np.array_equal(2 * y + 3 * np.array([10, 20, 30], np.float32),
                          np.array([44, 72, 154], np.float32))


def test_fp32_chain_is_fma_in_stored_order():
    """Row 0 = [1 at col 0, a at col 1]; X[:,0] = [c, x] with
    a*x = 1 + 2^-11 + 2^-24 and c = -(1 + 2^-11).  The stored-order FMA chain
    gives fma(a, x, c) = 2^-24; an unfused multiply-add or the reverse order
    gives 0."""
    a = np.float32(1 + 2.0 ** -12)
    x = np.float32(1 + 2.0 ** -12)
    c = np.float32(-(1 + 2.0 ** -11))
    indptr = np.array([0, 2, 3], np.int64)
    indices = np.array([0, 1, 1], np.int64)
    data = np.array([1.0, a, 1.0], np.float32)
    X = np.array([[c], [x]], np.float32)
    Y = orc.spmm_step_f32(indptr, indices, data, X)
    assert Y[0, 0] == np.float32(2.0 ** -24)
    v = orc.diffuse_vector_f32(indptr, indices, data, X[:, 0], 1)
    assert v[0] == np.float32(2.0 ** -24)


def test_unsorted_slot_is_accumulated_in_stored_order():
    # -1e8 + 1 + 1e8 = 0 in float (1 is absorbed); 1 - 1e8 + 1e8 = 0 too;
    # 1e8 - 1e8 + 1 = 1: only the stored order [0, 2, 1] gives 1.
    X = np.array([[1e8], [1.0], [-1e8]], np.float32)
    ip = np.array([0, 3, 3, 3], np.int64)
    dv = np.ones(3, np.float32)
    assert orc.spmm_step_f32(ip, np.array([0, 2, 1], np.int64), dv, X)[0, 0] == 1.0
    assert orc.spmm_step_f32(ip, np.array([0, 1, 2], np.int64), dv, X)[0, 0] == 0.0


def _random_csr(rng, n, density, normalize=True, n_cols=None):
    m = n if n_cols is None else n_cols
    A = (rng.random((n, m)) < density) * rng.random((n, m))
    if normalize:
        s = A.sum(axis=1, keepdims=True)
        A = np.divide(A, s, out=np.zeros_like(A), where=s > 0)
    return A.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_brute_force_dense_power(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n, F, t = 50, 7, 4
    A = _random_csr(rng, n, 0.15)
    ip, ix, d = csr_from_dense(A)
    X = rng.standard_normal((n, F)).astype(np.float32)
    ref = np.linalg.matrix_power(A, t) @ X.astype(np.float64)
    y64 = orc.diffuse_f64(ip, ix, d, X, t)
    np.testing.assert_allclose(y64, ref, rtol=1e-12, atol=1e-12)
    y32 = orc.diffuse_matrix_f32(ip, ix, d, X, t)
    assert orc.rel_err(y32, ref) < 1e-5
    yt = orc.diffuse_f64(ip, ix, d, X, t, transpose=True)
    np.testing.assert_allclose(yt, np.linalg.matrix_power(A.T, t) @ X, rtol=1e-12,
                               atol=1e-12)
    # the vector path is the F == 1 case
    v = orc.diffuse_vector_f32(ip, ix, d, X[:, 0], t)
    assert np.array_equal(v, orc.diffuse_matrix_f32(ip, ix, d, X[:, :1], t)[:, 0])


def test_constant_columns_preserved():
    g = synth.uniform_knn(2000, 15, seed=7).normalized()
    X = np.full((2000, 3), 2.5, np.float32)
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X, 5)
    # P rows sum to 1 within 15 fp32 roundings
    assert np.abs(y64 - 2.5).max() < 2.5 * 16 * 2.0 ** -24 * 5
    y32 = orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X, 5)
    assert np.abs(y32 - 2.5).max() < 2.5 * 1e-5


def test_pi_weighted_means_preserved_on_symmetric_graph():
    g = synth.knn_graph(600, 10, seed=3)
    W = dense_from_csr(600, g.indptr, g.indices, g.data)
    assert np.array_equal(W, W.T)  # union-max symmetrisation
    deg = W.sum(axis=1)
    P = W / deg[:, None]
    ip, ix, _ = csr_from_dense(P)
    X = np.random.Generator(np.random.PCG64(5)).standard_normal((600, 4))
    pi = deg / deg.sum()
    vals = P[P != 0]  # row-major order == stored order
    Y = orc.diffuse_f64(ip, ix, vals, X, 6)
    np.testing.assert_allclose(pi @ Y, pi @ X, rtol=1e-11, atol=1e-12)


def test_transition_f32_follows_reference_formula():
    # rows: [1,2,3], empty, tiny (sum <= 1e-15 -> 0), single
    indptr = np.array([0, 3, 3, 5, 6], np.int64)
    indices = np.array([0, 1, 3, 0, 2, 1], np.int64)
    vals = np.array([1, 2, 3, 4e-16, 5e-16, 7], np.float32)
    out = orc.transition_f32(indptr, indices, vals, symmetric=False)
    inv = np.float32(1) / np.float32(6)
    assert np.array_equal(out[:3], np.array([1, 2, 3], np.float32) * inv)
    assert np.array_equal(out[3:5], np.zeros(2, np.float32))
    assert out[5] == np.float32(7) * (np.float32(1) / np.float32(7))
    # symmetric: v *= d_i^-1/2 * d_j^-1/2, all in float
    s = np.array([6, 0, np.float32(4e-16) + np.float32(5e-16), 7], np.float32)
    with np.errstate(divide="ignore"):
        dis = np.where(s > np.float32(1e-15), np.float32(1) / np.sqrt(s), np.float32(0))
    out = orc.transition_f32(indptr, indices, vals, symmetric=True)
    exp = []
    for i in range(4):
        for k in range(indptr[i], indptr[i + 1]):
            exp.append(np.float32(vals[k] * np.float32(dis[i] * dis[indices[k]])))
    assert np.array_equal(out, np.array(exp, np.float32))


def test_transition_f64_formula_and_stochastic_rows():
    g = synth.uniform_knn(500, 15, seed=11)
    out, deg = orc.transition_f64(g.indptr, g.indices, g.data)
    exp = np.empty_like(g.data)
    for i in range(500):
        s = 0.0
        for k in range(g.indptr[i], g.indptr[i + 1]):
            s += float(g.data[k])
        assert deg[i] == s
        for k in range(g.indptr[i], g.indptr[i + 1]):
            exp[k] = np.float32(float(g.data[k]) / s)
    assert np.array_equal(out, exp)
    rs = np.add.reduceat(out.astype(np.float64), g.indptr[:-1])
    assert np.abs(rs - 1).max() < 15 * 2.0 ** -24
    # f32 reference formula differs by a few ulps only
    ref32 = orc.transition_f32(g.indptr, g.indices, g.data)
    assert np.abs(ref32.astype(np.float64) - out).max() / out.max() < 20 * 2.0 ** -24
    # symmetric mode: v * r_i * r_j with r = 1/sqrt(d) in FP64
    sym, deg2 = orc.transition_f64(g.indptr, g.indices, g.data, symmetric=True)
    r = 1 / np.sqrt(deg2)
    rows = np.repeat(np.arange(500), np.diff(g.indptr))
    exp = ((g.data.astype(np.float64) * r[rows]) * r[g.indices]).astype(np.float32)
    assert np.array_equal(sym, exp)


def test_grad_matches_finite_differences_and_adjoint_identity():
    rng = np.random.Generator(np.random.PCG64(9))
    n, F, t = 30, 3, 3
    A = _random_csr(rng, n, 0.2)
    ip, ix, d = csr_from_dense(A)
    vals = d.astype(np.float64)
    X = rng.standard_normal((n, F))
    T = rng.standard_normal((n, F))
    L, Y, dX, dP = orc.grad_f64(ip, ix, vals, X, T, t)
    np.testing.assert_allclose(Y, np.linalg.matrix_power(A, t) @ X, rtol=1e-12, atol=1e-12)
    assert abs(L - 0.5 * ((Y - T) ** 2).sum()) < 1e-10
    eps = 1e-6
    for (i, c) in [(0, 0), (5, 2), (17, 1), (29, 0)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[i, c] += eps
        Xm[i, c] -= eps
        fd = (orc.loss_f64(ip, ix, vals, Xp, T, t) - orc.loss_f64(ip, ix, vals, Xm, T, t)) / (2 * eps)
        assert abs(fd - dX[i, c]) < 1e-6 * max(1, abs(fd))
    for k in [0, len(vals) // 3, len(vals) - 1]:
        vp, vm = vals.copy(), vals.copy()
        vp[k] += eps
        vm[k] -= eps
        fd = (orc.loss_f64(ip, ix, vp, X, T, t) - orc.loss_f64(ip, ix, vm, X, T, t)) / (2 * eps)
        assert abs(fd - dP[k]) < 1e-6 * max(1, abs(fd))
    # <P^t X, G> == <X, (P^T)^t G>
    G = rng.standard_normal((n, F))
    lhs = (orc.diffuse_f64(ip, ix, vals, X, t) * G).sum()
    rhs = (X * orc.diffuse_f64(ip, ix, vals, G, t, transpose=True)).sum()
    assert abs(lhs - rhs) < 1e-10


def test_stationarity_residual():
    g = synth.knn_graph(400, 8, seed=4)
    P, deg = orc.transition_f64(g.indptr, g.indices, g.data)
    r = orc.stationarity_residual_f64(g.indptr, g.indices, P, deg)
    assert r < 1e-6  # fp32 rounding of P only
    # dense FP64 evaluation of the same formula on an asymmetric operator
    u = synth.uniform_knn(300, 6, seed=4)
    Pu, du = orc.transition_f64(u.indptr, u.indices, u.data)
    Pd = dense_from_csr(300, u.indptr, u.indices, Pu)
    pi = du / du.sum()
    exp = np.abs(Pd.T @ pi - pi).sum()
    got = orc.stationarity_residual_f64(u.indptr, u.indices, Pu, du)
    assert got > 1e-3
    assert abs(got - exp) < 1e-12 * max(1, exp)


def test_relative_error_convention():
    a = np.array([1.0, -3.0, 2.0], np.float32)
    b = np.array([1.0, -2.0, 2.0], np.float64)
    assert orc.rel_err(a, b) == pytest.approx(1.0 / 3.0)
    z = np.zeros(3, np.float32)
    assert orc.rel_err(z, np.full(3, 1e-16)) == pytest.approx(1e-16)


def test_bf16_rounding_matches_library_rne():
    import torch
    rng = np.random.Generator(np.random.PCG64(17))
    x = np.concatenate([rng.standard_normal(100_000).astype(np.float32),
                        np.float32([1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0,
                                    np.inf, -np.inf, 3.4e38, 1e-40])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = orc.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    one = [orc.load().orc_bf16_round(float(v)) for v in x[-9:-4]]
    assert one == [1.0, 1 + 2 ** -6, -1.0, 0.0, -0.0]


def test_bf16_restatement_reduces_to_fp32_when_nothing_rounds():
    fx = load_fixture(GOLDEN / "impute_3x3.npz")  # 0.5 weights, small integers
    y = orc.diffuse_bf16_f32(fx["indptr"], fx["indices"], fx["data"], fx["X"], 2,
                             bf16_values=True)
    assert np.array_equal(y, fx["expected_f32"])
    g = synth.uniform_knn(800, 9, seed=2).normalized()
    X = synth.normal(800, 6, seed=2)
    # t = 1: no iterate is rounded
    assert np.array_equal(orc.diffuse_bf16_f32(g.indptr, g.indices, g.data, X, 1),
                          orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X, 1))
    # t = 2 by hand: round the first iterate, then an fp32 step
    mid = orc.bf16_round(orc.spmm_step_f32(g.indptr, g.indices, g.data, X))
    exp = orc.spmm_step_f32(g.indptr, g.indices, g.data, mid)
    assert np.array_equal(orc.diffuse_bf16_f32(g.indptr, g.indices, g.data, X, 2), exp)
    # bf16 values: the same with P rounded first
    pb = orc.bf16_round(g.data)
    mid = orc.bf16_round(orc.spmm_step_f32(g.indptr, g.indices, pb, X))
    exp = orc.spmm_step_f32(g.indptr, g.indices, pb, mid)
    assert np.array_equal(orc.diffuse_bf16_f32(g.indptr, g.indices, g.data, X, 2,
                                               bf16_values=True), exp)
    # and the error against FP64 stays within a few bf16 roundings
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X, 4)
    e = orc.rel_err(orc.diffuse_bf16_f32(g.indptr, g.indices, g.data, X, 4), y64)
    assert 1e-5 < e < 3 * 2.0 ** -8


def test_fp16_rounding_matches_numpy():
    rng = np.random.Generator(np.random.PCG64(19))
    x = np.concatenate([
        rng.standard_normal(20_000).astype(np.float32),
        (rng.standard_normal(5_000) * 1e-6).astype(np.float32),     # subnormal fp16
        (rng.standard_normal(2_000) * 3e4).astype(np.float32),      # near overflow
        np.float32([65504, 65519, 65520, 7e4, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                    1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, -0.0, 0.0])])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    got = orc.fp16_round(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_fp16_restatement_scaling_is_exact_and_reduces():
    g = synth.uniform_knn(600, 8, seed=6).normalized()
    X = synth.normal(600, 4, seed=6)
    # t = 1 or nothing to round: equals the FP32 restatement for any scale
    assert np.array_equal(orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, 1, "fp16", scale_exp=7),
                          orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X, 1))
    # t = 2 by hand with a scale of 2^5
    mid = orc.spmm_step_f32(g.indptr, g.indices, g.data, X)
    mid = (orc.fp16_round(mid * np.float32(32.0)) / np.float32(32.0)).astype(np.float32)
    exp = orc.spmm_step_f32(g.indptr, g.indices, g.data, mid)
    got = orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, 2, "fp16", scale_exp=5)
    assert np.array_equal(got, exp)
    # fp16 iterates are ~8x closer to FP64 than bf16 iterates
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X, 4)
    e16 = orc.rel_err(orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, 4, "fp16",
                                           scale_exp=14), y64)
    eb = orc.rel_err(orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, 4, "bf16"), y64)
    assert e16 < eb / 4 and e16 < 1e-3


def test_rounded_input_restatement_by_hand():
    """round_input = one more rounding, of X itself, before step 1: equal to
    the plain 16-bit restatement run on the rounded X, and to t = 1 done by
    hand."""
    g = synth.uniform_knn(500, 7, seed=8).normalized()
    X = synth.normal(500, 6, seed=8) * np.float32(5.0)
    for fmt, k in (("bf16", 0), ("fp16", 9)):
        s = np.float32(2.0 ** k)
        Xq = (orc.fp16_round(X * s) / s).astype(np.float32) if fmt == "fp16" else orc.bf16_round(X)
        assert not np.array_equal(Xq, X)
        for t in (1, 3):
            got = orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, t, fmt, scale_exp=k,
                                       round_input=True)
            assert np.array_equal(got, orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, Xq,
                                                            t, fmt, scale_exp=k))
        assert np.array_equal(
            orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, 1, fmt, scale_exp=k,
                                 round_input=True),
            orc.spmm_step_f32(g.indptr, g.indices, g.data, Xq))


@pytest.mark.parametrize("symmetric", [False, True])
def test_fp64_normalisation_within_fp32_formula_error(symmetric):
    """The FP64 normalisation (device path) and the reference's FP32 formula
    differ only by the FP32 formula's own rounding: every entry lies within
    the first-order bound of that error.  A wrong term would not: dividing by
    the wrong row's degree moves entries by far more than the bound."""
    g = synth.knn_graph(20_000, 15, seed=31)
    e64, _ = orc.transition_f64(g.indptr, g.indices, g.data, symmetric)
    r32 = orc.transition_f32(g.indptr, g.indices, g.data, symmetric)
    bound = orc.ulps_bound_fp32_transition(g.indptr, g.indices, symmetric)
    ulps = orc.ulp_distance(e64, r32)
    assert np.all(ulps <= bound) and ulps.max() >= 1
    if not symmetric:   # planted mistake: the next row's degree
        lens = np.diff(g.indptr)
        rs = np.add.reduceat(g.data.astype(np.float64), g.indptr[:-1])
        wrong = (g.data / np.repeat(np.roll(rs, -1), lens)).astype(np.float32)
        assert np.any(orc.ulp_distance(wrong, r32) > bound)


def test_ulp_distance_counts_representable_steps():
    a = np.array([1.0, -1.0, 0.0, -0.0, 1.0], np.float32)
    b = np.array([np.nextafter(np.float32(1), np.float32(2)), -1.0, -0.0,
                  np.float32(1e-45), np.nextafter(np.nextafter(np.float32(1), np.float32(0)),
                                                  np.float32(0))], np.float32)
    assert orc.ulp_distance(a, b).tolist() == [1, 0, 0, 1, 2]


def test_compensated_bf16_rounding_properties():
    """orc_comp_round: one rounding to 24 bits (15 mantissa bits, nearest
    even).  Pinned by what a wrong bit position or direction breaks: the low
    8 bits of the result are zero; it is the nearer of the two 24-bit
    neighbours of x (so the error is at most 2^-16 relative, and reaches it);
    exact halves go to the even neighbour; idempotent, odd, monotone; inf
    stays inf, nan stays nan and the largest finite float rounds to inf."""
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        (rng.standard_normal(5000) * 1e20).astype(np.float32),
                        (rng.standard_normal(5000) * 1e-30).astype(np.float32)])
    c = orc.comp_round(x)
    cb = c.view(np.uint32)
    assert np.all(cb & 0xFF == 0)
    ulp = 256.0 * np.abs(np.spacing(c.astype(np.float32))).astype(np.float64)
    d = np.abs(c.astype(np.float64) - x)
    assert np.all(d <= ulp / 2 * (1 + 1e-12))       # nearest 24-bit value
    rel = d / np.abs(x.astype(np.float64))
    normal = np.abs(x) > 1e-37
    assert rel[normal].max() <= 2.0 ** -16 and rel[normal].max() > 2.0 ** -17
    # exact halves between two 24-bit values go to the even one
    lo = (rng.integers(0x3F800000, 0x40000000, 1000, dtype=np.int64) & ~0xFF).astype(np.uint32)
    half = (lo | 0x80).view(np.float32)
    got = orc.comp_round(half).view(np.uint32)
    even = np.where((lo >> 8) & 1, lo + 0x100, lo).astype(np.uint32)
    assert np.array_equal(got, even)
    assert np.array_equal(orc.comp_round(c), c)
    assert np.array_equal(orc.comp_round(-x), -c)
    xs = np.sort(x)
    assert np.all(np.diff(orc.comp_round(xs).astype(np.float64)) >= 0)
    special = np.array([np.inf, -np.inf, np.finfo(np.float32).max], np.float32)
    assert np.array_equal(orc.comp_round(special), np.array([np.inf, -np.inf, np.inf],
                                                            np.float32))
    assert np.isnan(orc.comp_round(np.array([np.nan], np.float32)))[0]


def test_compensated_restatement_reduces_and_beats_bf16():
    g = synth.uniform_knn(800, 9, seed=3).normalized()
    X = synth.normal(800, 8, seed=3)
    ip, ix, p = g.indptr, g.indices, g.data
    # t = 2 by hand, values and input compensated
    pc, Xc = orc.comp_round(p), orc.comp_round(X)
    mid = orc.comp_round(orc.spmm_step_f32(ip, ix, pc, Xc))
    exp = orc.spmm_step_f32(ip, ix, pc, mid)
    got = orc.diffuse_lowp_f32(ip, ix, p, X, 2, "comp", bf16_values=True, round_input=True)
    assert np.array_equal(got, exp)
    y64 = orc.diffuse_f64(ip, ix, p, X, 4)
    ec = orc.rel_err(orc.diffuse_lowp_f32(ip, ix, p, X, 4, "comp", bf16_values=True,
                                          round_input=True), y64)
    eb = orc.rel_err(orc.diffuse_lowp_f32(ip, ix, p, X, 4, "bf16", bf16_values=True,
                                          round_input=True), y64)
    assert ec < 2.0 ** -13 < eb
