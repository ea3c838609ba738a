"""GPU parity of the reduced-precision diffusion (bf16, fp16 with a
power-of-two scale, and compensated bf16: bf16 plus an 8-bit residual)
against its CPU restatement (oracle orc_diffuse_lowp_f32: the same fp32 FMA
chains in stored order, the same rounding of every stored value), bit for
bit, plus the error against FP64 stated for the uniform15 workload shape.
The precision bar is 1e-3 relative to FP64 (max-norm convention)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scl_amd import CsrMatrix, synth, torch_ops

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("F,t,vb", [(2, 3, False), (8, 2, True), (50, 4, False),
                                    (64, 4, True), (256, 4, False), (256, 4, True),
                                    (300, 3, False), (1000, 2, True), (128, 1, True)])
def test_bf16_path_matches_restatement(gpu, F, t, vb):
    n = 20_000 if F <= 256 else 5_000
    g = synth.uniform_knn(n, 15, seed=F + t).normalized()
    X = synth.normal(n, F, seed=F)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Y = torch_ops.diffuse_bf16(P, torch.from_numpy(X).cuda(), t, bf16_values=vb)
    ref = orc.diffuse_bf16_f32(g.indptr, g.indices, g.data, X, t, bf16_values=vb)
    assert np.array_equal(Y.cpu().numpy(), ref)


@pytest.mark.parametrize("F,t,vb", [(2, 3, False), (64, 4, True), (256, 4, False),
                                    (300, 3, True), (1000, 2, False)])
def test_fp16_path_matches_restatement(gpu, F, t, vb):
    n = 20_000 if F <= 256 else 5_000
    g = synth.uniform_knn(n, 15, seed=F + t + 1).normalized()
    X = synth.normal(n, F, seed=F + 1) * np.float32(37.0)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd = torch.from_numpy(X).cuda()
    k = torch_ops.lowp_scale_exp(P, Xd, t)
    # 2^k * max|X| stays within 2^15 and is not needlessly small
    assert 2.0 ** 14 < 2.0 ** k * np.abs(X).max() * 1.0000001 <= 2.0 ** 15
    Y = torch_ops.diffuse_lowp(P, Xd, t, "fp16", bf16_values=vb, scale_exp=k)
    ref = orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, t, "fp16",
                               bf16_values=vb, scale_exp=k)
    assert np.array_equal(Y.cpu().numpy(), ref)


@pytest.mark.parametrize("fmt,F,t,vb", [("bf16", 2, 1, False), ("bf16", 64, 3, True),
                                        ("fp16", 256, 4, False), ("fp16", 300, 1, True),
                                        ("fp16", 1000, 2, False)])
def test_rounded_input_matches_restatement(gpu, fmt, F, t, vb):
    """X stored in 16 bits before the first step: bit-identical to the
    restatement, also for t = 1 and for F not a multiple of 4."""
    n = 20_000 if F <= 256 else 5_000
    g = synth.uniform_knn(n, 15, seed=F + t + 2).normalized()
    X = synth.normal(n, F, seed=F + 2) * np.float32(11.0)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd = torch.from_numpy(X).cuda()
    k = torch_ops.lowp_scale_exp(P, Xd, t) if fmt == "fp16" else 0
    Y = torch_ops.diffuse_lowp(P, Xd, t, fmt, bf16_values=vb, scale_exp=k, round_input=True)
    ref = orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, t, fmt, bf16_values=vb,
                               scale_exp=k, round_input=True)
    assert np.array_equal(Y.cpu().numpy(), ref)


@pytest.mark.parametrize("F,t,vb,ri", [(4, 1, False, False), (8, 2, True, False),
                                       (12, 3, True, True), (64, 4, False, True),
                                       (256, 4, True, True), (256, 3, False, False),
                                       (300, 2, True, True), (1000, 4, True, False),
                                       (272, 2, True, True), (512, 3, False, True)])
def test_compensated_path_matches_restatement(gpu, F, t, vb, ri):
    """Compensated bf16 iterates (and X with round_input, P's values with
    bf16_values): bit-identical to the restatement, for 16 values per lane
    (F a multiple of 16 from 256, 272 ending inside a slab), 8 (F a multiple
    of 8) and 4 (F a multiple of 4 only), t = 1 included."""
    n = 20_000 if F <= 256 else 5_000
    g = synth.uniform_knn(n, 15, seed=F + t + 3).normalized()
    X = synth.normal(n, F, seed=F + 3) * np.float32(5.0)
    X[0, 0], X[1, 0] = np.float32(1e-35), np.float32(-3e38)   # tiny and huge
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Y = torch_ops.diffuse_lowp(P, torch.from_numpy(X).cuda(), t, "comp", bf16_values=vb,
                               round_input=ri)
    ref = orc.diffuse_lowp_f32(g.indptr, g.indices, g.data, X, t, "comp", bf16_values=vb,
                               round_input=ri)
    assert np.array_equal(Y.cpu().numpy(), ref, equal_nan=True)


def test_bf16_error_vs_fp64_uniform15(gpu):
    """n = 200,000 (the 1M-cell shape scaled down 5x), k = 15 uniform
    neighbours, F = 256 N(0,1), t = 4."""
    n, F, t = 200_000, 256, 4
    g = synth.uniform_knn(n, 15, seed=4).normalized()
    X = synth.normal(n, F, seed=4)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd = torch.from_numpy(X).cuda()
    cols = np.arange(0, F, 16)
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X[:, cols], t)
    y64q = orc.diffuse_f64(g.indptr, g.indices, g.data, orc.bf16_round(X[:, cols]), t)
    rows = []
    for fmt, vb, ri in (("bf16", False, False), ("bf16", True, False), ("fp16", False, False),
                        ("fp16", True, False), ("fp16", False, True), ("comp", True, True),
                        ("bf16", True, True)):
        Y = torch_ops.diffuse_lowp(P, Xd, t, fmt, bf16_values=vb,
                                   round_input=ri)[:, torch.from_numpy(cols).cuda()]
        Yh = Y.cpu().numpy()
        rows.append((fmt + (" rounded input" if ri else ""), vb, orc.rel_err(Yh, y64),
                     orc.median_rel_err(Yh, y64)))
    Y32 = torch_ops.diffuse(P, Xd, t)[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    e32 = orc.rel_err(Y32, y64)
    Xq = torch.from_numpy(orc.bf16_round(X)).cuda()
    Yq = torch_ops.diffuse_bf16(P, Xq, t)[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    eq = orc.rel_err(Yq, y64q)
    print(f"\nfp32 kernel vs FP64: {e32:.2e}")
    for fmt, vb, emax, emed in rows:
        print(f"{fmt} iterates, {'bf16' if vb else 'fp32'} values: max {emax:.2e} median {emed:.2e}")
    print(f"bf16 iterates on bf16-rounded input vs FP64 of that input: {eq:.2e}")
    assert e32 <= 1e-5
    # P values and X stored in compensated bf16 meet the 1e-3 bar with a
    # wide margin; fp16 iterates with fp32 values meet it too, also with X
    # in fp16; plain bf16 storage does not (its rows are reported above)
    assert rows[5][2] <= 1e-4
    assert rows[2][2] <= 1e-3
    assert rows[4][2] <= 1e-3
    assert rows[6][2] > 1e-3
