"""Checks at the full sizes of the differentiation, 16-bit and batching
workloads, through properties that need no CPU run at that size: columns
never mix, so sampled columns are compared with the restatement and the
FP64 oracle run on those columns alone; the loss is homogeneous of degree t
in P's values, so sum_k dP_k P_k = t <Y - T, Y> (Euler) checks the whole
per-nonzero gradient at once; graphs of a batch never mix, so each sampled
graph is diffused alone and compared bit for bit.

The operators are uniform random kNN graphs (k distinct neighbours per
row); the properties do not depend on the graph's geometry, and the
geometric generator of the headline workload would dominate the test time
at these sizes."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scl_amd import CsrMatrix, autograd, batch, synth, torch_ops

pytestmark = pytest.mark.gpu

COLS = np.array([0, 1, 17, 63, 64, 100, 126, 127])


def test_differentiation_500k_f128_t6(gpu):
    n, k, F, t = 500_000, 30, 128, 6
    g = synth.uniform_knn(n, k, seed=3).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    X = synth.normal(n, F, seed=3)
    T = synth.normal(n, F, seed=103)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev).requires_grad_(True)
    Td = torch.from_numpy(T).to(dev)
    vals = torch.from_numpy(g.data.copy()).to(dev).requires_grad_(True)
    Y = autograd.diffuse(P, Xd, t, values=vals, checkpoint_every=2)
    (0.5 * ((Y - Td) ** 2).sum()).backward()
    Yh = Y.detach().cpu().numpy()
    # forward: sampled columns bit-identical to the restatement, FP64 error
    assert np.array_equal(Yh[:, COLS],
                          orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, X[:, COLS], t))
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X[:, COLS], t)
    assert orc.rel_err(Yh[:, COLS], y64) <= 1e-5
    # dX: the sampled columns against the FP64 adjoint of the FP64 residual
    dx64 = orc.diffuse_f64(g.indptr, g.indices, g.data, y64 - T[:, COLS], t, transpose=True)
    ex = orc.rel_err(Xd.grad[:, torch.from_numpy(COLS).to(dev)].cpu().numpy(), dx64)
    # dP: Euler's identity for the degree-t homogeneous loss
    lhs = float(np.dot(vals.grad.double().cpu().numpy(), g.data.astype(np.float64)))
    R = Yh.astype(np.float64) - T
    rhs = t * float((R * Yh).sum())
    print(f"\n500k x 128, t=6: dX err {ex:.2e}; sum dP*P {lhs:.6e} vs t<Y-T,Y> {rhs:.6e}")
    assert ex <= 1e-3
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)


def test_precision_1m_f256_t4(gpu):
    n, k, F, t = 1_000_000, 15, 256, 4
    g = synth.uniform_knn(n, k, seed=4).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    X = synth.normal(n, F, seed=4)
    Xd = torch.from_numpy(X).cuda()
    cols = np.arange(0, 256, 16)
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X[:, cols], t)
    Y32 = torch_ops.diffuse(P, Xd, t)[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    assert np.array_equal(Y32, orc.diffuse_matrix_f32(g.indptr, g.indices, g.data,
                                                      X[:, cols], t))
    errs = {}
    for name, kw in (("fp16, 16-bit X", dict(fmt="fp16", round_input=True)),
                     ("fp16", dict(fmt="fp16")),
                     ("bf16", dict(fmt="bf16"))):
        Y = torch_ops.diffuse_lowp(P, Xd, t, **kw)[:, torch.from_numpy(cols).cuda()]
        errs[name] = orc.rel_err(Y.cpu().numpy(), y64)
    print("\n1M x 256, t=4 vs FP64: fp32 %.2e, %s" % (
        orc.rel_err(Y32, y64), ", ".join(f"{k} {v:.2e}" for k, v in errs.items())))
    assert orc.rel_err(Y32, y64) <= 1e-5
    assert errs["fp16, 16-bit X"] <= 1e-3 and errs["fp16"] <= 1e-3


def test_batching_256_graphs(gpu):
    F, t = 64, 5
    sizes, ks = synth.batch_sizes(256, seed=5)
    graphs = [synth.uniform_knn(int(n), int(k), seed=5000 + i).normalized()
              for i, (n, k) in enumerate(zip(sizes, ks))]
    mats = [CsrMatrix.from_arrays(gr.indptr, gr.indices, gr.data, (gr.n, gr.n))
            for gr in graphs]
    B = batch.block_diag(mats)
    off = batch.offsets(sizes)
    N = int(off[-1])
    X = synth.normal(N, F, seed=5)
    T = synth.normal(N, F, seed=105)
    plan = batch.DiffusionPlan(B, N, F, t, every=1)
    plan.X.copy_(torch.from_numpy(X).cuda())
    plan.T.copy_(torch.from_numpy(T).cuda())
    plan.launch()
    torch.cuda.synchronize()
    Y = plan.Y.cpu().numpy()
    dX = plan.dX.cpu().numpy()
    dP = plan.dP.cpu().numpy()
    plan.close()
    nz = np.concatenate([[0], np.cumsum([gr.nnz for gr in graphs])])
    rng = np.random.Generator(np.random.PCG64(5))
    sample = rng.choice(256, 6, replace=False)
    for i in sample:
        a, b = int(off[i]), int(off[i + 1])
        gr = graphs[i]
        alone = torch_ops.diffuse(mats[i], torch.from_numpy(X[a:b]).cuda(), t).cpu().numpy()
        assert np.array_equal(Y[a:b], alone), i
    # gradients of the smallest sampled graphs against the FP64 oracle
    for i in sorted(sample, key=lambda j: sizes[j])[:3]:
        a, b = int(off[i]), int(off[i + 1])
        gr = graphs[i]
        _, _, dx64, dp64 = orc.grad_f64(gr.indptr, gr.indices, gr.data, X[a:b], T[a:b], t)
        assert orc.rel_err(dX[a:b], dx64) <= 1e-3, i
        assert orc.rel_err(dP[int(nz[i]):int(nz[i + 1])], dp64) <= 1e-3, i
    print(f"\n256 graphs, {N} rows, {B.nnz} nonzeros: sampled graphs bit-identical "
          "alone; gradients within 1e-3 of FP64")
