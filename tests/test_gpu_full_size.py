"""Checks at the full sizes of the differentiation, 16-bit and batching
workloads, through properties that need no CPU run at that size: columns
never mix, so sampled columns are compared with the restatement and the
FP64 oracle run on those columns alone; the loss is homogeneous of degree t
in P's values, so sum_k dP_k P_k = t <Y - T, Y> (Euler) checks the whole
per-nonzero gradient at once; graphs of a batch never mix, so each sampled
graph is diffused alone and compared bit for bit.

The differentiation workload runs on its own geometric kNN graph and is
also checked against the FP64 oracle gradient over the whole problem and
against finite differences; the precision and batching workloads use
uniform random kNN graphs (k distinct neighbours per row), as their configs
say."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scl_amd import CsrMatrix, autograd, batch, synth, torch_ops

pytestmark = pytest.mark.gpu

COLS = np.array([0, 1, 17, 63, 64, 100, 126, 127])


def test_differentiation_knn30_500k_f128_t6(gpu):
    """The differentiation workload itself: n = 500,000 geometric kNN (k = 30,
    union-max symmetrised, nnz 22.1 M), X and T N(0,1) 500,000 x 128, t = 6,
    checkpoints every 2.  dX and every per-nonzero dP are checked against
    the FP64 oracle gradient over the whole problem; 32 X entries (a
    one-column FP64 loss each) and 32 P entries (the loss of one column,
    backpropagated on its own) against FP64 central differences with
    eps = 1e-3."""
    n, k, F, t, every, eps = 500_000, 30, 128, 6, 2, 1e-3
    g = synth.knn_graph(n, k, seed=3).normalized()
    ip, ix, p = g.indptr, g.indices, g.data
    P = CsrMatrix.from_arrays(ip, ix, p, (n, n))
    X = synth.normal(n, F, seed=3)
    T = synth.normal(n, F, seed=103)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev).requires_grad_(True)
    Td = torch.from_numpy(T).to(dev)
    vals = torch.from_numpy(p.copy()).to(dev).requires_grad_(True)
    Y = autograd.diffuse(P, Xd, t, values=vals, checkpoint_every=every)
    (0.5 * ((Y - Td) ** 2).sum()).backward()
    dX, dP = Xd.grad.cpu().numpy(), vals.grad.cpu().numpy()
    Yh = Y.detach().cpu().numpy()
    L64, Y64, dX64, dP64 = orc.grad_f64(ip, ix, p, X, T, t)
    ey, ex, ep = orc.rel_err(Yh, Y64), orc.rel_err(dX, dX64), orc.rel_err(dP, dP64)
    mx, mp = orc.median_rel_err(dX, dX64), orc.median_rel_err(dP, dP64)
    # forward columns bit-identical to the restatement
    assert np.array_equal(Yh[:, COLS], orc.diffuse_matrix_f32(ip, ix, p, X[:, COLS], t))
    # Euler: the loss is homogeneous of degree t in P's values
    R = Yh.astype(np.float64) - T
    lhs = float(np.dot(dP.astype(np.float64), p.astype(np.float64)))
    rhs = t * float((R * Yh).sum())
    # finite differences of one column's loss
    rng = np.random.Generator(np.random.PCG64(33))
    p64 = p.astype(np.float64)

    def col_loss(c, Xc, v):
        return orc.loss_f64(ip, ix, v, Xc, T[:, [c]], t)

    fx, gx = [], []
    for c, i in zip(rng.choice(F, 32), rng.integers(0, n, 32)):
        Xc = X[:, [c]].astype(np.float64)
        Xc[i, 0] += eps
        up = col_loss(c, Xc, p64)
        Xc[i, 0] -= 2 * eps
        fx.append((up - col_loss(c, Xc, p64)) / (2 * eps))
        gx.append(dX[i, c])
    c0 = int(rng.integers(0, F))
    Xc0 = X[:, [c0]].astype(np.float64)
    Xd.grad = None
    vals.grad = None
    Y = autograd.diffuse(P, Xd, t, values=vals, checkpoint_every=every)
    (0.5 * ((Y[:, c0] - Td[:, c0]) ** 2).sum()).backward()
    dPc = vals.grad.cpu().numpy()
    fp, gp = [], []
    for e in rng.choice(len(ix), 32, replace=False):
        v = p64.copy()
        v[e] += eps
        up = col_loss(c0, Xc0, v)
        v[e] -= 2 * eps
        fp.append((up - col_loss(c0, Xc0, v)) / (2 * eps))
        gp.append(dPc[e])
    efx = float(np.max(np.abs(np.subtract(gx, fx))) / np.max(np.abs(fx)))
    efp = float(np.max(np.abs(np.subtract(gp, fp))) / np.max(np.abs(fp)))
    print(f"\nknn30_500k (nnz {g.nnz}) x 128, t=6: Y err {ey:.2e}; dX err {ex:.2e} "
          f"(median {mx:.2e}); dP err {ep:.2e} (median {mp:.2e}) over all nonzeros; "
          f"FD: 32 X entries {efx:.2e}, 32 P entries {efp:.2e}; Euler {lhs:.6e} vs {rhs:.6e}")
    assert g.nnz == 22_122_568
    assert ey <= 1e-5 and ex <= 1e-3 and ep <= 1e-3
    assert efx <= 1e-3 and efp <= 1e-3
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
    for name, kw in (("compensated bf16 P and X", dict(fmt="comp", bf16_values=True,
                                                       round_input=True)),
                     ("bf16 P and X", dict(fmt="bf16", bf16_values=True, round_input=True)),
                     ("fp16, 16-bit X", dict(fmt="fp16", round_input=True)),
                     ("fp16", dict(fmt="fp16")),
                     ("bf16", dict(fmt="bf16"))):
        Y = torch_ops.diffuse_lowp(P, Xd, t, **kw)[:, torch.from_numpy(cols).cuda()]
        Yh = Y.cpu().numpy()
        errs[name] = (orc.rel_err(Yh, y64), orc.median_rel_err(Yh, y64))
    print("\n1M x 256, t=4 vs FP64 (max / median): fp32 %.2e / %.2e, %s" % (
        orc.rel_err(Y32, y64), orc.median_rel_err(Y32, y64),
        ", ".join(f"{k} {a:.2e} / {b:.2e}" for k, (a, b) in errs.items())))
    assert orc.rel_err(Y32, y64) <= 1e-5
    # the precision workload as specified (P values and X stored in bf16,
    # with the compensating residual) meets 1e-3; so does fp16 + 2^k
    assert errs["compensated bf16 P and X"][0] <= 1e-3
    assert errs["fp16, 16-bit X"][0] <= 1e-3 and errs["fp16"][0] <= 1e-3


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
