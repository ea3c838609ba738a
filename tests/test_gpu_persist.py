"""The persistent batched plan: every graph of a block-diagonal stack runs its
forward steps, dY = Y - T, its fused adjoint + edge-gradient steps and the dP
permutation inside one kernel launch, pinned to one XCD.

Its outputs must equal the replayed launch sequence (scl_gpu_plan_diffusion
with every iterate stored and the default adjoint shape) bit for bit, on every
replay, for any item size and chunking of the queues; each launch must report
every work item complete; and the gradients are anchored to the FP64 oracle
on sampled graphs."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scl_amd import CsrMatrix, batch, synth, torch_ops

pytestmark = pytest.mark.gpu


def make_batch(n_graphs, seed, lo=200, hi=3000, empty=False):
    sizes, ks = synth.batch_sizes(n_graphs, seed, lo=lo, hi=hi)
    graphs = [synth.uniform_knn(int(n), int(k), seed=seed * 1000 + i).normalized()
              for i, (n, k) in enumerate(zip(sizes, ks))]
    mats = [CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n)) for g in graphs]
    if empty:
        # a block with no stored entries (all its rows diffuse to zero)
        m = 77
        mats.insert(3, CsrMatrix.from_arrays(np.zeros(m + 1, np.int64), np.zeros(0, np.int64),
                                             np.zeros(0, np.float32), (m, m)))
        sizes = list(sizes[:3]) + [m] + list(sizes[3:])
    return [int(s) for s in sizes], mats


def replayed(B, N, F, t, X, T, backward=True):
    plan = batch.DiffusionPlan(B, N, F, t, every=1, backward=backward)
    try:
        plan.X.copy_(X)
        if backward:
            plan.T.copy_(T)
        plan.launch()
        torch.cuda.synchronize()
        out = {"Y": plan.Y.clone()}
        if backward:
            out.update(dY=plan.dY.clone(), dX=plan.dX.clone(), dP=plan.dP[:B.nnz].clone())
        return out
    finally:
        plan.close()


def assert_same(bp, ref, what, rep=None):
    for k, v in ref.items():
        got = getattr(bp, k)
        if k == "dP":
            got = got[:v.numel()]
        assert torch.equal(got, v), (what, k, rep)


@pytest.mark.parametrize("F,t", [(32, 3), (64, 5), (96, 2), (128, 1), (256, 3), (512, 2)])
def test_batched_plan_matches_replayed_plan(gpu, F, t):
    sizes, mats = make_batch(20, seed=21, empty=True)
    B = batch.block_diag(mats)
    N = B.shape[0]
    X = torch.from_numpy(synth.normal(N, F, seed=F + t)).cuda()
    T = torch.from_numpy(synth.normal(N, F, seed=F + t + 1)).cuda()
    ref = replayed(B, N, F, t, X, T)
    bp = batch.BatchedPlan(B, sizes, F, t)
    try:
        for rep in range(3):
            bp.X.copy_(X)
            bp.T.copy_(T)
            bp.dX.zero_()
            bp.dP.zero_()
            bp.launch()
            bp.check()
            assert_same(bp, ref, (F, t), rep)
            # unrelated eager work between replays
            _ = torch_ops.spmm(B, X).sum().item()
    finally:
        bp.close()


@pytest.mark.parametrize("rows_per_item,graphs_per_chunk", [(4, 0), (7, 1), (64, 3), (5000, 0)])
def test_batched_plan_items_and_chunks(gpu, rows_per_item, graphs_per_chunk):
    sizes, mats = make_batch(17, seed=23, lo=100, hi=4000)
    B = batch.block_diag(mats)
    N, F, t = B.shape[0], 64, 4
    X = torch.from_numpy(synth.normal(N, F, seed=3)).cuda()
    T = torch.from_numpy(synth.normal(N, F, seed=4)).cuda()
    ref = replayed(B, N, F, t, X, T)
    bp = batch.BatchedPlan(B, sizes, F, t, rows_per_item=rows_per_item,
                           graphs_per_chunk=graphs_per_chunk)
    try:
        for rep in range(2):
            bp.X.copy_(X)
            bp.T.copy_(T)
            bp.launch()
            bp.check()
            assert_same(bp, ref, (rows_per_item, graphs_per_chunk), rep)
    finally:
        bp.close()


def test_batched_plan_forward_only_and_without_dp(gpu):
    sizes, mats = make_batch(9, seed=25)
    B = batch.block_diag(mats)
    N, F, t = B.shape[0], 64, 3
    X = torch.from_numpy(synth.normal(N, F, seed=5)).cuda()
    T = torch.from_numpy(synth.normal(N, F, seed=6)).cuda()
    ref = replayed(B, N, F, t, X, T)
    fwd = batch.BatchedPlan(B, sizes, F, t, backward=False)
    nodp = batch.BatchedPlan(B, sizes, F, t, want_dp=False)
    try:
        fwd.X.copy_(X)
        fwd.launch()
        fwd.check()
        assert torch.equal(fwd.Y, ref["Y"])
        nodp.X.copy_(X)
        nodp.T.copy_(T)
        nodp.launch()
        nodp.check()
        assert nodp.dP is None
        for k in ("Y", "dY", "dX"):
            assert torch.equal(getattr(nodp, k), ref[k]), k
    finally:
        fwd.close()
        nodp.close()


def test_batched_plan_gradients_vs_fp64(gpu):
    sizes, mats = make_batch(12, seed=27, lo=300, hi=1500)
    B = batch.block_diag(mats)
    off = batch.offsets(sizes)
    N, F, t = B.shape[0], 32, 5
    X = synth.normal(N, F, seed=7)
    T = synth.normal(N, F, seed=8)
    bp = batch.BatchedPlan(B, sizes, F, t)
    try:
        bp.X.copy_(torch.from_numpy(X))
        bp.T.copy_(torch.from_numpy(T))
        bp.launch()
        bp.check()
        Y, dX, dP = bp.Y.cpu().numpy(), bp.dX.cpu().numpy(), bp.dP.cpu().numpy()
    finally:
        bp.close()
    ip, ix, pv = B.export()
    nz = ip[off]
    for i in (0, 6, 11):
        a, b = int(off[i]), int(off[i + 1])
        gi = ip[a:b + 1] - ip[a]
        gx = ix[ip[a]:ip[b]] - a
        gv = pv[ip[a]:ip[b]]
        _, y64, dX64, dP64 = orc.grad_f64(gi, gx, gv, X[a:b], T[a:b], t)
        assert orc.rel_err(Y[a:b], y64) <= 1e-5
        assert orc.rel_err(dX[a:b], dX64) <= 1e-5
        assert orc.rel_err(dP[nz[i]:nz[i + 1]], dP64) <= 1e-5


def test_batched_plan_refuses_bad_batches(gpu):
    sizes, mats = make_batch(5, seed=29)
    B = batch.block_diag(mats)
    with pytest.raises(Exception):
        batch.BatchedPlan(B, sizes, 50, 2)          # F not a multiple of 32
    with pytest.raises(Exception):
        batch.BatchedPlan(B, sizes[:-1], 64, 2)     # blocks do not cover the rows
    with pytest.raises(Exception):
        merged = [sizes[0] + sizes[1]] + sizes[2:]
        batch.BatchedPlan(B, [merged[0] - 1, 1] + merged[1:], 64, 2)  # entries cross blocks
