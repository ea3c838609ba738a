"""Batched diffusion of many uneven graphs: block-diagonal stacking, one
launch per step, and the hipGraph plan (forward + loss gradient + reverse).

Every graph's result in the batch must be bit-identical to diffusing it
alone (rows of different blocks never mix); the plan's replays must equal
the eager sequence on every replay, with eager work between replays; the
gradients are checked against the FP64 oracle for a sample of graphs."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scl_amd import CsrMatrix, autograd, batch, synth, torch_ops

pytestmark = pytest.mark.gpu


def make_batch(n_graphs, seed, lo=200, hi=3000):
    sizes, ks = synth.batch_sizes(n_graphs, seed, lo=lo, hi=hi)
    graphs = [synth.uniform_knn(int(n), int(k), seed=seed * 1000 + i).normalized()
              for i, (n, k) in enumerate(zip(sizes, ks))]
    mats = [CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n)) for g in graphs]
    return graphs, mats


def test_block_diag_forward_matches_each_graph(gpu):
    graphs, mats = make_batch(24, seed=5)
    B = batch.block_diag(mats)
    off = batch.offsets([g.n for g in graphs])
    N, F, t = int(off[-1]), 64, 5
    assert B.shape == (N, N) and B.nnz == sum(g.nnz for g in graphs)
    X = synth.normal(N, F, seed=9)
    Y = torch_ops.diffuse(B, torch.from_numpy(X).cuda(), t).cpu().numpy()
    for i, g in enumerate(graphs):
        Xi = X[off[i]:off[i + 1]]
        assert np.array_equal(Y[off[i]:off[i + 1]],
                              orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, Xi, t)), i


@pytest.mark.parametrize("every", [1, 2])
def test_plan_replays_match_eager(gpu, every):
    graphs, mats = make_batch(16, seed=7)
    B = batch.block_diag(mats)
    N, F, t = B.shape[0], 64, 5
    X = torch.from_numpy(synth.normal(N, F, seed=1)).cuda()
    T = torch.from_numpy(synth.normal(N, F, seed=2)).cuda()
    # eager reference through autograd
    Xr = X.clone().requires_grad_(True)
    vals = torch.from_numpy(B.export()[2]).cuda().requires_grad_(True)
    Y = autograd.diffuse(B, Xr, t, values=vals, checkpoint_every=every)
    (0.5 * ((Y - T) ** 2).sum()).backward()
    plan = batch.DiffusionPlan(B, N, F, t, every=every)
    try:
        for rep in range(3):
            plan.X.copy_(X)
            plan.T.copy_(T)
            plan.dX.zero_()
            plan.launch()
            torch.cuda.synchronize()
            assert torch.equal(plan.Y, Y.detach()), rep
            assert torch.equal(plan.dX, Xr.grad), rep
            assert torch.equal(plan.dP[:B.nnz], vals.grad), rep
            # unrelated eager work between replays
            _ = torch_ops.spmm(B, X).sum().item()
    finally:
        plan.close()


def test_batched_gradients_vs_fp64(gpu):
    graphs, mats = make_batch(12, seed=11, lo=300, hi=1500)
    B = batch.block_diag(mats)
    off = batch.offsets([g.n for g in graphs])
    N, F, t = B.shape[0], 16, 5
    X = synth.normal(N, F, seed=3)
    T = synth.normal(N, F, seed=4)
    plan = batch.DiffusionPlan(B, N, F, t, every=2)
    try:
        plan.X.copy_(torch.from_numpy(X))
        plan.T.copy_(torch.from_numpy(T))
        plan.launch()
        torch.cuda.synchronize()
        dX, dP = plan.dX.cpu().numpy(), plan.dP.cpu().numpy()
    finally:
        plan.close()
    nz = np.concatenate([[0], np.cumsum([g.nnz for g in graphs])])
    for i in (0, 5, 11):
        g = graphs[i]
        _, _, dX64, dP64 = orc.grad_f64(g.indptr, g.indices, g.data,
                                        X[off[i]:off[i + 1]], T[off[i]:off[i + 1]], t)
        assert orc.rel_err(dX[off[i]:off[i + 1]], dX64) <= 1e-3
        assert orc.rel_err(dP[nz[i]:nz[i + 1]], dP64) <= 1e-3
