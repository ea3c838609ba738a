"""GPU parity of the reverse-mode diffusion (dX, per-nonzero dP) against the
FP64 oracle gradient (oracle/oracle.c orc_grad_f64, itself pinned by finite
differences in test_oracle.py) and against central finite differences.

Tolerance: relative error <= 1e-3 in the reference's max-norm convention
(fp32 gradients of an fp32 forward vs the FP64 gradient)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scl_amd import CsrMatrix, autograd, synth, torch_ops

pytestmark = pytest.mark.gpu


def _problem(n, k, F, seed, knn=True):
    g = (synth.knn_graph(n, k, seed=seed) if knn else synth.uniform_knn(n, k, seed=seed)).normalized()
    X = synth.normal(n, F, seed=seed)
    T = synth.normal(n, F, seed=seed + 100)
    return g, X, T


@pytest.mark.parametrize("n,k,F,t,every", [
    (3000, 10, 16, 3, 2), (4000, 8, 128, 6, 2), (2500, 6, 50, 4, 3),
    (2000, 5, 3, 2, 1), (5000, 7, 300, 5, 2), (1500, 12, 1100, 2, 2)])
def test_gradients_match_fp64(gpu, n, k, F, t, every):
    g, X, T = _problem(n, k, F, seed=n + F)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev).requires_grad_(True)
    Td = torch.from_numpy(T).to(dev)
    vals = torch.from_numpy(g.data.copy()).to(dev).requires_grad_(True)
    Y = autograd.diffuse(P, Xd, t, values=vals, checkpoint_every=every)
    L = 0.5 * ((Y - Td) ** 2).sum()
    L.backward()
    L64, Y64, dX64, dP64 = orc.grad_f64(g.indptr, g.indices, g.data, X, T, t)
    # forward identical to the plain kernel
    assert torch.equal(Y.detach(), torch_ops.diffuse(P, Xd.detach(), t))
    ex = orc.rel_err(Xd.grad.cpu().numpy(), dX64)
    ep = orc.rel_err(vals.grad.cpu().numpy(), dP64)
    print(f"\nn={n} F={F} t={t} every={every}: dX err {ex:.2e}, dP err {ep:.2e}")
    assert ex <= 1e-3 and ep <= 1e-3


def test_dx_only_and_constant_operator(gpu):
    g, X, T = _problem(3000, 8, 32, seed=5)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (3000, 3000))
    Xd = torch.from_numpy(X).cuda().requires_grad_(True)
    Y = autograd.diffuse(P, Xd, 4)
    (0.5 * ((Y - torch.from_numpy(T).cuda()) ** 2).sum()).backward()
    _, _, dX64, _ = orc.grad_f64(g.indptr, g.indices, g.data, X, T, 4)
    assert orc.rel_err(Xd.grad.cpu().numpy(), dX64) <= 1e-3
    # the adjoint of the adjoint: dX of <P^t X, G> is (P^T)^t G exactly as the
    # transposed SpMM computes it step by step
    G = torch.from_numpy(T).cuda()
    Xd.grad = None
    Y = autograd.diffuse(P, Xd, 2, checkpoint_every=1)
    (Y * G).sum().backward()
    ref = torch_ops.spmm(P, torch_ops.spmm(P, G, transpose=True), transpose=True)
    assert torch.equal(Xd.grad, ref)


def test_finite_differences(gpu):
    """Central differences of the FP64 loss (eps 1e-3) at 32 entries of X and
    32 nonzeros of P, against the GPU gradient (FP32)."""
    n, F, t = 2000, 8, 6
    g, X, T = _problem(n, 8, F, seed=33)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd = torch.from_numpy(X).cuda().requires_grad_(True)
    vals = torch.from_numpy(g.data.copy()).cuda().requires_grad_(True)
    Y = autograd.diffuse(P, Xd, t, values=vals, checkpoint_every=2)
    (0.5 * ((Y - torch.from_numpy(T).cuda()) ** 2).sum()).backward()
    gX, gP = Xd.grad.cpu().numpy(), vals.grad.cpu().numpy()
    rng = np.random.Generator(np.random.PCG64(3))
    eps = 1e-3
    v64 = g.data.astype(np.float64)
    X64 = X.astype(np.float64)
    errs = []
    for _ in range(32):
        i, c = rng.integers(n), rng.integers(F)
        Xp, Xm = X64.copy(), X64.copy()
        Xp[i, c] += eps
        Xm[i, c] -= eps
        fd = (orc.loss_f64(g.indptr, g.indices, v64, Xp, T, t) -
              orc.loss_f64(g.indptr, g.indices, v64, Xm, T, t)) / (2 * eps)
        errs.append(abs(fd - gX[i, c]) / max(1e-2, abs(fd)))
    for _ in range(32):
        kk = rng.integers(len(v64))
        vp, vm = v64.copy(), v64.copy()
        vp[kk] += eps
        vm[kk] -= eps
        fd = (orc.loss_f64(g.indptr, g.indices, vp, X64, T, t) -
              orc.loss_f64(g.indptr, g.indices, vm, X64, T, t)) / (2 * eps)
        errs.append(abs(fd - gP[kk]) / max(1e-2, abs(fd)))
    print(f"\nfinite differences: max rel deviation {max(errs):.2e}")
    assert max(errs) < 1e-3


def test_unfused_building_blocks_agree_with_fused(gpu):
    """Forward iterates kept, adjoint by the transposed SpMM, edge gradient by
    its own kernel: the same dX bit for bit, and dP to fp32 rounding."""
    n, F, t = 6000, 96, 4
    g, X, T = _problem(n, 10, F, seed=77)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    its = [Xd]
    for _ in range(t):
        its.append(torch_ops.spmm(P, its[-1]))
    G = its[-1] - Td
    dP = None
    for s in range(t, 0, -1):
        dP = torch_ops.edge_gradient(P, G, its[s - 1], dP, accumulate=s != t)
        G = torch_ops.spmm(P, G, transpose=True)
    Xr = Xd.clone().requires_grad_(True)
    vals = torch.from_numpy(g.data.copy()).cuda().requires_grad_(True)
    Y = autograd.diffuse(P, Xr, t, values=vals, checkpoint_every=2)
    (0.5 * ((Y - Td) ** 2).sum()).backward()
    assert torch.equal(G, Xr.grad)
    assert torch.allclose(dP, vals.grad, rtol=1e-5, atol=1e-6 * vals.grad.abs().max().item())


@pytest.mark.parametrize("F", [64, 128, 300])
def test_every_adjoint_shape_agrees(gpu, F):
    """Every (vec, group) of the fused adjoint kernel, with and without an
    occupancy cap: dX bit-identical, dP equal up to the order of its
    per-feature sums, and both within 1e-3 of the FP64 gradient; autotuning
    picks one of them."""
    n, t = 17_000, 3   # above the size that gets a locality schedule
    g, X, T = _problem(n, 9, F, seed=F)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    _, _, dX64, dP64 = orc.grad_f64(g.indptr, g.indices, g.data, X, T, t)
    ref = None
    shapes = [(v, gr, sc) for v in (4, 2, 1) if F % v == 0 for gr in (4, 8, 16, 32, 64)
              for sc in (True, False)]
    for i, (v, gr, sc) in enumerate(shapes):
        occ = 4 if i % 2 else 0
        torch_ops.set_shape(P, F, v, gr, schedule=sc, adjoint=True, wg_per_cu=occ)
        got = torch_ops.get_shape(P, F, adjoint=True)
        assert (got["vec"], got["group"], got["schedule"], got["wg_per_cu"]) == (v, gr, sc, occ)
        Xr = Xd.clone().requires_grad_(True)
        vals = torch.from_numpy(g.data.copy()).cuda().requires_grad_(True)
        Y = autograd.diffuse(P, Xr, t, values=vals, checkpoint_every=2)
        (0.5 * ((Y - Td) ** 2).sum()).backward()
        if ref is None:
            ref = (Xr.grad.clone(), vals.grad.clone())
        assert torch.equal(Xr.grad, ref[0]), (v, gr)
        assert torch.allclose(vals.grad, ref[1], rtol=1e-5,
                              atol=1e-6 * ref[1].abs().max().item()), (v, gr)
        assert orc.rel_err(vals.grad.cpu().numpy(), dP64) <= 1e-3
        assert orc.rel_err(Xr.grad.cpu().numpy(), dX64) <= 1e-3
    torch_ops.autotune(P, Xd, adjoint=True)
    best = torch_ops.get_shape(P, F, adjoint=True)
    assert (best["vec"], best["group"], best["schedule"]) in shapes


def test_two_forwards_with_different_values_then_one_backward(gpu):
    """Each forward reads its own values tensor (saved for backward); P's
    host and device values are never written, so two forwards on one P with
    different values, then one backward through both, give each forward's
    own FP64 gradient."""
    n, F, t = 3000, 24, 4
    g, X, T = _problem(n, 9, F, seed=41)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    before = P.export()[2].copy()
    rng = np.random.default_rng(42)
    other = (g.data * rng.uniform(0.5, 1.5, g.nnz)).astype(np.float32)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Td = torch.from_numpy(T).to(dev)
    va = torch.from_numpy(g.data.copy()).to(dev).requires_grad_(True)
    vb = torch.from_numpy(other).to(dev).requires_grad_(True)
    Xa = Xd.clone().requires_grad_(True)
    Xb = Xd.clone().requires_grad_(True)
    Ya = autograd.diffuse(P, Xa, t, values=va)
    Yb = autograd.diffuse(P, Xb, t, values=vb)
    L = 0.5 * ((Ya - Td) ** 2).sum() + 0.5 * ((Yb - Td) ** 2).sum()
    L.backward()
    for vals, Xg, vg in ((g.data, Xa.grad, va.grad), (other, Xb.grad, vb.grad)):
        _, _, dX64, dP64 = orc.grad_f64(g.indptr, g.indices, vals, X, T, t)
        ex = orc.rel_err(Xg.cpu().numpy(), dX64)
        ep = orc.rel_err(vg.cpu().numpy(), dP64)
        assert ex <= 1e-3 and ep <= 1e-3, (ex, ep)
    assert np.array_equal(P.export()[2], before)
    # and the forward with P's own values still matches the plain kernel
    assert torch.equal(torch_ops.diffuse(P, Xd, t), Ya.detach())


def test_constant_operator_changed_before_backward_is_refused(gpu):
    n, F = 2000, 8
    g, X, T = _problem(n, 6, F, seed=43)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    Xd = torch.from_numpy(X).cuda().requires_grad_(True)
    Y = autograd.diffuse(P, Xd, 3)
    torch_ops.set_values(P, torch.from_numpy(g.data * 0.5).cuda())
    with pytest.raises(RuntimeError, match="changed between forward and backward"):
        Y.sum().backward()
