"""Differentiable diffusion: ``Y = P^t X`` as a ``torch.autograd.Function``.

Forward runs the checkpointed device forward (checkpoints every ``every``
steps); backward runs the library's reverse pass, which recomputes the
missing iterates and applies one fused adjoint + edge-gradient kernel per
step, giving

    dX = (P^T)^t dY,      dP_k = sum_{s=1..t} <G_s[i,:], X_{s-1}[j,:]>

for every stored nonzero k = (i, j) of P (the entries are treated as free
parameters: no renormalisation is differentiated through).

The reference (scl-core) has no gradient; the forward semantics are those of
scl/kernel/diffusion.hpp:638-664.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from .sparse import SparseMatrixBase


def _stream(t: torch.Tensor):
    return torch.cuda.current_stream(t.device).cuda_stream


def _generation(P) -> int:
    g = ctypes.c_uint64(0)
    _lib.check(_lib.load().scl_gpu_sparse_generation(P.handle, ctypes.byref(g)))
    return g.value


class _Diffuse(torch.autograd.Function):
    """With ``values`` the forward and the reverse pass read P's pattern with
    those values (saved for backward, so an in-place edit between the two
    trips autograd's version check); P's own values are never written.
    Without them both read P's values, and backward refuses to run when P
    changed in between."""

    @staticmethod
    def forward(ctx, X, values, P, n_steps, every):
        lib = _lib.load()
        Xc = X.detach().contiguous()
        v = values.detach().contiguous() if values is not None else None
        n, F = Xc.shape
        n_ckpt = (n_steps + every - 1) // every - 1
        Y = torch.empty_like(Xc)
        ckpt = torch.empty((max(n_ckpt, 1), n, F), dtype=Xc.dtype, device=Xc.device) \
            if n_ckpt > 0 else None
        work = torch.empty((2, n, F), dtype=Xc.dtype, device=Xc.device) \
            if n_steps >= 2 else None
        _lib.check(lib.scl_gpu_diffuse_forward_ckpt_values_device(
            P.handle, None if v is None else v.data_ptr(), Xc.data_ptr(), Y.data_ptr(),
            None if ckpt is None else ckpt.data_ptr(),
            None if work is None else work.data_ptr(), n, F, n_steps, every,
            _stream(Xc)))
        del work
        ctx.P, ctx.n_steps, ctx.every = P, n_steps, every
        ctx.has_values = v is not None
        ctx.generation = None if v is not None else _generation(P)
        ctx.save_for_backward(Xc, ckpt if ckpt is not None else Xc.new_empty(0),
                              v if v is not None else Xc.new_empty(0))
        ctx.n_ckpt = n_ckpt
        return Y

    @staticmethod
    def backward(ctx, dY):
        lib = _lib.load()
        Xc, ckpt, v = ctx.saved_tensors
        n, F = Xc.shape
        P = ctx.P
        if not ctx.has_values and _generation(P) != ctx.generation:
            raise RuntimeError("the operator changed between forward and backward; pass "
                               "values= to differentiate through a changing operator")
        want_dp = ctx.has_values and ctx.needs_input_grad[1]
        dYc = dY.contiguous()
        nbytes = ctypes.c_size_t(0)
        query = (lib.scl_gpu_diffuse_backward_values_workspace if ctx.has_values
                 else lib.scl_gpu_diffuse_backward_workspace)
        _lib.check(query(P.handle, n, F, ctx.every, 1 if want_dp else 0,
                         ctypes.byref(nbytes)))
        work = torch.empty(nbytes.value, dtype=torch.uint8, device=Xc.device)
        dX = torch.empty_like(Xc)
        dP = torch.empty(P.nnz, dtype=Xc.dtype, device=Xc.device) if want_dp else None
        _lib.check(lib.scl_gpu_diffuse_backward_values_device(
            P.handle, v.data_ptr() if ctx.has_values else None, Xc.data_ptr(),
            ckpt.data_ptr() if ctx.n_ckpt > 0 else None,
            dYc.data_ptr(), dX.data_ptr(), None if dP is None else dP.data_ptr(),
            work.data_ptr(), nbytes.value, n, F, ctx.n_steps, ctx.every, _stream(Xc)))
        return (dX if ctx.needs_input_grad[0] else None, dP, None, None, None)


def diffuse(P: SparseMatrixBase, X: torch.Tensor, n_steps: int,
            values: Optional[torch.Tensor] = None, checkpoint_every: int = 2
            ) -> torch.Tensor:
    """Differentiable ``P^t X``.

    ``values`` (nnz float32 on X's device, stored order of P) stands in for
    P's entries (P itself is not modified) and receives ``dL/dP`` when it
    requires grad; without it P is a constant and only ``dL/dX`` is produced.
    """
    if X.dtype != torch.float32 or X.dim() != 2 or not X.is_cuda:
        raise ValueError("expected a (n, F) float32 device tensor")
    if n_steps < 1 or checkpoint_every < 1:
        raise ValueError("n_steps and checkpoint_every must be >= 1")
    return _Diffuse.apply(X, values, P, int(n_steps), int(checkpoint_every))
