"""Device-tensor entry points (scl_diffusion_gpu.h) for torch tensors.

Tensors must be float32, C-contiguous and on the current HIP device; work is
enqueued on torch's current stream, so it orders with surrounding torch ops.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .sparse import SparseMatrixBase


# rows per group of the shape flags' group code (bits 16..17)
_ROWS_PER_GROUP = (0, 4, 8, 64)
TILED = 128  # rows_per_group code of the LDS-tiled form

def _stream_ptr(t: torch.Tensor):
    return torch.cuda.current_stream(t.device).cuda_stream


def _check_dense(X: torch.Tensor) -> None:
    if not X.is_cuda:
        raise ValueError("expected a device tensor")
    if X.dtype != torch.float32 or not X.is_contiguous() or X.dim() != 2:
        raise ValueError("expected a contiguous (n, F) float32 tensor")


def diffuse(P: SparseMatrixBase, X: torch.Tensor, n_steps: int) -> torch.Tensor:
    """Returns P^t X (X unchanged)."""
    _check_dense(X)
    n, F = X.shape
    Y = torch.empty_like(X)
    work = torch.empty_like(X) if n_steps >= 2 else None
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().scl_gpu_diffuse_matrix_device(
            P.handle, X.data_ptr(), Y.data_ptr(),
            None if work is None else work.data_ptr(), n, F, n_steps,
            _stream_ptr(X)))
    return Y


def lowp_scale_exp(P: SparseMatrixBase, X: torch.Tensor, n_steps: int) -> int:
    """Power-of-two exponent that keeps fp16 iterates of P^t X in range."""
    _check_dense(X)
    n, F = X.shape
    k = ctypes.c_int(0)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().scl_gpu_lowp_scale_exp(
            P.handle, X.data_ptr(), n, F, n_steps, ctypes.byref(k), _stream_ptr(X)))
    return k.value


def diffuse_lowp(P: SparseMatrixBase, X: torch.Tensor, n_steps: int,
                 fmt: str = "bf16", bf16_values: bool = False, schedule: bool = True,
                 scale_exp=None, round_input: bool = False) -> torch.Tensor:
    """Returns P^t X (fp32) with reduced-precision intermediate iterates and
    fp32 accumulation: "bf16", "fp16" (scaled by 2^scale_exp, computed when
    not given) or "comp" (compensated bf16: bf16 plus an 8-bit residual of
    its rounding, 3 bytes per value).  round_input stores X in the same
    format first, so every step gathers reduced rows; bf16_values reads P's
    values as bf16 (compensated with "comp").  schedule=False forces the
    natural row order (otherwise the order follows the fp32 SpMM's tuned
    shape)."""
    _check_dense(X)
    n, F = X.shape
    fp16 = fmt == "fp16"
    comp = fmt == "comp"
    if fmt not in ("bf16", "fp16", "comp"):
        raise ValueError("fmt must be 'bf16', 'fp16' or 'comp'")
    if scale_exp is None:
        scale_exp = lowp_scale_exp(P, X, n_steps) if fp16 else 0
    Y = torch.empty_like(X)
    flags = ((1 if bf16_values else 0) | (0 if schedule else 2) | (4 if fp16 else 0)
             | (8 if round_input else 0) | (16 if comp else 0))
    work = None
    if n_steps >= 2 or round_input:
        nb = ctypes.c_size_t(0)
        _lib.check(_lib.load().scl_gpu_lowp_workspace(n, F, flags, ctypes.byref(nb)))
        work = torch.empty(nb.value, dtype=torch.uint8, device=X.device)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().scl_gpu_diffuse_matrix_lowp_device(
            P.handle, X.data_ptr(), Y.data_ptr(),
            None if work is None else work.data_ptr(), n, F, n_steps, flags,
            int(scale_exp), _stream_ptr(X)))
    return Y


def diffuse_bf16(P: SparseMatrixBase, X: torch.Tensor, n_steps: int,
                 bf16_values: bool = False, schedule: bool = True) -> torch.Tensor:
    """P^t X with bf16 intermediate iterates (see diffuse_lowp)."""
    return diffuse_lowp(P, X, n_steps, "bf16", bf16_values, schedule)


def edge_gradient(P: SparseMatrixBase, G: torch.Tensor, Xprev: torch.Tensor,
                  dP: torch.Tensor = None, accumulate: bool = False) -> torch.Tensor:
    """dP[k] (+)= <G[i,:], Xprev[j,:]> for every nonzero k = (i, j) of P."""
    _check_dense(G)
    _check_dense(Xprev)
    n, F = G.shape
    if dP is None:
        dP = torch.empty(max(P.nnz, 1), dtype=torch.float32, device=G.device)
    with torch.cuda.device(G.device):
        _lib.check(_lib.load().scl_gpu_edge_gradient_device(
            P.handle, G.data_ptr(), Xprev.data_ptr(), dP.data_ptr(), n, F,
            1 if accumulate else 0, _stream_ptr(G)))
    return dP


def _direction(transpose: bool, adjoint: bool) -> int:
    return 2 if adjoint else (1 if transpose else 0)


def autotune(P: SparseMatrixBase, X: torch.Tensor, transpose: bool = False,
             adjoint: bool = False) -> float:
    """Picks the fastest launch shape for P and X's width; returns ms/launch.

    transpose tunes the P^T SpMM; adjoint tunes the fused adjoint +
    edge-gradient step of the reverse pass."""
    _check_dense(X)
    n, F = X.shape
    Y = torch.empty_like(X)
    best = ctypes.c_float(0.0)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().scl_gpu_spmm_autotune(
            P.handle, X.data_ptr(), Y.data_ptr(), n, F, _direction(transpose, adjoint),
            _stream_ptr(X), ctypes.addressof(best)))
    return best.value


def get_shape(P: SparseMatrixBase, n_features: int, transpose: bool = False,
              adjoint: bool = False) -> dict:
    v, g, u, f = (ctypes.c_int() for _ in range(4))
    _lib.check(_lib.load().scl_gpu_spmm_get_shape(
        P.handle, n_features, _direction(transpose, adjoint), ctypes.addressof(v),
        ctypes.addressof(g), ctypes.addressof(u), ctypes.addressof(f)))
    return {"vec": v.value, "group": g.value, "unroll": u.value,
            "nontemporal": bool(f.value & 1), "schedule": not (f.value & 2),
            "xcd_rows": bool(f.value & 4), "scalar_rows": bool(f.value & 8),
            "wg_per_cu": (f.value >> 8) & 0xF, "merge": (f.value >> 12) & 0xF,
            "xcd_block": (f.value >> 20) & 0xF,
            "rows_per_group": (TILED if f.value & (1 << 18)
                               else _ROWS_PER_GROUP[(f.value >> 16) & 0x3])}


def set_shape(P: SparseMatrixBase, n_features: int, vec: int, group: int,
              unroll: int = 8, nontemporal: bool = False, schedule: bool = True,
              transpose: bool = False, xcd_rows: bool = False,
              scalar_rows: bool = False, adjoint: bool = False,
              wg_per_cu: int = 0, merge: int = 0, rows_per_group: int = 0,
              xcd_block: int = 0) -> None:
    """merge=m > 0 selects the merge-path SpMM with 256 << (m - 1) path items
    per wave (group 64 only); rows_per_group=4 or 8 the row-grouped form, 64
    the LDS-staged groups (vec 4, group 64), TILED (128) the LDS tiles
    (vec 4; group = lanes per row, 4 or 8; wg_per_cu = LDS budget, 1 or 2;
    unroll = slabs per workgroup)."""
    code = ((1 << 18) if rows_per_group == TILED
            else _ROWS_PER_GROUP.index(rows_per_group) << 16)
    flags = ((1 if nontemporal else 0) | (0 if schedule else 2) | (4 if xcd_rows else 0)
             | (8 if scalar_rows else 0) | ((wg_per_cu & 0xF) << 8) | ((merge & 0xF) << 12)
             | code | ((xcd_block & 0xF) << 20))
    _lib.check(_lib.load().scl_gpu_spmm_set_shape(
        P.handle, n_features, _direction(transpose, adjoint), vec, group, unroll, flags))


def spmm(P: SparseMatrixBase, X: torch.Tensor, transpose: bool = False
         ) -> torch.Tensor:
    """Returns P X, or P^T X when transpose is set."""
    _check_dense(X)
    n, F = X.shape
    Y = torch.empty_like(X)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().scl_gpu_spmm_device(
            P.handle, X.data_ptr(), Y.data_ptr(), n, F, 1 if transpose else 0,
            _stream_ptr(X)))
    return Y


def set_values(P: SparseMatrixBase, values: torch.Tensor) -> None:
    """Replaces P's values (nnz float32 on the device, stored order) in its
    device copies and host arrays (scl_gpu_sparse_set_values_device)."""
    v = values.detach().contiguous()
    if v.dtype != torch.float32 or v.numel() != P.nnz or not v.is_cuda:
        raise ValueError("expected nnz float32 values on the device")
    with torch.cuda.device(v.device):
        _lib.check(_lib.load().scl_gpu_sparse_set_values_device(
            P.handle, v.data_ptr(), _stream_ptr(v)))
