"""Many independent graphs diffused together.

The graphs are stacked into one block-diagonal operator
(``scl_sparse_block_diag``), so each diffusion step over the whole batch is a
single launch of the same row-split kernel (rows of different blocks never
mix, so every graph's result is bit-identical to diffusing it alone), and the
full forward + loss-gradient + reverse sequence is recorded once into a
hipGraph (``scl_gpu_plan_diffusion``) and replayed with one launch.
:class:`BatchedPlan` runs the same sequence for all graphs inside one
persistent kernel (``scl_gpu_plan_diffusion_batched``), every graph on one
XCD with its steps ordered by its own progress.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .sparse import CsrMatrix, SparseMatrixBase


def block_diag(mats: Sequence[SparseMatrixBase]) -> SparseMatrixBase:
    """Deep-copied block-diagonal stack of CSR (or CSC) handles."""
    if not mats:
        raise ValueError("empty batch")
    arr = (ctypes.c_void_p * len(mats))(*[m.handle.value for m in mats])
    out = ctypes.c_void_p()
    _lib.check(_lib.load().scl_sparse_block_diag(arr, len(mats), ctypes.byref(out)))
    return type(mats[0])(out)


def offsets(sizes: Sequence[int]) -> np.ndarray:
    """Row offsets of the blocks in the stacked operator."""
    return np.concatenate([[0], np.cumsum(np.asarray(sizes, np.int64))])


class DiffusionPlan:
    """fwd (+ dY = Y - T) (+ reverse) over fixed device buffers, as one hipGraph.

    Buffers are allocated here (torch) and stay bound to the graph: write the
    input into ``X`` (and ``T``), call :meth:`launch`, read ``Y``, ``dX``,
    ``dP``.  The operator must outlive the plan (call :meth:`close` first).
    """

    def __init__(self, P: SparseMatrixBase, n: int, F: int, n_steps: int,
                 every: int = 2, backward: bool = True, want_dp: bool = True,
                 device: Optional[torch.device] = None):
        dev = device or torch.device("cuda", torch.cuda.current_device())
        f32 = dict(dtype=torch.float32, device=dev)
        self.P, self.n, self.F, self.t, self.every = P, n, F, n_steps, every
        n_ckpt = (n_steps + every - 1) // every - 1
        self.X = torch.zeros((n, F), **f32)
        self.Y = torch.empty((n, F), **f32)
        self.ckpt = torch.empty((max(n_ckpt, 1), n, F), **f32)
        self.fwd_work = torch.empty((2, n, F), **f32)
        self.T = self.dY = self.dX = self.dP = self.bwd_work = None
        lib = _lib.load()
        bwd_bytes = 0
        if backward:
            self.T = torch.zeros((n, F), **f32)
            self.dY = torch.empty((n, F), **f32)
            self.dX = torch.empty((n, F), **f32)
            if want_dp:
                self.dP = torch.empty(max(P.nnz, 1), **f32)
            nb = ctypes.c_size_t(0)
            _lib.check(lib.scl_gpu_diffuse_backward_workspace(
                P.handle, n, F, every, 1 if want_dp else 0, ctypes.byref(nb)))
            bwd_bytes = nb.value
            self.bwd_work = torch.empty(bwd_bytes, dtype=torch.uint8, device=dev)
        ptr = lambda t: None if t is None else t.data_ptr()
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(dev):
            torch.cuda.synchronize(dev)
            _lib.check(lib.scl_gpu_plan_diffusion(
                P.handle, ptr(self.X), ptr(self.Y), ptr(self.ckpt), ptr(self.fwd_work),
                ptr(self.T), ptr(self.dY), ptr(self.dX), ptr(self.dP), ptr(self.bwd_work),
                bwd_bytes, n, F, n_steps, every, ctypes.byref(self._plan)))
        self.device = dev

    def launch(self) -> None:
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.load().scl_gpu_plan_launch(self._plan, stream))

    def close(self) -> None:
        if self._plan is not None and self._plan.value:
            _lib.check(_lib.load().scl_gpu_plan_destroy(ctypes.byref(self._plan)))
        self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedPlan:
    """fwd (+ dY = Y - T) (+ reverse) of every block in ONE persistent launch.

    ``sizes`` are the rows of the stacked blocks.  Buffers are allocated here
    (torch) and bound to the plan: write ``X`` (and ``T``), call
    :meth:`launch`, read ``Y``, ``dY``, ``dX``, ``dP``.  The plan owns its
    intermediates; :meth:`check` waits for the last launch and raises if a
    work item did not complete.  F must be a multiple of 32 (at most 512).
    """

    def __init__(self, P: SparseMatrixBase, sizes: Sequence[int], F: int, n_steps: int,
                 backward: bool = True, want_dp: bool = True, rows_per_item: int = 0,
                 graphs_per_chunk: int = 0, device: Optional[torch.device] = None):
        dev = device or torch.device("cuda", torch.cuda.current_device())
        f32 = dict(dtype=torch.float32, device=dev)
        n = int(sum(int(x) for x in sizes))
        self.P, self.n, self.F, self.t = P, n, F, n_steps
        self.X = torch.zeros((n, F), **f32)
        self.Y = torch.empty((n, F), **f32)
        self.T = self.dY = self.dX = self.dP = None
        if backward:
            self.T = torch.zeros((n, F), **f32)
            self.dY = torch.empty((n, F), **f32)
            self.dX = torch.empty((n, F), **f32)
            if want_dp:
                self.dP = torch.empty(max(P.nnz, 1), **f32)
        rows = np.ascontiguousarray(np.asarray(sizes, dtype=np.int64))
        ptr = lambda t: None if t is None else t.data_ptr()
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(dev):
            torch.cuda.synchronize(dev)
            _lib.check(_lib.load().scl_gpu_plan_diffusion_batched(
                P.handle, rows.ctypes.data, len(rows), ptr(self.X), ptr(self.Y), ptr(self.T),
                ptr(self.dY), ptr(self.dX), ptr(self.dP), F, n_steps, rows_per_item,
                graphs_per_chunk, ctypes.byref(self._plan)))
        self.device = dev

    def launch(self) -> None:
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.load().scl_gpu_plan_launch(self._plan, stream))

    def check(self) -> None:
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.load().scl_gpu_plan_check(self._plan, stream))

    def close(self) -> None:
        if self._plan is not None and self._plan.value:
            _lib.check(_lib.load().scl_gpu_plan_destroy(ctypes.byref(self._plan)))
        self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
