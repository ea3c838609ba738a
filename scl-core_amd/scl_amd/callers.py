"""Host-array wrappers of the other SpMM callers: scl_impute.h,
scl_pseudotime.h and scl_algebra.h.

Each function calls the C ABI exactly as a ctypes binding of the reference's
``scl/binding/c_api/impute.h``, ``pseudotime.h`` and ``algebra.h`` would;
outputs are fresh float32 arrays, except that ``algebra_spmm`` updates ``Y``
in place like the C buffer.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .sparse import SparseMatrixBase


def impute_diffusion(matrix: SparseMatrixBase, transition: SparseMatrixBase, n_cells: int,
                     n_genes: int, n_steps: int) -> np.ndarray:
    """scl_impute_diffusion: P^n_steps applied to the dense expression matrix."""
    out = np.empty((n_cells, n_genes), np.float32)
    _lib.check(_lib.load().scl_impute_diffusion(matrix.handle, transition.handle, n_cells,
                                                n_genes, n_steps, out.ctypes.data))
    return out


def impute_magic(matrix: SparseMatrixBase, affinity: SparseMatrixBase, n_cells: int,
                 n_genes: int, t: int) -> np.ndarray:
    """scl_impute_magic: t steps of the row-normalised affinity."""
    out = np.empty((n_cells, n_genes), np.float32)
    _lib.check(_lib.load().scl_impute_magic(matrix.handle, affinity.handle, n_cells, n_genes,
                                            t, out.ctypes.data))
    return out


def pseudotime_diffusion(adjacency: SparseMatrixBase, root_cell: int, n_dcs: int = 10,
                         n: int = -1) -> np.ndarray:
    """scl_pseudotime_diffusion (n defaults to the operator's rows)."""
    n = adjacency.shape[0] if n < 0 else n
    out = np.zeros(n, np.float32)
    _lib.check(_lib.load().scl_pseudotime_diffusion(adjacency.handle, root_cell,
                                                    out.ctypes.data, n, n_dcs))
    return out


def algebra_spmm(A: SparseMatrixBase, X: np.ndarray, Y: np.ndarray, alpha: float = 1.0,
                 beta: float = 0.0) -> np.ndarray:
    """scl_algebra_spmm: Y = alpha A X + beta Y, Y updated in place."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    if not (isinstance(Y, np.ndarray) and Y.dtype == np.float32 and Y.flags.c_contiguous
            and Y.flags.writeable):
        raise TypeError("Y must be a writeable C-contiguous float32 numpy array")
    n_cols = Y.shape[1] if Y.ndim == 2 else 1
    _lib.check(_lib.load().scl_algebra_spmm(A.handle, X.ctypes.data, n_cols, Y.ctypes.data,
                                            alpha, beta))
    return Y
