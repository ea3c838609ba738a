"""Host-array wrappers of the operator-construction entry points:
scl_kernel.h (kernel weights of kNN distances) and scl_transition.h
(row normalisation, symmetrisation, velocity transition, stationary
distribution), called exactly as a ctypes binding of the reference's
``scl/binding/c_api/kernel.h`` and ``transition.h`` would."""
from __future__ import annotations

import numpy as np

from . import _lib
from .sparse import SparseMatrixBase

KERNELS = {"gaussian": 0, "epanechnikov": 1, "cosine": 2, "linear": 3, "polynomial": 4,
           "laplacian": 5, "cauchy": 6, "sigmoid": 7, "uniform": 8, "triangular": 9}


def _kind(kernel) -> int:
    return KERNELS[kernel] if isinstance(kernel, str) else int(kernel)


def kernel_matrix(distances: SparseMatrixBase, bandwidth: float = 1.0,
                  kernel="gaussian") -> np.ndarray:
    """scl_kernel_compute_matrix: the kernel value of every stored distance."""
    out = np.empty(distances.nnz, np.float32)
    _lib.check(_lib.load().scl_kernel_compute_matrix(distances.handle, out.ctypes.data,
                                                     bandwidth, _kind(kernel)))
    return out


def kernel_row_sums(distances: SparseMatrixBase, bandwidth: float = 1.0,
                    kernel="gaussian") -> np.ndarray:
    """scl_kernel_row_sums (n_points = the handle's rows)."""
    n = distances.shape[0]
    out = np.zeros(n, np.float32)
    _lib.check(_lib.load().scl_kernel_row_sums(distances.handle, out.ctypes.data, n,
                                               bandwidth, _kind(kernel)))
    return out


def matrix_from_velocity(graph: SparseMatrixBase) -> np.ndarray:
    """scl_transition_matrix_from_velocity: dense n x n row-stochastic matrix."""
    n = graph.shape[0]
    out = np.empty((n, n), np.float32)
    _lib.check(_lib.load().scl_transition_matrix_from_velocity(graph.handle, n,
                                                               out.ctypes.data))
    return out


def row_normalize(matrix: SparseMatrixBase, n: int = -1) -> np.ndarray:
    """scl_transition_row_normalize_to_stochastic over the first n rows."""
    n = matrix.shape[0] if n < 0 else n
    ip, _, _ = matrix.export()
    # rows past the handle's are the library's error to report
    out = np.empty(int(ip[min(n, len(ip) - 1)] - ip[0]), np.float32)
    _lib.check(_lib.load().scl_transition_row_normalize_to_stochastic(
        matrix.handle, n, out.ctypes.data, len(out)))
    return out


def symmetrize(matrix: SparseMatrixBase) -> np.ndarray:
    """scl_transition_symmetrize: dense (T + T^T)/2, rows normalised."""
    n = matrix.shape[0]
    out = np.empty((n, n), np.float32)
    _lib.check(_lib.load().scl_transition_symmetrize(matrix.handle, n, out.ctypes.data))
    return out


def stationary_distribution(matrix: SparseMatrixBase, max_iter: int = 1000,
                            tol: float = 1e-6) -> np.ndarray:
    """scl_transition_stationary_distribution (the reference's kernel defaults)."""
    n = matrix.shape[0]
    out = np.empty(n, np.float32)
    _lib.check(_lib.load().scl_transition_stationary_distribution(
        matrix.handle, n, out.ctypes.data, max_iter, tol))
    return out
