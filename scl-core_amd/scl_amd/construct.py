"""Host-array wrappers of the operator-construction entry points:
scl_neighbors.h (exact k-nearest neighbours), scl_kernel.h (kernel weights
of kNN distances) and scl_transition.h (row normalisation, symmetrisation,
velocity transition, stationary distribution), called exactly as a ctypes
binding of the reference's ``scl/binding/c_api/neighbors.h``, ``kernel.h``
and ``transition.h`` would."""
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
    prim = matrix.shape[0] if matrix._is_csr else matrix.shape[1]
    if n == prim:
        size = matrix.nnz     # every stored entry: no need to export the arrays
    else:
        ip, _, _ = matrix.export()
        # rows past the handle's are the library's error to report
        size = int(ip[min(n, len(ip) - 1)] - ip[0])
    out = np.empty(size, np.float32)
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


def neighbor_norms(matrix: SparseMatrixBase) -> np.ndarray:
    """scl_neighbors_compute_norms: squared norm of every sample."""
    out = np.zeros(matrix.shape[0], np.float32)
    _lib.check(_lib.load().scl_neighbors_compute_norms(matrix.handle, out.ctypes.data))
    return out


def knn(matrix: SparseMatrixBase, k: int, norms_sq: np.ndarray = None):
    """scl_knn (with norms) or scl_knn_simple: (indices n x k int64,
    distances n x k float32) of the samples (primary slots)."""
    n = matrix.shape[0] if matrix._is_csr else matrix.shape[1]
    idx = np.empty((n, k), np.int64)
    dist = np.empty((n, k), np.float32)
    lib = _lib.load()
    if norms_sq is None:
        _lib.check(lib.scl_knn_simple(matrix.handle, n, k, idx.ctypes.data, dist.ctypes.data))
    else:
        nr = np.ascontiguousarray(norms_sq, dtype=np.float32)
        _lib.check(lib.scl_knn(matrix.handle, nr.ctypes.data, n, k, idx.ctypes.data,
                               dist.ctypes.data))
    return idx, dist
