"""Host-array wrappers of the drop-in diffusion entry points (scl_diffusion.h).

Each function calls the C ABI exactly as a ctypes binding of the reference's
``scl/binding/c_api/diffusion.h`` would (the reference's codegen would emit
such a module into ``python/scl/_bindings/diffusion.py``); arrays are
float32 / C-contiguous and are modified in place like the C buffers are.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import _lib
from .sparse import SparseMatrixBase


def _f32_inplace(a: np.ndarray) -> np.ndarray:
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32
            and a.flags.c_contiguous and a.flags.writeable):
        raise TypeError("expected a writeable C-contiguous float32 numpy array")
    return a


def compute_transition_matrix(adjacency: SparseMatrixBase, symmetric: bool = False
                              ) -> None:
    """scl_diffusion_compute_transition_matrix (values normalised in place)."""
    _lib.check(_lib.load().scl_diffusion_compute_transition_matrix(
        adjacency.handle, 1 if symmetric else 0))


def compute_transition_matrix_ex(adjacency: SparseMatrixBase,
                                 symmetric: bool = False, residual: bool = True
                                 ) -> Tuple[np.ndarray, Optional[float]]:
    """Normalises in place and returns (FP64 degrees, FP64 stationarity
    residual); the residual (square matrices only) is None when not asked
    for."""
    rows = adjacency.shape[0] if adjacency._is_csr else adjacency.shape[1]
    deg = np.empty(rows, np.float64)
    res = ctypes.c_double(0.0)
    _lib.check(_lib.load().scl_gpu_compute_transition_matrix_ex(
        adjacency.handle, 1 if symmetric else 0, deg.ctypes.data,
        ctypes.addressof(res) if residual else None))
    return deg, (res.value if residual else None)


def diffuse_matrix(transition: SparseMatrixBase, X: np.ndarray, n_steps: int,
                   n_nodes: Optional[int] = None,
                   n_features: Optional[int] = None) -> np.ndarray:
    """scl_diffusion_diffuse_matrix: X <- P^t X in place; returns X."""
    X = _f32_inplace(X)
    n = X.shape[0] if n_nodes is None else n_nodes
    F = (X.shape[1] if X.ndim == 2 else 1) if n_features is None else n_features
    _lib.check(_lib.load().scl_diffusion_diffuse_matrix(
        transition.handle, X.ctypes.data, n, F, n_steps))
    return X


def magic_impute(transition: SparseMatrixBase, X: np.ndarray, t: int = 0
                 ) -> np.ndarray:
    """scl_diffusion_magic_impute (t == 0 means 3)."""
    X = _f32_inplace(X)
    _lib.check(_lib.load().scl_diffusion_magic_impute(
        transition.handle, X.ctypes.data, X.shape[0],
        X.shape[1] if X.ndim == 2 else 1, t))
    return X


def diffuse_vector(transition: SparseMatrixBase, x: np.ndarray, n_steps: int
                   ) -> np.ndarray:
    """scl_diffusion_diffuse_vector: x <- P^t x in place; returns x."""
    x = _f32_inplace(x)
    _lib.check(_lib.load().scl_diffusion_diffuse_vector(
        transition.handle, x.ctypes.data, x.shape[0], n_steps))
    return x


# ------------------------------------------------ vector-side consumers --
# Host-array wrappers of the scl_diffusion_* walks (0 selects the reference's
# defaults, as at the C ABI).

def _n(transition: SparseMatrixBase) -> int:
    return transition.shape[0]


def _idx(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def compute_dpt(transition: SparseMatrixBase, root_cell: int, max_iter: int = 0,
                tol: float = 0.0) -> np.ndarray:
    """Diffusion pseudotime from one root (scl_diffusion_compute_dpt)."""
    out = np.empty(_n(transition), np.float32)
    _lib.check(_lib.load().scl_diffusion_compute_dpt(
        transition.handle, root_cell, out.ctypes.data, len(out), max_iter, tol))
    return out


def compute_dpt_multi_root(transition: SparseMatrixBase, root_cells, max_iter: int = 0
                           ) -> np.ndarray:
    """Mean pseudotime over several roots (scl_diffusion_compute_dpt_multi_root)."""
    roots = _idx(root_cells)
    out = np.empty(_n(transition), np.float32)
    _lib.check(_lib.load().scl_diffusion_compute_dpt_multi_root(
        transition.handle, roots.ctypes.data, len(roots), out.ctypes.data, len(out),
        max_iter))
    return out


def random_walk_with_restart(transition: SparseMatrixBase, seed_nodes, alpha: float = 0.0,
                             max_iter: int = 0, tol: float = 0.0) -> np.ndarray:
    """scl_diffusion_random_walk_with_restart."""
    seeds = _idx(seed_nodes)
    out = np.empty(_n(transition), np.float32)
    _lib.check(_lib.load().scl_diffusion_random_walk_with_restart(
        transition.handle, seeds.ctypes.data, len(seeds), out.ctypes.data, len(out), alpha,
        max_iter, tol))
    return out


def personalized_pagerank(transition: SparseMatrixBase, seed_node: int, alpha: float = 0.0,
                          max_iter: int = 0, tol: float = 0.0) -> np.ndarray:
    """scl_diffusion_personalized_pagerank."""
    out = np.empty(_n(transition), np.float32)
    _lib.check(_lib.load().scl_diffusion_personalized_pagerank(
        transition.handle, seed_node, out.ctypes.data, len(out), alpha, max_iter, tol))
    return out


def diffusion_map_embedding(transition: SparseMatrixBase, n_components: int,
                            n_iter: int = 0) -> np.ndarray:
    """n x n_components embedding (scl_diffusion_diffusion_map_embedding)."""
    n = _n(transition)
    out = np.empty((n, n_components), np.float32)
    _lib.check(_lib.load().scl_diffusion_diffusion_map_embedding(
        transition.handle, out.ctypes.data, n, n_components, n_iter))
    return out


def heat_kernel_signature(transition: SparseMatrixBase, t: float = 1.0, n_steps: int = 0
                          ) -> np.ndarray:
    """scl_diffusion_heat_kernel_signature."""
    out = np.empty(_n(transition), np.float32)
    _lib.check(_lib.load().scl_diffusion_heat_kernel_signature(
        transition.handle, out.ctypes.data, len(out), t, n_steps))
    return out


def diffusion_distance(transition: SparseMatrixBase, n_steps: int = 3) -> np.ndarray:
    """n x n diffusion distances (scl_diffusion_diffusion_distance)."""
    n = _n(transition)
    out = np.empty((n, n), np.float32)
    _lib.check(_lib.load().scl_diffusion_diffusion_distance(
        transition.handle, out.ctypes.data, n, n_steps))
    return out


def lazy_random_walk(transition: SparseMatrixBase, x: np.ndarray, n_steps: int,
                     laziness: float = 0.0) -> np.ndarray:
    """x <- laziness x + (1 - laziness) P x, n_steps times, in place."""
    x = _f32_inplace(x)
    _lib.check(_lib.load().scl_diffusion_lazy_random_walk(
        transition.handle, x.ctypes.data, x.shape[0], n_steps, laziness))
    return x
