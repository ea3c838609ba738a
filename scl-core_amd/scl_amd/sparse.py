"""CSR / CSC handles over the C ABI.

Follows the reference's ``python/scl/core/sparse.py`` (CsrMatrix 247-650,
CscMatrix 651-): ``copy`` makes a deep copy through ``scl_sparse_create``,
``wrap`` makes a zero-copy view through ``scl_sparse_wrap`` (the numpy arrays
must stay alive and are kept referenced here), indices are int64 and values
float32 (the f32_i64 variant).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import _lib


def _as_i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.int64)


def _as_f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.float32)


class SparseMatrixBase:
    _is_csr = True

    def __init__(self, handle: ctypes.c_void_p, keepalive=None):
        self._handle = handle
        self._keepalive = keepalive

    # lifecycle -----------------------------------------------------------
    def destroy(self) -> None:
        if self._handle is not None and self._handle.value:
            _lib.check(_lib.load().scl_sparse_destroy(ctypes.byref(self._handle)))
        self._handle = None
        self._keepalive = None

    close = destroy

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._handle is None or not self._handle.value:
            raise ValueError("sparse handle has been destroyed")
        return self._handle

    # queries ---------------------------------------------------------------
    def _q(self, name, ctype):
        out = ctype()
        _lib.check(getattr(_lib.load(), name)(self.handle, ctypes.byref(out)))
        return out.value

    @property
    def shape(self) -> Tuple[int, int]:
        return (self._q("scl_sparse_rows", _lib.c_index),
                self._q("scl_sparse_cols", _lib.c_index))

    @property
    def nnz(self) -> int:
        return self._q("scl_sparse_nnz", _lib.c_index)

    @property
    def is_contiguous(self) -> bool:
        return bool(self._q("scl_sparse_is_contiguous", _lib.c_bool))

    @property
    def format(self) -> str:
        return "csr" if self._is_csr else "csc"

    def export(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(indptr, indices, data) copies of the stored primary-slot arrays."""
        rows, cols = self.shape
        prim = rows if self._is_csr else cols
        nnz = self.nnz
        indptr = np.empty(prim + 1, np.int64)
        indices = np.empty(max(nnz, 1), np.int64)
        data = np.empty(max(nnz, 1), np.float32)
        _lib.check(_lib.load().scl_sparse_export(
            self.handle, indptr.ctypes.data, indices.ctypes.data, data.ctypes.data))
        return indptr, indices[:nnz], data[:nnz]

    def clone(self):
        out = ctypes.c_void_p()
        _lib.check(_lib.load().scl_sparse_clone(self.handle, ctypes.byref(out)))
        return type(self)(out)

    def locality_order(self) -> np.ndarray:
        """The row visiting order the device kernels use (host computation)."""
        rows, cols = self.shape
        order = np.empty(rows if self._is_csr else cols, np.int64)
        _lib.check(_lib.load().scl_sparse_locality_order(self.handle, order.ctypes.data))
        return order

    def prepare_device(self) -> None:
        """Uploads the operator to the current GPU now."""
        _lib.check(_lib.load().scl_gpu_sparse_prepare(self.handle))

    def invalidate_device(self) -> None:
        _lib.check(_lib.load().scl_gpu_sparse_invalidate(self.handle))

    # construction ------------------------------------------------------------
    @classmethod
    def _create(cls, indptr, indices, data, shape, wrap: bool):
        indptr, indices, data = _as_i64(indptr), _as_i64(indices), _as_f32(data)
        rows, cols = int(shape[0]), int(shape[1])
        nnz = int(indices.shape[0])
        if data.shape[0] != nnz:
            raise ValueError("indices and data lengths differ")
        out = ctypes.c_void_p()
        # the ABI rejects NULL arrays even when nnz == 0
        idx_buf = indices if nnz else np.zeros(1, np.int64)
        dat_buf = data if nnz else np.zeros(1, np.float32)
        fn = _lib.load().scl_sparse_wrap if wrap else _lib.load().scl_sparse_create
        _lib.check(fn(ctypes.byref(out), rows, cols, nnz, indptr.ctypes.data,
                      idx_buf.ctypes.data, dat_buf.ctypes.data,
                      1 if cls._is_csr else 0))
        keep = (indptr, idx_buf, dat_buf) if wrap else None
        return cls(out, keepalive=keep)

    @classmethod
    def from_arrays(cls, indptr, indices, data, shape):
        """Deep copy (scl_sparse_create)."""
        return cls._create(indptr, indices, data, shape, wrap=False)

    @classmethod
    def wrap(cls, indptr, indices, data, shape):
        """Zero-copy view (scl_sparse_wrap); arrays are converted to int64 /
        float32 first if needed, and the converted arrays are what is shared."""
        return cls._create(indptr, indices, data, shape, wrap=True)

    @classmethod
    def copy(cls, sp):
        """From a scipy.sparse matrix of the same orientation."""
        fmt = "csr" if cls._is_csr else "csc"
        sp = sp.asformat(fmt)
        return cls.from_arrays(sp.indptr, sp.indices, sp.data, sp.shape)

    def to_scipy(self):
        import scipy.sparse as sps
        indptr, indices, data = self.export()
        ctor = sps.csr_matrix if self._is_csr else sps.csc_matrix
        return ctor((data, indices, indptr), shape=self.shape)


class CsrMatrix(SparseMatrixBase):
    _is_csr = True

    def transpose(self) -> "CscMatrix":
        out = ctypes.c_void_p()
        _lib.check(_lib.load().scl_sparse_transpose(self.handle, ctypes.byref(out)))
        return CscMatrix(out)


class CscMatrix(SparseMatrixBase):
    _is_csr = False

    def transpose(self) -> CsrMatrix:
        out = ctypes.c_void_p()
        _lib.check(_lib.load().scl_sparse_transpose(self.handle, ctypes.byref(out)))
        return CsrMatrix(out)


SparseMatrix = CsrMatrix
