"""Seeded synthetic workloads (ctypes over lib/libscl_synth.so).

The five workloads the diffusion path is measured on, by their shapes:

  uniform15_10k  n=10,000, k=15 uniform neighbours, F=50, t=3
  knn30_200k     n=200,000, k=30 geometric kNN + union-max symmetrisation
                 (nnz about 8.9e6, rows 30..~100), F=2000 log1p(Poisson), t=8
  knn30_500k     n=500,000, same graph law, F=128 N(0,1), t=6 (gradients)
  uniform15_1m   n=1,000,000, k=15 uniform, F=256 N(0,1), t=4 (bf16)
  batch256       256 graphs, n_i log-uniform in [2000, 50000], k_i in 10..30,
                 F=64 N(0,1), t=5

Edge weights: w = exp(-d^2), d = |N(0,1)| (sigma = 1).  Rows of W are
normalised in FP64 to give P (stored fp32).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

_LIB = Path(__file__).resolve().parent.parent / "lib" / "libscl_synth.so"
_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
U64 = ctypes.c_uint64


def load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            raise ImportError(f"{_LIB} is missing: run `make -C scl-core_amd`")
        lib = ctypes.CDLL(os.fspath(_LIB))
        lib.synth_uniform_knn.restype = ctypes.c_int
        lib.synth_uniform_knn.argtypes = [I64, I64, U64, P, P, P]
        lib.synth_knn_graph.restype = I64
        lib.synth_knn_graph.argtypes = [I64, I64, I64, I64, ctypes.c_double, U64, P, P, P]
        lib.synth_row_normalize.restype = None
        lib.synth_row_normalize.argtypes = [I64, P, P]
        lib.synth_fill_normal.restype = None
        lib.synth_fill_normal.argtypes = [I64, I64, U64, P, I64, P]
        lib.synth_fill_log1p_poisson.restype = None
        lib.synth_fill_log1p_poisson.argtypes = [I64, I64, U64, P, I64, P]
        _lib = lib
    return _lib


@dataclass
class Graph:
    n: int
    indptr: np.ndarray   # int64, n + 1
    indices: np.ndarray  # int64, nnz
    data: np.ndarray     # float32, nnz

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def normalized(self) -> "Graph":
        """Row-stochastic copy: FP64 row sums, p = (float)(w / d)."""
        d = self.data.copy()
        load().synth_row_normalize(self.n, self.indptr.ctypes.data, d.ctypes.data)
        return Graph(self.n, self.indptr, self.indices, d)


def uniform_knn(n: int, k: int, seed: int) -> Graph:
    """k distinct neighbours per row, uniform over {0..n-1} minus i, sorted."""
    indptr = np.empty(n + 1, np.int64)
    indices = np.empty(n * k, np.int64)
    w = np.empty(n * k, np.float32)
    rc = load().synth_uniform_knn(n, k, seed, indptr.ctypes.data,
                                  indices.ctypes.data, w.ctypes.data)
    if rc != 0:
        raise ValueError("bad uniform_knn arguments")
    return Graph(n, indptr, indices, w)


def knn_graph(n: int, k: int, seed: int, dim: int = 10, n_clusters: int = 20,
              spread: float = 4.0) -> Graph:
    """Exact kNN of Gaussian-mixture points in R^dim, union symmetrised (max)."""
    indptr = np.empty(n + 1, np.int64)
    indices = np.empty(2 * n * k, np.int64)
    w = np.empty(2 * n * k, np.float32)
    nnz = load().synth_knn_graph(n, dim, k, n_clusters, spread, seed,
                                 indptr.ctypes.data, indices.ctypes.data,
                                 w.ctypes.data)
    if nnz < 0:
        raise ValueError("bad knn_graph arguments")
    return Graph(n, indptr, indices[:nnz].copy(), w[:nnz].copy())


def normal(n: int, F: int, seed: int, cols: Optional[Sequence[int]] = None,
           out: Optional[np.ndarray] = None) -> np.ndarray:
    """X[i, g] ~ N(0,1); with cols, only those columns (bit-identical)."""
    c = None if cols is None else np.ascontiguousarray(cols, np.int64)
    m = F if c is None else len(c)
    if out is None:
        out = np.empty((n, m), np.float32)
    load().synth_fill_normal(n, F, seed, None if c is None else c.ctypes.data,
                             m, out.ctypes.data)
    return out


def log1p_poisson(n: int, F: int, seed: int, cols: Optional[Sequence[int]] = None,
                  out: Optional[np.ndarray] = None) -> np.ndarray:
    """X[i, g] = log1p(Poisson(lambda_g)), lambda_g ~ LogNormal(-1, 1.5)."""
    c = None if cols is None else np.ascontiguousarray(cols, np.int64)
    m = F if c is None else len(c)
    if out is None:
        out = np.empty((n, m), np.float32)
    load().synth_fill_log1p_poisson(n, F, seed, None if c is None else c.ctypes.data,
                                    m, out.ctypes.data)
    return out


def batch_sizes(n_graphs: int, seed: int, lo: int = 2000, hi: int = 50000,
                kmin: int = 10, kmax: int = 30):
    """(n_i, k_i): n_i = round(exp(U(ln lo, ln hi))), k_i ~ U{kmin..kmax}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n_graphs))).astype(np.int64)
    k = rng.integers(kmin, kmax + 1, n_graphs).astype(np.int64)
    return n, k
