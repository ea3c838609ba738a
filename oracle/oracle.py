"""ctypes loader of the CPU restatement / FP64 oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product never imports it.

The AVX-512 build (liboracle_v4.so) is used when the host CPU has the
x86-64-v4 features, else the AVX2 build; both give identical results.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
D = ctypes.c_double
SZ = ctypes.c_size_t


def _has_v4() -> bool:
    try:
        flags = set()
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
        return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    except OSError:
        return False


def load():
    global _lib
    if _lib is not None:
        return _lib
    name = "liboracle_v4.so" if _has_v4() else "liboracle_v3.so"
    path = HERE / name
    if not path.exists():
        raise ImportError(f"{path} is missing: run `make -C oracle`")
    lib = ctypes.CDLL(os.fspath(path))
    sig = {
        "orc_spmm_step_f32": (None, [I64, I64, P, P, P, P, P]),
        "orc_diffuse_matrix_f32": (ctypes.c_int, [I64, I64, I64, P, P, P, P]),
        "orc_diffuse_columns_f32": (ctypes.c_int, [I64, I64, I64, P, P, P, P]),
        "orc_diffuse_vector_f32": (ctypes.c_int, [I64, I64, P, P, P, P]),
        "orc_transition_f32": (None, [I64, P, P, P, ctypes.c_int]),
        "orc_transition_f64": (None, [I64, P, P, P, ctypes.c_int, P]),
        "orc_diffuse_f64": (ctypes.c_int, [I64, I64, I64, P, P, P, P, P]),
        "orc_diffuse_transpose_f64": (ctypes.c_int, [I64, I64, I64, P, P, P, P, P]),
        "orc_grad_f64": (D, [I64, I64, I64, P, P, P, P, P, P, P, P]),
        "orc_loss_f64": (D, [I64, I64, I64, P, P, P, P, P]),
        "orc_stationarity_residual_f64": (D, [I64, P, P, P, P]),
        "orc_rel_err_f32_f64": (D, [SZ, P, P]),
        "orc_rel_err_f32_f32": (D, [SZ, P, P]),
        "orc_median_rel_err_f32_f64": (D, [SZ, P, P]),
        "orc_num_threads": (ctypes.c_int, []),
        "orc_diffuse_bf16_f32": (ctypes.c_int, [I64, I64, I64, P, P, P, ctypes.c_int, P, P]),
        "orc_bf16_round": (ctypes.c_float, [ctypes.c_float]),
        "orc_fp16_round": (ctypes.c_float, [ctypes.c_float]),
        "orc_comp_round": (ctypes.c_float, [ctypes.c_float]),
        "orc_set_num_threads": (None, [ctypes.c_int]),
        "orc_diffuse_lowp_f32": (ctypes.c_int, [I64, I64, I64, P, P, P, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 P, P]),
        "orc_lazy_walk_f32": (ctypes.c_int, [I64, P, P, P, P, I64, ctypes.c_float]),
        "orc_heat_kernel_f32": (ctypes.c_int, [I64, P, P, P, P, ctypes.c_float, I64]),
        "orc_rwr_f32": (I64, [I64, P, P, P, P, I64, P, ctypes.c_float, I64, ctypes.c_float]),
        "orc_dpt_f32": (ctypes.c_int, [I64, P, P, P, I64, P, I64, ctypes.c_float]),
        "orc_dpt_multi_f32": (ctypes.c_int, [I64, P, P, P, P, I64, P, I64]),
        "orc_dmap_init_f32": (None, [I64, I64, P]),
        "orc_dmap_f32": (ctypes.c_int, [I64, P, P, P, P, I64, I64]),
        "orc_dmap_f64": (ctypes.c_int, [I64, P, P, P, P, I64, I64]),
        "orc_pseudotime_f64": (ctypes.c_int, [I64, P, P, P, I64, I64, I64, P]),
        "orc_mgs_f32": (ctypes.c_int, [I64, I64, P]),
        "orc_diffusion_distance_f32": (ctypes.c_int, [I64, P, P, P, P, I64]),
        "orc_densify_f32": (None, [I64, I64, P, P, P, ctypes.c_int, P]),
        "orc_impute_diffusion_f32": (ctypes.c_int, [I64, I64, I64, P, P, P, P]),
        "orc_magic_f32": (ctypes.c_int, [I64, I64, I64, P, P, P, P]),
        "orc_fastrng_seed": (None, [ctypes.c_uint64, P, P]),
        "orc_fastrng_step": (None, [P, I64, P]),
        "orc_fastrng_uniform": (None, [ctypes.c_uint64, I64, P]),
        "orc_pt_orthonormalise_f32": (ctypes.c_int, [I64, I64, P]),
        "orc_pseudotime_f32": (ctypes.c_int, [I64, P, P, P, I64, I64, I64, P]),
        "orc_algebra_spmm_f32": (None, [I64, I64, P, P, P, P, P, ctypes.c_float,
                                        ctypes.c_float]),
        "orc_kernel_matrix_f32": (None, [I64, P, ctypes.c_float, ctypes.c_int, P]),
        "orc_kernel_row_sums_f32": (None, [I64, P, P, ctypes.c_float, ctypes.c_int, P]),
        "orc_vector_sum8_f32": (ctypes.c_float, [P, I64]),
        "orc_row_normalize_f32": (None, [I64, P, P, P]),
        "orc_symmetrize_f32": (None, [I64, P, P, P, P]),
        "orc_from_velocity_f32": (None, [I64, P, P, P, P]),
        "orc_stationary_f32": (I64, [I64, P, P, P, I64, ctypes.c_float, P]),
        "orc_sum_squared_f32": (ctypes.c_float, [P, I64]),
        "orc_norms_f32": (None, [I64, P, P, P]),
        "orc_knn_f32": (ctypes.c_int, [I64, P, P, P, P, I64, P, P]),
    }
    for k, (res, args) in sig.items():
        fn = getattr(lib, k)
        fn.restype = res
        fn.argtypes = args
    lib.variant = name
    _lib = lib
    return lib


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def spmm_step_f32(indptr, indices, vals, X):
    X = _f32(X)
    n, F = X.shape
    Y = np.empty_like(X)
    ip, ix, v = _i64(indptr), _i64(indices), _f32(vals)
    load().orc_spmm_step_f32(n, F, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                             X.ctypes.data, Y.ctypes.data)
    return Y


def diffuse_matrix_f32(indptr, indices, vals, X, t):
    """Reference-order FP32 P^t X (returns a new array)."""
    Y = _f32(X).copy()
    n = Y.shape[0]
    F = Y.shape[1] if Y.ndim == 2 else 1
    ip, ix, v = _i64(indptr), _i64(indices), _f32(vals)
    rc = load().orc_diffuse_matrix_f32(n, F, t, ip.ctypes.data, ix.ctypes.data,
                                       v.ctypes.data, Y.ctypes.data)
    if rc:
        raise MemoryError("oracle allocation failed")
    return Y


def diffuse_bf16_f32(indptr, indices, vals, X, t, bf16_values=False):
    """bf16-iterate restatement (fp32 accumulate, fp32 in and out)."""
    X = _f32(X)
    n, F = X.shape
    Y = np.empty_like(X)
    ip, ix, v = _i64(indptr), _i64(indices), _f32(vals)
    rc = load().orc_diffuse_bf16_f32(n, F, t, ip.ctypes.data, ix.ctypes.data,
                                     v.ctypes.data, 1 if bf16_values else 0,
                                     X.ctypes.data, Y.ctypes.data)
    if rc:
        raise MemoryError("oracle allocation failed")
    return Y


def diffuse_lowp_f32(indptr, indices, vals, X, t, fmt="bf16", bf16_values=False,
                     scale_exp=0, round_input=False):
    """Reduced-precision-iterate restatement ("bf16", scaled "fp16" or
    "comp", compensated bf16); round_input stores X in the same format before
    the first step."""
    X = _f32(X)
    n, F = X.shape
    Y = np.empty_like(X)
    ip, ix, v = _i64(indptr), _i64(indices), _f32(vals)
    code = {"bf16": 0, "fp16": 1, "comp": 2}[fmt]
    rc = load().orc_diffuse_lowp_f32(n, F, t, ip.ctypes.data, ix.ctypes.data,
                                     v.ctypes.data, 1 if bf16_values else 0,
                                     code, int(scale_exp),
                                     1 if round_input else 0,
                                     X.ctypes.data, Y.ctypes.data)
    if rc:
        raise MemoryError("oracle allocation failed")
    return Y


def fp16_round(a):
    """Nearest-even binary16 rounding of a float32 array via the oracle."""
    a = _f32(a)
    f = load().orc_fp16_round
    return np.array([f(float(x)) for x in a.ravel()], np.float32).reshape(a.shape)


def comp_round(a):
    """Compensated-bf16 round trip of a float32 array via the oracle."""
    a = _f32(a)
    f = load().orc_comp_round
    return np.array([f(float(x)) for x in a.ravel()], np.float32).reshape(a.shape)


def bf16_round(a):
    """Nearest-even bf16 rounding of a float32 array (returned as float32)."""
    a = _f32(a)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + (((u >> 16) & 1) + 0x7FFF)) & 0xFFFF0000)
    r = np.where(nan, (u | 0x00400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32)


def diffuse_vector_f32(indptr, indices, vals, x, t):
    y = _f32(x).copy()
    ip, ix, v = _i64(indptr), _i64(indices), _f32(vals)
    rc = load().orc_diffuse_vector_f32(y.shape[0], t, ip.ctypes.data, ix.ctypes.data,
                                       v.ctypes.data, y.ctypes.data)
    if rc:
        raise MemoryError("oracle allocation failed")
    return y


def transition_f32(indptr, indices, vals, symmetric=False):
    v = _f32(vals).copy()
    ip, ix = _i64(indptr), _i64(indices)
    load().orc_transition_f32(len(ip) - 1, ip.ctypes.data, ix.ctypes.data,
                              v.ctypes.data, 1 if symmetric else 0)
    return v


def transition_f64(indptr, indices, vals, symmetric=False):
    v = _f32(vals).copy()
    ip, ix = _i64(indptr), _i64(indices)
    deg = np.empty(len(ip) - 1, np.float64)
    load().orc_transition_f64(len(ip) - 1, ip.ctypes.data, ix.ctypes.data,
                              v.ctypes.data, 1 if symmetric else 0, deg.ctypes.data)
    return v, deg


def diffuse_f64(indptr, indices, vals, X, t, transpose=False):
    X = _f64(X)
    n = X.shape[0]
    F = X.shape[1] if X.ndim == 2 else 1
    Y = np.empty_like(X)
    ip, ix, v = _i64(indptr), _i64(indices), _f64(vals)
    fn = load().orc_diffuse_transpose_f64 if transpose else load().orc_diffuse_f64
    rc = fn(n, F, t, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
            X.ctypes.data, Y.ctypes.data)
    if rc:
        raise MemoryError("oracle allocation failed")
    return Y


def grad_f64(indptr, indices, vals, X, T, t):
    """Returns (L, Y, dX, dP) of L = 0.5 ||P^t X - T||^2 in FP64."""
    X, T = _f64(X), _f64(T)
    n, F = X.shape
    ip, ix, v = _i64(indptr), _i64(indices), _f64(vals)
    Y, dX = np.empty_like(X), np.empty_like(X)
    dP = np.empty(max(len(ix), 1), np.float64)
    L = load().orc_grad_f64(n, F, t, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                            X.ctypes.data, T.ctypes.data, Y.ctypes.data,
                            dX.ctypes.data, dP.ctypes.data)
    return L, Y, dX, dP[:len(ix)]


def loss_f64(indptr, indices, vals, X, T, t):
    X, T = _f64(X), _f64(T)
    n, F = X.shape
    ip, ix, v = _i64(indptr), _i64(indices), _f64(vals)
    return load().orc_loss_f64(n, F, t, ip.ctypes.data, ix.ctypes.data,
                               v.ctypes.data, X.ctypes.data, T.ctypes.data)


def stationarity_residual_f64(indptr, indices, vals, deg):
    ip, ix, v, d = _i64(indptr), _i64(indices), _f32(vals), _f64(deg)
    return load().orc_stationarity_residual_f64(len(ip) - 1, ip.ctypes.data,
                                                ix.ctypes.data, v.ctypes.data,
                                                d.ctypes.data)


def ulp_distance(a, b) -> np.ndarray:
    """Per-element distance in fp32 ulps between two float32 arrays (the
    number of representable floats between them; +0 and -0 coincide)."""
    def ordered(x):
        i = _f32(x).ravel().view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7FFFFFFF), i)
    return np.abs(ordered(a) - ordered(b))


def ulps_bound_fp32_transition(indptr, indices, symmetric=False) -> np.ndarray:
    """First-order bound, in ulps, of how far the reference's FP32
    normalisation (scl/kernel/diffusion.hpp:452-603: sequential fp32 row sum
    d_i, then a_ij / d_i or a_ij d_i^-1/2 d_j^-1/2) can lie from the
    correctly rounded FP64 result: a sum of L terms is off by at most
    (L - 1) u relative, each further operation by u, and a relative error e
    is at most 2^24 e ulps.  Row mode: L_i + 2; symmetric: (L_i + L_j)/2 + 4."""
    ip, ix = _i64(indptr), _i64(indices)
    lens = np.diff(ip)
    row = np.repeat(np.arange(len(lens)), lens)
    if symmetric:
        return (lens[row] + lens[ix] + 1) // 2 + 4
    return lens[row] + 2


def rel_err(a, b) -> float:
    """max|a-b| / max(max|a|, max|b|) (reference precision.hpp:102-120)."""
    a = _f32(a).ravel()
    if np.asarray(b).dtype == np.float64:
        b = _f64(b).ravel()
        return load().orc_rel_err_f32_f64(a.size, a.ctypes.data, b.ctypes.data)
    b = _f32(b).ravel()
    return load().orc_rel_err_f32_f32(a.size, a.ctypes.data, b.ctypes.data)


def median_rel_err(a, b) -> float:
    a, b = _f32(a).ravel(), _f64(b).ravel()
    return load().orc_median_rel_err_f32_f64(a.size, a.ctypes.data, b.ctypes.data)


def num_threads() -> int:
    return load().orc_num_threads()


def set_num_threads(n: int) -> None:
    load().orc_set_num_threads(int(n))


# ---------------------------------------------------- vector-side consumers --

def _csr(indptr, indices, vals):
    return _i64(indptr), _i64(indices), _f32(vals)


def lazy_walk_f32(indptr, indices, vals, x, n_steps, laziness):
    ip, ix, v = _csr(indptr, indices, vals)
    y = _f32(x).copy()
    if load().orc_lazy_walk_f32(len(ip) - 1, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                                y.ctypes.data, n_steps, laziness):
        raise MemoryError
    return y


def heat_kernel_f32(indptr, indices, vals, t, n_steps):
    ip, ix, v = _csr(indptr, indices, vals)
    sig = np.empty(len(ip) - 1, np.float32)
    if load().orc_heat_kernel_f32(len(ip) - 1, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                                  sig.ctypes.data, t, n_steps):
        raise MemoryError
    return sig


def rwr_f32(indptr, indices, vals, seeds, alpha, max_iter, tol):
    """Returns (scores, iterations run)."""
    ip, ix, v = _csr(indptr, indices, vals)
    sd = _i64(seeds)
    out = np.empty(len(ip) - 1, np.float32)
    it = load().orc_rwr_f32(len(ip) - 1, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                            sd.ctypes.data, len(sd), out.ctypes.data, alpha, max_iter, tol)
    if it < 0:
        raise MemoryError
    return out, it


def dpt_f32(indptr, indices, vals, root, max_iter, tol):
    ip, ix, v = _csr(indptr, indices, vals)
    out = np.empty(len(ip) - 1, np.float32)
    if load().orc_dpt_f32(len(ip) - 1, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, root,
                          out.ctypes.data, max_iter, tol):
        raise MemoryError
    return out


def dpt_multi_f32(indptr, indices, vals, roots, max_iter):
    ip, ix, v = _csr(indptr, indices, vals)
    r = _i64(roots)
    out = np.empty(len(ip) - 1, np.float32)
    if load().orc_dpt_multi_f32(len(ip) - 1, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                                r.ctypes.data, len(r), out.ctypes.data, max_iter):
        raise MemoryError
    return out


def dmap_init_f32(n, K):
    Q = np.empty((n, K), np.float32)
    load().orc_dmap_init_f32(n, K, Q.ctypes.data)
    return Q


def dmap_f64(indptr, indices, vals, K, n_iter):
    """dmap_f32's steps in FP64 from the same fp32 start vectors (an anchor)."""
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    out = np.empty((n, K), np.float64)
    if load().orc_dmap_f64(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, out.ctypes.data,
                           K, n_iter):
        raise MemoryError
    return out


def dmap_f32(indptr, indices, vals, K, n_iter):
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    out = np.empty((n, K), np.float32)
    if load().orc_dmap_f32(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, out.ctypes.data,
                           K, n_iter):
        raise MemoryError
    return out


def mgs_f32(Q):
    Q = _f32(Q).copy()
    if load().orc_mgs_f32(Q.shape[0], Q.shape[1], Q.ctypes.data):
        raise MemoryError
    return Q


def diffusion_distance_f32(indptr, indices, vals, n_steps):
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    D = np.empty((n, n), np.float32)
    if load().orc_diffusion_distance_f32(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                                         D.ctypes.data, n_steps):
        raise MemoryError
    return D


# ------------------------------------------------ other callers of the SpMM --

def densify_f32(indptr, indices, vals, n_cells, n_genes, is_csr=True):
    """Dense cells x genes copy of a sparse matrix whose primary slots are
    cells (CSR) or genes (CSC)."""
    ip, ix, v = _csr(indptr, indices, vals)
    out = np.empty((n_cells, n_genes), np.float32)
    load().orc_densify_f32(n_cells, n_genes, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                           int(bool(is_csr)), out.ctypes.data)
    return out


def impute_diffusion_f32(indptr, indices, vals, X, n_steps):
    ip, ix, v = _csr(indptr, indices, vals)
    X = _f32(X).copy()
    if load().orc_impute_diffusion_f32(X.shape[0], X.shape[1], n_steps, ip.ctypes.data,
                                       ix.ctypes.data, v.ctypes.data, X.ctypes.data):
        raise MemoryError
    return X


def magic_f32(indptr, indices, vals, X, t):
    ip, ix, v = _csr(indptr, indices, vals)
    X = _f32(X).copy()
    if load().orc_magic_f32(X.shape[0], X.shape[1], t, ip.ctypes.data, ix.ctypes.data,
                            v.ctypes.data, X.ctypes.data):
        raise MemoryError
    return X


def fastrng_seed(seed):
    """(the four 64-bit seeding values z, the 4-word state)."""
    z = np.empty(4, np.uint64)
    st = np.empty(4, np.uint32)
    load().orc_fastrng_seed(seed, z.ctypes.data, st.ctypes.data)
    return z, st


def fastrng_step(state, m):
    """m raw outputs from a 4-word state; returns (outputs, new state)."""
    st = np.ascontiguousarray(state, dtype=np.uint32).copy()
    out = np.empty(m, np.uint32)
    load().orc_fastrng_step(st.ctypes.data, m, out.ctypes.data)
    return out, st


def fastrng_uniform(seed, m):
    out = np.empty(m, np.float32)
    load().orc_fastrng_uniform(seed, m, out.ctypes.data)
    return out


def pt_orthonormalise_f32(Q):
    Q = _f32(Q).copy()
    if load().orc_pt_orthonormalise_f32(Q.shape[0], Q.shape[1], Q.ctypes.data):
        raise MemoryError
    return Q


def pseudotime_f64(indptr, indices, vals, root, n_dcs, n_iter=100):
    """pseudotime_f32's steps in FP64 from the same fp32 start (an anchor)."""
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    out = np.empty(n, np.float64)
    if load().orc_pseudotime_f64(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, root,
                                 n_dcs, n_iter, out.ctypes.data):
        raise MemoryError
    return out


def pseudotime_f32(indptr, indices, vals, root, n_dcs, n_iter=100):
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    out = np.empty(n, np.float32)
    if load().orc_pseudotime_f32(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, root,
                                 n_dcs, n_iter, out.ctypes.data):
        raise MemoryError
    return out


def algebra_spmm_f32(indptr, indices, vals, X, Y, alpha, beta):
    """Y = alpha A X + beta Y for A with len(indptr) - 1 rows; returns a new Y."""
    ip, ix, v = _csr(indptr, indices, vals)
    X = _f32(X)
    Y = _f32(Y).copy()
    F = Y.shape[1] if Y.ndim == 2 else 1
    load().orc_algebra_spmm_f32(len(ip) - 1, F, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                                X.ctypes.data, Y.ctypes.data, alpha, beta)
    return Y


# ------------------------------------------------ construction of the operator --

def kernel_matrix_f32(d, bandwidth, kind):
    d = _f32(d)
    out = np.empty_like(d)
    load().orc_kernel_matrix_f32(len(d), d.ctypes.data, bandwidth, kind, out.ctypes.data)
    return out


def kernel_row_sums_f32(indptr, d, bandwidth, kind):
    ip, d = _i64(indptr), _f32(d)
    out = np.empty(len(ip) - 1, np.float32)
    load().orc_kernel_row_sums_f32(len(ip) - 1, ip.ctypes.data, d.ctypes.data, bandwidth, kind,
                                   out.ctypes.data)
    return out


def vector_sum8_f32(v):
    v = _f32(v)
    return float(np.float32(load().orc_vector_sum8_f32(v.ctypes.data, len(v))))


def row_normalize_f32(indptr, vals):
    ip, v = _i64(indptr), _f32(vals)
    out = np.empty_like(v)
    load().orc_row_normalize_f32(len(ip) - 1, ip.ctypes.data, v.ctypes.data, out.ctypes.data)
    return out


def symmetrize_f32(indptr, indices, vals):
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    S = np.empty((n, n), np.float32)
    load().orc_symmetrize_f32(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, S.ctypes.data)
    return S


def from_velocity_f32(indptr, indices, vals):
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    S = np.empty((n, n), np.float32)
    load().orc_from_velocity_f32(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data,
                                 S.ctypes.data)
    return S


def stationary_f32(indptr, indices, vals, max_iter, tol):
    """Returns (pi, iterations run)."""
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    pi = np.empty(n, np.float32)
    it = load().orc_stationary_f32(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, max_iter,
                                   tol, pi.ctypes.data)
    if it < 0:
        raise MemoryError
    return pi, it


# ------------------------------------------------------ k-nearest neighbours --

def sum_squared_f32(v):
    v = _f32(v)
    return float(np.float32(load().orc_sum_squared_f32(v.ctypes.data, len(v))))


def norms_f32(indptr, vals):
    ip, v = _i64(indptr), _f32(vals)
    out = np.empty(len(ip) - 1, np.float32)
    load().orc_norms_f32(len(ip) - 1, ip.ctypes.data, v.ctypes.data, out.ctypes.data)
    return out


def knn_f32(indptr, indices, vals, k, norms=None):
    """(indices n x k int64, distances n x k float32), as the reference."""
    ip, ix, v = _csr(indptr, indices, vals)
    n = len(ip) - 1
    nr = norms_f32(ip, v) if norms is None else _f32(norms)
    idx = np.empty((n, k), np.int64)
    dist = np.empty((n, k), np.float32)
    if load().orc_knn_f32(n, ip.ctypes.data, ix.ctypes.data, v.ctypes.data, nr.ctypes.data, k,
                          idx.ctypes.data, dist.ctypes.data):
        raise MemoryError
    return idx, dist
