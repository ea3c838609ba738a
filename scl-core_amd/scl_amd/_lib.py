"""ctypes binding of libscl_f32_i64.so (the C ABI declared in include/*.h).

The library is the HIP build in ``scl-core_amd/lib``.  There is no fallback:
if it is missing, loading raises immediately, and every compute entry point
returns SCL_ERROR_FEATURE_UNAVAILABLE (41) when no GPU is present.

Mirrors the loader role of the reference's ``python/scl/core/config.py``
(107-162: pick ``libscl_{real}_{index}.so``; index type forced to int64).
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_DIR = PKG_ROOT / "lib"
LIB_NAME = "libscl_f32_i64.so"

# scl_core.h error codes (reference core.h:128-166)
SCL_OK = 0
SCL_ERROR_UNKNOWN = 1
SCL_ERROR_INTERNAL = 2
SCL_ERROR_OUT_OF_MEMORY = 3
SCL_ERROR_NULL_POINTER = 4
SCL_ERROR_INVALID_ARGUMENT = 10
SCL_ERROR_DIMENSION_MISMATCH = 11
SCL_ERROR_INDEX_OUT_OF_BOUNDS = 14
SCL_ERROR_NOT_IMPLEMENTED = 40
SCL_ERROR_FEATURE_UNAVAILABLE = 41

c_index = ctypes.c_int64
c_real = ctypes.c_float
c_size = ctypes.c_size_t
c_bool = ctypes.c_int
c_err = ctypes.c_int32
P = ctypes.c_void_p
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every function declared in include/*.h
SIGNATURES = {
    # scl_core.h
    "scl_get_version": (ctypes.c_char_p, []),
    "scl_get_build_config": (ctypes.c_char_p, []),
    "scl_get_last_error": (ctypes.c_char_p, []),
    "scl_get_last_error_code": (c_err, []),
    "scl_clear_error": (None, []),
    "scl_is_ok": (c_bool, [c_err]),
    "scl_is_error": (c_bool, [c_err]),
    # scl_sparse.h
    "scl_sparse_create": (c_err, [PP, c_index, c_index, c_index, P, P, P, c_bool]),
    "scl_sparse_create_with_strategy": (
        c_err, [PP, c_index, c_index, c_index, P, P, P, c_bool, ctypes.c_int]),
    "scl_sparse_wrap": (c_err, [PP, c_index, c_index, c_index, P, P, P, c_bool]),
    "scl_sparse_wrap_and_own": (c_err, [PP, c_index, c_index, c_index, P, P, P, c_bool]),
    "scl_sparse_clone": (c_err, [P, PP]),
    "scl_sparse_destroy": (c_err, [PP]),
    "scl_sparse_rows": (c_err, [P, ctypes.POINTER(c_index)]),
    "scl_sparse_cols": (c_err, [P, ctypes.POINTER(c_index)]),
    "scl_sparse_nnz": (c_err, [P, ctypes.POINTER(c_index)]),
    "scl_sparse_is_csr": (c_err, [P, ctypes.POINTER(c_bool)]),
    "scl_sparse_is_csc": (c_err, [P, ctypes.POINTER(c_bool)]),
    "scl_sparse_is_valid": (c_err, [P, ctypes.POINTER(c_bool)]),
    "scl_sparse_is_contiguous": (c_err, [P, ctypes.POINTER(c_bool)]),
    "scl_sparse_export": (c_err, [P, P, P, P]),
    "scl_sparse_get_lengths": (c_err, [P, PP, ctypes.POINTER(c_size)]),
    "scl_sparse_get_indices": (c_err, [P, PP, ctypes.POINTER(c_size)]),
    "scl_sparse_get_data": (c_err, [P, PP, ctypes.POINTER(c_size)]),
    "scl_sparse_transpose": (c_err, [P, PP]),
    "scl_sparse_to_contiguous": (c_err, [P, PP]),
    "scl_sparse_block_diag": (c_err, [P, c_size, PP]),
    # scl_diffusion.h
    "scl_diffusion_compute_transition_matrix": (c_err, [P, ctypes.c_int]),
    "scl_diffusion_diffuse_vector": (c_err, [P, P, c_size, c_index]),
    "scl_diffusion_diffuse_matrix": (c_err, [P, P, c_index, c_index, c_index]),
    "scl_diffusion_magic_impute": (c_err, [P, P, c_index, c_index, c_index]),
    "scl_diffusion_compute_dpt": (c_err, [P, c_index, P, c_size, c_index, c_real]),
    "scl_diffusion_compute_dpt_multi_root": (c_err, [P, P, c_size, P, c_size, c_index]),
    "scl_diffusion_random_walk_with_restart": (
        c_err, [P, P, c_size, P, c_size, c_real, c_index, c_real]),
    "scl_diffusion_personalized_pagerank": (
        c_err, [P, c_index, P, c_size, c_real, c_index, c_real]),
    "scl_diffusion_diffusion_map_embedding": (c_err, [P, P, c_index, c_index, c_index]),
    "scl_diffusion_heat_kernel_signature": (c_err, [P, P, c_size, c_real, c_index]),
    "scl_diffusion_diffusion_distance": (c_err, [P, P, c_size, c_index]),
    "scl_diffusion_lazy_random_walk": (c_err, [P, P, c_size, c_index, c_real]),
    # scl_impute.h, scl_pseudotime.h, scl_algebra.h
    "scl_impute_diffusion": (c_err, [P, P, c_index, c_index, c_index, P]),
    "scl_impute_magic": (c_err, [P, P, c_index, c_index, c_index, P]),
    "scl_pseudotime_diffusion": (c_err, [P, c_index, P, c_index, c_index]),
    "scl_algebra_spmm": (c_err, [P, P, c_index, P, c_real, c_real]),
    # scl_neighbors.h
    "scl_neighbors_compute_norms": (c_err, [P, P]),
    "scl_knn": (c_err, [P, P, c_size, c_size, P, P]),
    "scl_knn_simple": (c_err, [P, c_size, c_size, P, P]),
    # scl_kernel.h, scl_transition.h
    "scl_kernel_compute_matrix": (c_err, [P, P, c_real, ctypes.c_int]),
    "scl_kernel_row_sums": (c_err, [P, P, c_size, c_real, ctypes.c_int]),
    "scl_transition_matrix_from_velocity": (c_err, [P, c_index, P]),
    "scl_transition_row_normalize_to_stochastic": (c_err, [P, c_index, P, c_size]),
    "scl_transition_symmetrize": (c_err, [P, c_index, P]),
    "scl_transition_stationary_distribution": (c_err, [P, c_index, P, c_index, c_real]),
    # scl_diffusion_gpu.h
    "scl_gpu_device_count": (c_err, [ctypes.POINTER(ctypes.c_int)]),
    "scl_gpu_sparse_invalidate": (c_err, [P]),
    "scl_sparse_locality_order": (c_err, [P, P]),
    "scl_gpu_sparse_prepare": (c_err, [P]),
    "scl_gpu_diffuse_matrix_device": (
        c_err, [P, P, P, P, c_index, c_index, c_index, P]),
    "scl_gpu_spmm_device": (c_err, [P, P, P, c_index, c_index, ctypes.c_int, P]),
    "scl_gpu_compute_transition_matrix_ex": (c_err, [P, ctypes.c_int, P, P]),
    "scl_gpu_spmm_set_shape": (
        c_err, [P, c_index, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                ctypes.c_int]),
    "scl_gpu_spmm_get_shape": (c_err, [P, c_index, ctypes.c_int, P, P, P, P]),
    "scl_gpu_spmm_autotune": (c_err, [P, P, P, c_index, c_index, ctypes.c_int, P, P]),
    "scl_gpu_sparse_set_values_device": (c_err, [P, P, P]),
    "scl_gpu_diffuse_matrix_lowp_device": (
        c_err, [P, P, P, P, c_index, c_index, c_index, ctypes.c_int, ctypes.c_int, P]),
    "scl_gpu_lowp_scale_exp": (
        c_err, [P, P, c_index, c_index, c_index, ctypes.POINTER(ctypes.c_int), P]),
    "scl_gpu_diffuse_forward_ckpt_device": (
        c_err, [P, P, P, P, P, c_index, c_index, c_index, c_index, P]),
    "scl_gpu_edge_gradient_device": (
        c_err, [P, P, P, P, c_index, c_index, ctypes.c_int, P]),
    "scl_gpu_diffuse_backward_workspace": (
        c_err, [P, c_index, c_index, c_index, ctypes.c_int, ctypes.POINTER(c_size)]),
    "scl_gpu_plan_diffusion": (
        c_err, [P, P, P, P, P, P, P, P, P, P, c_size, c_index, c_index, c_index,
                c_index, PP]),
    "scl_gpu_plan_launch": (c_err, [P, P]),
    "scl_gpu_plan_destroy": (c_err, [PP]),
    "scl_gpu_plan_diffusion_batched": (
        c_err, [P, P, c_size, P, P, P, P, P, P, c_index, c_index, c_index, c_index, PP]),
    "scl_gpu_plan_check": (c_err, [P, P]),
    "scl_gpu_diffuse_backward_device": (
        c_err, [P, P, P, P, P, P, P, c_size, c_index, c_index, c_index, c_index, P]),
    "scl_gpu_diffuse_forward_ckpt_values_device": (
        c_err, [P, P, P, P, P, P, c_index, c_index, c_index, c_index, P]),
    "scl_gpu_diffuse_backward_values_workspace": (
        c_err, [P, c_index, c_index, c_index, ctypes.c_int, ctypes.POINTER(c_size)]),
    "scl_gpu_diffuse_backward_values_device": (
        c_err, [P, P, P, P, P, P, P, P, c_size, c_index, c_index, c_index, c_index, P]),
    "scl_gpu_sparse_generation": (c_err, [P, ctypes.POINTER(ctypes.c_uint64)]),
    "scl_gpu_lowp_workspace": (c_err, [c_index, c_index, ctypes.c_int, ctypes.POINTER(c_size)]),
    "scl_gpu_diffuse_matrix_device_noalloc": (
        c_err, [P, P, P, c_index, c_index, c_index, P]),
}

_lock = threading.Lock()
_lib = None


class SCLError(RuntimeError):
    """Raised for a non-zero scl_error_t (reference python/scl/core/error.py)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


def lib_path() -> Path:
    return LIB_DIR / LIB_NAME


def load() -> ctypes.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not path.exists():
            raise ImportError(
                f"{path} is missing: build it with `make -C scl-core_amd` "
                "(or __graft_entry__.build())")
        # PyTorch's ROCm wheel ships its own libamdhip64.so (same SONAME as
        # /opt/rocm's).  If this library were loaded first, torch would later
        # load its copy under another file name and the process would hold two
        # HIP runtimes, which do not share devices or memory.  Importing torch
        # first makes the dynamic linker bind this library to torch's runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(os.fspath(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(code: int) -> None:
    if code != SCL_OK:
        msg = load().scl_get_last_error().decode(errors="replace")
        raise SCLError(int(code), msg)


def get_last_error() -> str:
    return load().scl_get_last_error().decode(errors="replace")


def clear_error() -> None:
    load().scl_clear_error()


def device_count() -> int:
    out = ctypes.c_int(0)
    check(load().scl_gpu_device_count(ctypes.byref(out)))
    return out.value
