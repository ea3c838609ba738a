"""scl_amd -- MI355X build of scl-core's graph-diffusion path.

Python face of ``libscl_f32_i64.so`` (C ABI in ``include/``): sparse handles
(``CsrMatrix``/``CscMatrix``), the drop-in host-array diffusion calls
(``diffusion``), the other SpMM callers (``callers``: imputation,
pseudotime, algebra spmm), the operator's construction (``construct``:
kernel weights, transition matrices, stationary distribution), device-tensor
calls for torch (``torch_ops``), and seeded workload generators (``synth``).  Importing this package loads the HIP
library; there is no CPU fallback.
"""
from ._lib import (SCLError, check, clear_error, device_count, get_last_error,
                   lib_path, load)
from .sparse import CscMatrix, CsrMatrix, SparseMatrix, SparseMatrixBase
from . import callers, construct, diffusion

__version__ = "1.0.0"

load()

__all__ = [
    "SCLError", "check", "clear_error", "device_count", "get_last_error",
    "lib_path", "load", "CsrMatrix", "CscMatrix", "SparseMatrix",
    "SparseMatrixBase", "callers", "construct", "diffusion", "__version__",
]
