"""Test configuration: import paths, the `gpu` marker, shared helpers."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "scl-core_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available() -> bool:
    try:
        import scl_amd
        return scl_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips GPU tests loudly only when explicitly run without a device."""
    import scl_amd
    n = scl_amd.device_count()
    if n == 0:
        pytest.fail("GPU test run on a host without a HIP device")
    return n


def csr_from_dense(A: np.ndarray):
    """(indptr, indices, data) of a dense matrix, stored order = column order."""
    indptr = [0]
    indices, data = [], []
    for row in A:
        nz = np.nonzero(row)[0]
        indices.extend(nz.tolist())
        data.extend(row[nz].tolist())
        indptr.append(len(indices))
    return (np.array(indptr, np.int64), np.array(indices, np.int64),
            np.array(data, np.float32))


def dense_from_csr(n, indptr, indices, data, n_cols=None):
    A = np.zeros((n, n if n_cols is None else n_cols), np.float64)
    for i in range(n):
        for k in range(indptr[i], indptr[i + 1]):
            A[i, indices[k]] += float(data[k])
    return A
