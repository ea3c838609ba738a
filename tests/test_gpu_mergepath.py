"""GPU parity of the merge-path SpMM shape (SCL_GPU_SHAPE_MERGE_PATH) on
skewed patterns: hub rows of half the matrix, heavy-tailed row lengths and
empty rows.  Rows that one wave finishes are the row-split FMA chain and must
be bit-identical to the restatement; rows cut across waves are p chains added
in order and must agree to fp32 rounding."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scl_amd import CsrMatrix, CscMatrix, torch_ops

pytestmark = pytest.mark.gpu


def skewed(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    deg = np.minimum((rng.pareto(1.2, n) * 3).astype(np.int64), n - 1)
    deg[rng.integers(0, n, 3)] = n // 2          # hubs
    deg[rng.integers(0, n, 50)] = 0              # empty rows
    ip = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    ix = np.concatenate([np.sort(rng.choice(n, d, replace=False)) for d in deg]).astype(np.int64)
    return ip, ix, rng.random(len(ix)).astype(np.float32)


def cut_rows(ip, m):
    """Rows a merge-path launch with code m splits across waves (the path
    search of the device, on the host)."""
    n, nnz = len(ip) - 1, int(ip[-1])
    items = 256 << (m - 1)
    waves = max((n + nnz + items - 1) // items, 1)
    cut = set()
    for w in range(1, waves):
        d = min(w * items, n + nnz)
        lo, hi = max(d - nnz, 0), min(d, n)
        while lo < hi:
            mid = (lo + hi) // 2
            if ip[mid + 1] <= d - mid - 1:
                lo = mid + 1
            else:
                hi = mid
        i, k = lo, d - lo
        if i < n and ip[i] < k < ip[i + 1]:
            cut.add(i)
    return np.array(sorted(cut), np.int64)


@pytest.mark.parametrize("F", [1, 6, 64, 300])
@pytest.mark.parametrize("m", [1, 3, 8])
def test_merge_path_matches_restatement(gpu, F, m):
    n = 6000
    ip, ix, v = skewed(n, seed=F + m)
    X = np.random.Generator(np.random.PCG64(F)).standard_normal((n, F)).astype(np.float32)
    ref = orc.spmm_step_f32(ip, ix, v, X)
    P = CsrMatrix.from_arrays(ip, ix, v, (n, n))
    vec = 4 if F % 4 == 0 else (2 if F % 2 == 0 else 1)
    torch_ops.set_shape(P, F, vec, 64, 8, merge=m, nontemporal=m == 3)
    assert torch_ops.get_shape(P, F)["merge"] == m
    Y = torch_ops.spmm(P, torch.from_numpy(X).cuda()).cpu().numpy()
    cut = cut_rows(ip, m)
    whole = np.setdiff1d(np.arange(n), cut)
    assert len(cut) > 0                       # the hubs are always cut
    assert np.array_equal(Y[whole], ref[whole])
    # cut rows: within the 1e-5 bar of the CPU path, and no further from the
    # FP64 result than the stored-order chain itself is (hub rows of 3000
    # nonzeros put both ~1e-6 away)
    ref64 = orc.diffuse_f64(ip, ix, v, X, 1)
    assert orc.rel_err(Y, ref) <= 1e-5
    assert orc.rel_err(Y, ref64) <= 2 * max(orc.rel_err(ref, ref64), 1e-7)
    # a second launch on the same scratch gives the same bits
    Y2 = torch_ops.spmm(P, torch.from_numpy(X).cuda()).cpu().numpy()
    assert np.array_equal(Y, Y2)


def test_merge_path_transposed_and_diffusion(gpu):
    n, F, t = 5000, 32, 3
    ip, ix, v = skewed(n, seed=11)
    X = np.random.Generator(np.random.PCG64(3)).standard_normal((n, F)).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    # transposed direction of a CSR handle: P^T X over the transposed pattern
    P = CsrMatrix.from_arrays(ip, ix, v, (n, n))
    torch_ops.set_shape(P, F, 4, 64, 8, transpose=True, merge=2)
    via_t = torch_ops.spmm(P, Xd, transpose=True).cpu().numpy()
    ref_t = orc.diffuse_f64(ip, ix, v, X, 1, transpose=True)
    assert orc.rel_err(via_t, ref_t) <= 1e-5
    # a CSC handle over the arrays of a matrix M applies M^T (the reference's
    # visit()): with P's CSR arrays read as CSC, M^T = P again, walked through
    # the handle's own primary slot
    C = CscMatrix.from_arrays(ip, ix, v, (n, n))
    torch_ops.set_shape(C, F, 4, 64, 8, merge=2)
    via_csc = torch_ops.spmm(C, Xd).cpu().numpy()
    assert orc.rel_err(via_csc, orc.spmm_step_f32(ip, ix, v, X)) <= 1e-5
    # t steps through the device diffusion entry point
    torch_ops.set_shape(P, F, 4, 64, 8, merge=1)
    Y = torch_ops.diffuse(P, Xd, t).cpu().numpy()
    assert orc.rel_err(Y, orc.diffuse_matrix_f32(ip, ix, v, X, t)) <= 1e-5


def test_merge_path_shape_errors(gpu):
    n = 2000
    ip, ix, v = skewed(n, seed=5)
    P = CsrMatrix.from_arrays(ip, ix, v, (n, n))
    with pytest.raises(Exception):
        torch_ops.set_shape(P, 64, 4, 32, 8, merge=2)        # group must be 64
    with pytest.raises(Exception):
        torch_ops.set_shape(P, 64, 4, 64, 8, merge=2, adjoint=True)
    with pytest.raises(Exception):
        torch_ops.set_shape(P, 64, 4, 64, 8, merge=9)
