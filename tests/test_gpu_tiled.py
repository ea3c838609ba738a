"""LDS-tiled SpMM (rows cut into tiles whose column union fits LDS; every
slab of the union's X rows is copied to LDS once per tile and each row reads
its neighbours there in stored order).  Bar: bit-identical to the CPU
restatement of spmm_block (scl/kernel/diffusion.hpp:227-278 over axpy_simd
100-134) for every lane layout, LDS budget and slab chunk, including tiles
cut by the entry budget, empty rows, feature counts that end inside a slab,
rows in natural and in locality order, and after the values change."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from scl_amd import CsrMatrix, callers, diffusion, synth, torch_ops

pytestmark = pytest.mark.gpu

LAYOUTS = [(lpr, b, u, nt) for lpr in (4, 8) for b in (1, 2) for u, nt in ((4, False), (16, True))]


def with_empty_rows(g, every):
    """The graph with every `every`-th row emptied (its entries dropped)."""
    ip, ix, d = g.indptr, g.indices, g.data
    keep = np.ones(len(ix), bool)
    lens = np.diff(ip)
    for i in range(0, g.n, every):
        keep[ip[i]:ip[i + 1]] = False
        lens[i] = 0
    nip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    return nip, ix[keep], d[keep]


@pytest.mark.parametrize("F", [4, 36, 260, 2000])
def test_tiled_is_bit_identical(gpu, F):
    import torch
    cases = []
    g = synth.knn_graph(20_000, 12, seed=F)            # locality schedule
    cases.append(("knn20k", g.indptr, g.indices, g.data, g.n))
    g = synth.uniform_knn(3_001, 12, seed=F + 1)        # natural order
    cases.append(("uniform3001", g.indptr, g.indices, g.data, g.n))
    g = synth.uniform_knn(1_500, 150, seed=F + 2)       # tiles cut by entries
    cases.append(("long1500", g.indptr, g.indices, g.data, g.n))
    g = synth.knn_graph(17_000, 9, seed=F + 3)
    cases.append(("empty17k",) + with_empty_rows(g, 7) + (g.n,))
    for name, ip, ix, d, n in cases:
        A = CsrMatrix.from_arrays(ip, ix, d, (n, n))
        X = synth.normal(n, F, seed=F + n)
        Xd = torch.from_numpy(X).cuda()
        ref = orc.spmm_step_f32(ip, ix, d, X)
        for lpr, b, u, nt in LAYOUTS:
            torch_ops.set_shape(A, F, 4, lpr, u, rows_per_group=torch_ops.TILED,
                                wg_per_cu=b, nontemporal=nt)
            sh = torch_ops.get_shape(A, F)
            assert sh["rows_per_group"] == torch_ops.TILED and sh["group"] == lpr
            Y = torch_ops.spmm(A, Xd).cpu().numpy()
            assert np.array_equal(Y, ref), (name, F, lpr, b, u, nt)
    # values change in place: the tiled form re-gathers its copy of them
    g = synth.knn_graph(20_000, 12, seed=F + 4)
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n))
    torch_ops.set_shape(A, F, 4, 4, 16, rows_per_group=torch_ops.TILED, wg_per_cu=1)
    diffusion.compute_transition_matrix(A)
    ip, ix, p = A.export()
    X = synth.normal(g.n, F, seed=F + 5)
    Y = torch_ops.diffuse(A, torch.from_numpy(X).cuda(), 3).cpu().numpy()
    assert torch_ops.get_shape(A, F)["rows_per_group"] == torch_ops.TILED
    assert np.array_equal(Y, orc.diffuse_matrix_f32(ip, ix, p, X, 3))


def test_tiled_with_other_values_and_errors(gpu):
    """MAGIC runs the operator's pattern with its own weights: under a tiled
    shape (whose form holds a copy of the operator's values) it falls back to
    a shape that reads them, bit-identical.  The tiled form needs sorted rows,
    vec 4, the forward direction and rows that fit one tile."""
    n, G = 20_000, 64
    g = synth.knn_graph(n, 10, seed=23)
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    torch_ops.set_shape(A, G, 4, 4, 8, rows_per_group=torch_ops.TILED, wg_per_cu=2)
    X = synth.normal(n, G, seed=24)
    ip = np.arange(0, n * G + 1, G, dtype=np.int64)
    ix = np.tile(np.arange(G, dtype=np.int64), n)
    M = CsrMatrix.from_arrays(ip, ix, X.ravel(), (n, G))
    got = callers.impute_magic(M, A, n, G, 2)
    assert np.array_equal(got, orc.magic_f32(g.indptr, g.indices, g.data, X, 2))
    bad = CsrMatrix.from_arrays(np.array([0, 2, 3], np.int64), np.array([1, 0, 1], np.int64),
                                np.ones(3, np.float32), (2, 2))
    with pytest.raises(Exception):
        torch_ops.set_shape(bad, 8, 4, 4, 8, rows_per_group=torch_ops.TILED, wg_per_cu=1)
    with pytest.raises(Exception):   # vec 2
        torch_ops.set_shape(A, 8, 2, 4, 8, rows_per_group=torch_ops.TILED, wg_per_cu=1)
    with pytest.raises(Exception):   # budget 3
        torch_ops.set_shape(A, 8, 4, 4, 8, rows_per_group=torch_ops.TILED, wg_per_cu=3)
    with pytest.raises(Exception):   # transpose
        torch_ops.set_shape(A, 8, 4, 4, 8, rows_per_group=torch_ops.TILED, wg_per_cu=1,
                            transpose=True)
    # a row of 1000 nonzeros fits 1920 slots (4 lanes, budget 1) but not 480
    # (8 lanes, budget 2)
    m = 1200
    ipl = np.array([0, 1000] + [1000 + i for i in range(1, m)], np.int64)
    ixl = np.concatenate([np.arange(1000), np.arange(1, m) % m]).astype(np.int64)
    L = CsrMatrix.from_arrays(ipl, ixl, np.ones(len(ixl), np.float32), (m, m))
    torch_ops.set_shape(L, 8, 4, 4, 8, rows_per_group=torch_ops.TILED, wg_per_cu=1)
    with pytest.raises(Exception):
        torch_ops.set_shape(L, 8, 4, 8, 8, rows_per_group=torch_ops.TILED, wg_per_cu=2)


@pytest.mark.parametrize("b", [1, 3, 6])
def test_xcd_blocks_are_bit_identical(gpu, b):
    """Row-split and row-grouped shapes with XCD row blocks (each XCD runs
    its row range in blocks, every slab of a block before the next) give the
    restatement's result bit for bit, with blocks larger than an XCD's range
    included."""
    import torch
    g = synth.knn_graph(20_000, 12, seed=b)
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n))
    for F in (8, 520):
        X = synth.normal(g.n, F, seed=F + b)
        Xd = torch.from_numpy(X).cuda()
        ref = orc.spmm_step_f32(g.indptr, g.indices, g.data, X)
        for kw in (dict(rows_per_group=4), dict(rows_per_group=8), dict(), dict(group=16)):
            group = kw.pop("group", 64)
            torch_ops.set_shape(A, F, 4, group, 8, xcd_block=b, **kw)
            assert torch_ops.get_shape(A, F)["xcd_block"] == b
            Y = torch_ops.spmm(A, Xd).cpu().numpy()
            assert np.array_equal(Y, ref), (F, b, group, kw)
