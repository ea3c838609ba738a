"""Writes the small diffusion fixtures in tests/golden/*.npz.

Every fixture holds an operator (indptr, indices, data: the stored CSR of P),
an input X, a step count t and expected outputs:
  expected_f64  from a closed form or a dense FP64 matrix power computed here
                with numpy -- never from the oracle;
  expected_f32  (where present) the FP32 restatement's output, which the GPU
                path must reproduce bit for bit.

Known answers taken from the reference's own tests:
  algebra_3x3   tests/c_api/src/test_algebra.cpp:281-307 (spmm_basic) and
                309-325 (A x = [7, 6, 32] inside spmm_with_alpha_beta);
  impute_3x3    tests/c_api/src/test_impute.cpp:267-295 (the reference only
                asserts finiteness; the exact P^2 X is written out below).
Graph builders of tests/c_api/src/test_pseudotime.cpp:23-47 (path, ring)
give the path and ring fixtures.

Run:  python tests/golden/make_golden.py   (needs `make -C oracle` and
`make -C scl-core_amd` for the random fixture's generator and FP32 output).
"""
from __future__ import annotations

import sys
from math import comb
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "scl-core_amd"))


def csr(A):
    indptr, indices, data = [0], [], []
    for row in A:
        nz = np.nonzero(row)[0]
        indices += nz.tolist()
        data += row[nz].tolist()
        indptr.append(len(indices))
    return (np.array(indptr, np.int64), np.array(indices, np.int64),
            np.array(data, np.float32))


def save(name, A=None, *, indptr=None, indices=None, data=None, X, t,
         expected_f64, expected_f32=None, **extra):
    if A is not None:
        indptr, indices, data = csr(A)
    arrs = dict(indptr=indptr, indices=indices, data=data,
                X=np.asarray(X, np.float32), t=np.int64(t),
                expected_f64=np.asarray(expected_f64, np.float64), **extra)
    if expected_f32 is not None:
        arrs["expected_f32"] = np.asarray(expected_f32, np.float32)
    np.savez(HERE / f"{name}.npz", **arrs)
    print("wrote", name)


def main():
    # --- reference known answers (test_algebra.cpp:281-325) ---------------
    A = np.array([[1, 0, 2], [0, 3, 0], [4, 5, 6]], np.float64)
    X = np.array([[1, 2], [3, 4], [5, 6]], np.float64)
    save("algebra_3x3", A, X=X, t=1,
         expected_f64=[[11, 14], [9, 12], [49, 64]],
         expected_f32=[[11, 14], [9, 12], [49, 64]],
         x=np.array([1, 2, 3], np.float32),
         Ax=np.array([7, 6, 32], np.float32))

    # --- test_impute.cpp:267-295: P (all 0.5), X = the test's 3x3 input --
    Pm = np.array([[.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    Xi = np.array([[1, 2, 0], [3, 0, 0], [0, 4, 5]], np.float64)
    # P^2 = [[.5,.25,.25],[.25,.5,.25],[.25,.25,.5]] (hand computed), so
    # P^2 X is exact in binary floating point:
    hand = np.array([[1.25, 2.0, 1.25], [1.75, 1.5, 1.25], [1.0, 2.5, 2.5]])
    assert np.array_equal(np.linalg.matrix_power(Pm, 2) @ Xi, hand)
    save("impute_3x3", Pm, X=Xi, t=2, expected_f64=hand, expected_f32=hand)

    rng = np.random.Generator(np.random.PCG64(20260101))

    # --- ring C_n, weight 1/2 per side: binomial closed form ---------------
    n, t, F = 16, 3, 4
    R = np.zeros((n, n))
    for i in range(n):
        R[i, (i - 1) % n] = R[i, (i + 1) % n] = 0.5
    Xr = rng.standard_normal((n, F)).astype(np.float32).astype(np.float64)
    ring = np.zeros((n, F))
    for i in range(n):
        for m in range(t + 1):
            ring[i] += comb(t, m) * 2.0 ** -t * Xr[(i + 2 * m - t) % n]
    save("ring16_t3", R, X=Xr, t=t, expected_f64=ring)

    # --- complete graph K_n, uniform weights, no self loops ---------------
    # n = 9: the weight 1/(n-1) = 1/8 is exact in fp32
    n, t, F = 9, 2, 3
    K = (np.ones((n, n)) - np.eye(n)) / (n - 1)
    Xk = rng.standard_normal((n, F)).astype(np.float32).astype(np.float64)
    mu = Xk.mean(axis=0, keepdims=True)
    kn = mu + (-1.0 / (n - 1)) ** t * (Xk - mu)
    save("complete9_t2", K, X=Xk, t=t, expected_f64=kn)

    # --- path graph (row-normalised chain), dense FP64 power ---------------
    n, t, F = 12, 4, 2
    W = np.zeros((n, n))
    for i in range(n - 1):
        W[i, i + 1] = W[i + 1, i] = 1.0
    Pp = W / W.sum(axis=1, keepdims=True)
    Pp32 = Pp.astype(np.float32).astype(np.float64)
    Xp = rng.standard_normal((n, F)).astype(np.float32).astype(np.float64)
    save("path12_t4", Pp32, X=Xp, t=t,
         expected_f64=np.linalg.matrix_power(Pp32, t) @ Xp)

    # --- empty rows and a long row -------------------------------------------
    n, t, F = 40, 2, 5
    E = np.zeros((n, n))
    for i in range(0, n, 3):
        E[i, rng.choice(n, size=rng.integers(1, 6), replace=False)] = rng.random()
    E[7, :] = rng.random(n)
    E32 = E.astype(np.float32).astype(np.float64)
    Xe = rng.standard_normal((n, F)).astype(np.float32).astype(np.float64)
    save("ragged40_t2", E32, X=Xe, t=t,
         expected_f64=np.linalg.matrix_power(E32, t) @ Xe)

    # --- seeded uniform-kNN operator (generator + FP32 restatement) --------
    from scl_amd import synth
    from oracle import oracle as orc
    g = synth.uniform_knn(64, 5, seed=1).normalized()
    Pd = np.zeros((64, 64))
    for i in range(64):
        for k in range(g.indptr[i], g.indptr[i + 1]):
            Pd[i, g.indices[k]] = float(g.data[k])
    Xu = synth.normal(64, 8, seed=1)
    t = 3
    exp64 = np.linalg.matrix_power(Pd, t) @ Xu.astype(np.float64)
    exp32 = orc.diffuse_matrix_f32(g.indptr, g.indices, g.data, Xu, t)
    save("uniform64_k5_t3", indptr=g.indptr, indices=g.indices, data=g.data,
         X=Xu, t=t, expected_f64=exp64, expected_f32=exp32)


if __name__ == "__main__":
    main()
