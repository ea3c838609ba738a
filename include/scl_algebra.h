/*
 * scl_algebra.h -- the sparse x dense product of scl-core's algebra module
 * on the GPU.
 *
 * Replaces scl/binding/c_api/algebra.h:133-140 (implemented in
 * scl/binding/c_api/algebra.cpp:164-195 over scl/kernel/algebra.hpp:546-609).
 * Same signature, argument checks and error codes.
 *
 * Checks the reference lacks, added here:
 *   - nnz and the primary dimension must fit int32 on the device -> SCL_ERROR_INVALID_ARGUMENT
 *   - device allocation failure / no usable GPU                  -> 3 / 41
 */
#ifndef SCL_ALGEBRA_H
#define SCL_ALGEBRA_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* algebra.h:133-140 / algebra.cpp:164-195: Y = alpha A X + beta Y with A
 * primary_dim x secondary_dim (a CSC handle: its columns are the rows), X
 * secondary_dim x n_cols and Y primary_dim x n_cols, row-major host buffers.
 * beta == 0 clears Y (NaN included), beta == 1 keeps it; alpha == 0 stops
 * there; otherwise every element takes one FMA per stored entry of its row
 * in order with weight a_ij * alpha, and entries whose weight is zero are
 * skipped (algebra.hpp:110-111). */
SCL_EXPORT scl_error_t scl_algebra_spmm(scl_sparse_t A, const scl_real_t* X,
                                        scl_index_t n_cols, scl_real_t* Y,
                                        scl_real_t alpha, scl_real_t beta);

#ifdef __cplusplus
}
#endif

#endif /* SCL_ALGEBRA_H */
