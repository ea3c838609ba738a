/*
 * scl_impute.h -- the diffusion imputations of scl-core on the GPU.
 *
 * Replaces scl/binding/c_api/impute.h:41-58 (implemented in
 * scl/binding/c_api/impute.cpp:93-157 over scl/kernel/impute.hpp:599-741).
 * Same signatures, argument checks and error codes.  The expression handle
 * is densified on the device (zero, then every stored entry in range, in
 * stored order), the t steps run on the diffusion SpMM kernels and the
 * dense n_cells x n_genes result is written to the caller's host buffer.
 *
 * Checks the reference lacks, added here:
 *   - the transition / affinity matrix must be square with n_cells rows,
 *     and the expression handle must have n_cells rows (CSR) or n_genes
 *     columns (CSC)                                -> SCL_ERROR_DIMENSION_MISMATCH
 *   - nnz and n must fit int32 on the device       -> SCL_ERROR_INVALID_ARGUMENT
 *   - device allocation failure                    -> SCL_ERROR_OUT_OF_MEMORY
 *   - no usable GPU                                -> SCL_ERROR_FEATURE_UNAVAILABLE
 */
#ifndef SCL_IMPUTE_H
#define SCL_IMPUTE_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* impute.h:41-48 / impute.cpp:93-119 (diffusion_impute_sparse_transition,
 * impute.hpp:599-658): X_imputed = P^n_steps densify(matrix), each step the
 * stored-order FMA chain of every row (spmm_diffusion, impute.hpp:321-347).
 * A CSC transition handle applies its transpose, as the reference does. */
SCL_EXPORT scl_error_t scl_impute_diffusion(scl_sparse_t matrix,
                                            scl_sparse_t transition_matrix,
                                            scl_index_t n_cells, scl_index_t n_genes,
                                            scl_index_t n_steps, scl_real_t* X_imputed);

/* impute.h:51-58 / impute.cpp:125-157 (magic_impute, impute.hpp:669-741):
 * t steps of X <- D^-1 A X with D the FP32 row sums of the affinity in stored
 * order; a row summing below 1e-12 keeps its input row (spmm_normalized,
 * impute.hpp:355-389).  Only a CSR expression handle is densified; a CSC one
 * starts from zeros, as in the reference. */
SCL_EXPORT scl_error_t scl_impute_magic(scl_sparse_t matrix, scl_sparse_t affinity_matrix,
                                        scl_index_t n_cells, scl_index_t n_genes,
                                        scl_index_t t, scl_real_t* X_imputed);

#ifdef __cplusplus
}
#endif

#endif /* SCL_IMPUTE_H */
