/*
 * scl_neighbors.h -- exact k-nearest neighbours of the samples of a sparse
 * matrix, on the GPU.
 *
 * Replaces scl/binding/c_api/neighbors.h:19-45 (implemented in
 * scl/binding/c_api/neighbors.cpp:20-124 over
 * scl/kernel/neighbors.hpp:300-476).  Same signatures and error codes.
 * Samples are the handle's primary slots (rows of a CSR handle), features
 * its secondary dimension; column indices are taken as sorted and unique
 * within a sample, as the reference's merge-based dot products require.
 *
 * Distances are the reference's: squared norms in vectorize::sum_squared's
 * order, dot products as the reference's sequential FMA chain over matching
 * features, d^2 = n_i + n_j - 2 dot clamped at 0, then sqrt.  Each row's
 * neighbours come sorted by (distance, index); the reference's heap orders
 * equal distances its own way and, among candidates at exactly the k-th
 * distance, keeps the first it met.  Its Cauchy-Schwarz pruning is exact
 * except where rounding puts a computed distance below its bound for
 * near-duplicate samples; the search here has no pruning.
 *
 * Checks the reference lacks, added here:
 *   - n_samples must cover the matrix's samples     -> SCL_ERROR_DIMENSION_MISMATCH
 *   - scl_knn_simple with k == 0                     -> SCL_ERROR_INVALID_ARGUMENT
 *   - the dense samples x features matrix must fit   -> SCL_ERROR_OUT_OF_MEMORY
 *   - no usable GPU                                  -> SCL_ERROR_FEATURE_UNAVAILABLE
 */
#ifndef SCL_NEIGHBORS_H
#define SCL_NEIGHBORS_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* neighbors.h:19-22 / neighbors.cpp:20-38: squared norm of every sample
 * (the output has the handle's rows() entries). */
SCL_EXPORT scl_error_t scl_neighbors_compute_norms(scl_sparse_t matrix, scl_real_t* norms_sq);

/* neighbors.h:25-32 / neighbors.cpp:44-75: k nearest other samples of every
 * sample with the caller's squared norms; out_* are n_samples x k row-major,
 * missing neighbours are -1 / +inf. */
SCL_EXPORT scl_error_t scl_knn(scl_sparse_t matrix, const scl_real_t* norms_sq,
                               scl_size_t n_samples, scl_size_t k, scl_index_t* out_indices,
                               scl_real_t* out_distances);

/* neighbors.h:35-41 / neighbors.cpp:81-124: the same with the norms
 * computed internally. */
SCL_EXPORT scl_error_t scl_knn_simple(scl_sparse_t matrix, scl_size_t n_samples, scl_size_t k,
                                      scl_index_t* out_indices, scl_real_t* out_distances);

#ifdef __cplusplus
}
#endif

#endif /* SCL_NEIGHBORS_H */
