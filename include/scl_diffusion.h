/*
 * scl_diffusion.h -- the drop-in diffusion entry points.
 *
 * Replaces scl/binding/c_api/diffusion.h:15-35 and 99-105 (implemented in
 * scl/binding/c_api/diffusion.cpp:36-116 and 318-341).  Same signatures,
 * same argument checks and error codes, same in-place host-buffer semantics;
 * the work runs on the current HIP device and the call returns when the
 * result is back in the caller's buffer.
 *
 * Checks the reference lacks, added here (all forced from the host, none
 * touches the GPU):
 *   - n_nodes must equal the operator's rows and cols -> SCL_ERROR_DIMENSION_MISMATCH
 *   - n_nodes * n_features must fit in memory         -> SCL_ERROR_INVALID_ARGUMENT
 *   - nnz and n must fit int32 on the device          -> SCL_ERROR_INVALID_ARGUMENT
 *   - device allocation failure                       -> SCL_ERROR_OUT_OF_MEMORY
 *   - no usable GPU                                   -> SCL_ERROR_FEATURE_UNAVAILABLE
 *
 * A CSC handle is applied through its primary (column) slots, exactly as the
 * reference's visit() dispatch does (core/internal.hpp:133-151), so
 * diffusion through a CSC handle of P computes (P^T)^t X.
 */
#ifndef SCL_DIFFUSION_H
#define SCL_DIFFUSION_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* diffusion.h:15-18 / diffusion.cpp:36-62.  Normalises the handle's values in
 * place.  symmetric == 0: row-stochastic D^-1 A; symmetric != 0:
 * D^-1/2 A D^-1/2.  Row sums, reciprocals and square roots are FP64 on the
 * GPU and each value is rounded to float once; rows whose sum is <= 1e-15
 * become zero (diffusion.hpp:41, 565-600). */
SCL_EXPORT scl_error_t scl_diffusion_compute_transition_matrix(
    scl_sparse_t adjacency, int symmetric);

/* diffusion.h:21-26 / diffusion.cpp:68-88: x <- P^t x (x on the host). */
SCL_EXPORT scl_error_t scl_diffusion_diffuse_vector(scl_sparse_t transition,
                                                    scl_real_t* x,
                                                    scl_size_t n_nodes,
                                                    scl_index_t n_steps);

/* diffusion.h:29-35 / diffusion.cpp:94-116: X <- P^t X, X an n_nodes x
 * n_features row-major host buffer, overwritten in place. */
SCL_EXPORT scl_error_t scl_diffusion_diffuse_matrix(scl_sparse_t transition,
                                                    scl_real_t* X,
                                                    scl_index_t n_nodes,
                                                    scl_index_t n_features,
                                                    scl_index_t n_steps);

/* diffusion.h:99-105 / diffusion.cpp:318-341: alias of diffuse_matrix with
 * t == 0 meaning 3 (diffusion.cpp:333); t < 0 leaves X unchanged, as the
 * reference's unchecked loop does. */
SCL_EXPORT scl_error_t scl_diffusion_magic_impute(scl_sparse_t transition,
                                                  scl_real_t* X,
                                                  scl_index_t n_nodes,
                                                  scl_index_t n_features,
                                                  scl_index_t t);

#ifdef __cplusplus
}
#endif
#endif
