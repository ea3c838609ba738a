/*
 * scl_pseudotime.h -- diffusion-component pseudotime of scl-core on the GPU.
 *
 * Replaces scl/binding/c_api/pseudotime.h:41-47 (implemented in
 * scl/binding/c_api/pseudotime.cpp:80-118 over
 * scl/kernel/pseudotime.hpp:646-712).  Same signature, argument checks and
 * error codes.
 *
 * Checks the reference lacks, added here:
 *   - the adjacency must be square, n must cover its rows and root_cell must
 *     be one of them                               -> SCL_ERROR_DIMENSION_MISMATCH
 *                                                     (root: SCL_ERROR_INVALID_ARGUMENT)
 *   - n_dcs must not be negative                   -> SCL_ERROR_INVALID_ARGUMENT
 *   - device allocation failure / no usable GPU    -> 3 / 41
 */
#ifndef SCL_PSEUDOTIME_H
#define SCL_PSEUDOTIME_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pseudotime.h:41-47 / pseudotime.cpp:80-118: n_dcs diffusion components by
 * 100 steps of orthogonal iteration Q <- orthonormalise(A Q) from the
 * FastRNG(42) start (pseudotime.hpp:646-689), then each cell's distance to
 * root_cell in component space, divided by the largest distance.  The
 * SpMM runs on the diffusion kernels; the Gram-Schmidt sums spread over the
 * GPU, so results agree with a sequential CPU run to rounding. */
SCL_EXPORT scl_error_t scl_pseudotime_diffusion(scl_sparse_t adjacency,
                                                scl_index_t root_cell,
                                                scl_real_t* pseudotime, scl_index_t n,
                                                scl_index_t n_dcs);

#ifdef __cplusplus
}
#endif

#endif /* SCL_PSEUDOTIME_H */
