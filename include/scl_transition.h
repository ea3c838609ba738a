/*
 * scl_transition.h -- transition-matrix construction and the stationary
 * distribution on the GPU.
 *
 * Replaces scl/binding/c_api/transition.h:18-55 (implemented in
 * scl/binding/c_api/transition.cpp:29-150 over
 * scl/kernel/transition.hpp:449-623).  Same signatures and semantics; the
 * results follow the reference's FP32 arithmetic operation for operation as
 * built for AVX/AVX2 x86 (vector_sum's eight strided partials,
 * transition.hpp:165-179), so they are bit-identical to a CPU restatement of
 * it.
 *
 * Checks the reference lacks, added here (it reads and writes past the
 * buffers instead):
 *   - the matrix must be n x n (row_normalize: n at most its rows, and
 *     output_size at least the entries of those rows) -> SCL_ERROR_DIMENSION_MISMATCH
 *   - n must not be negative                           -> SCL_ERROR_INVALID_ARGUMENT
 *   - device allocation failure / no usable GPU        -> 3 / 41
 */
#ifndef SCL_TRANSITION_H
#define SCL_TRANSITION_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* transition.h:18-22 / transition.cpp:29-55 (transition.hpp:449-487): dense
 * row-major n x n; row i keeps its positive entries divided by their sum,
 * or is e_i when that sum is at most 1e-15. */
SCL_EXPORT scl_error_t scl_transition_matrix_from_velocity(scl_sparse_t velocity_graph,
                                                           scl_index_t n,
                                                           scl_real_t* row_stochastic_out);

/* transition.h:28-33 / transition.cpp:57-87 (transition.hpp:493-525): the
 * values of the first n rows divided by their row sums (stored-order FP32
 * sums; rows summing to at most 1e-15 become 0). */
SCL_EXPORT scl_error_t scl_transition_row_normalize_to_stochastic(scl_sparse_t input,
                                                                  scl_index_t n,
                                                                  scl_real_t* output_values,
                                                                  scl_size_t output_size);

/* transition.h:39-43 / transition.cpp:89-115 (transition.hpp:530-565):
 * dense row-major (T + T^T)/2 with every row of positive sum normalised. */
SCL_EXPORT scl_error_t scl_transition_symmetrize(scl_sparse_t transition_mat, scl_index_t n,
                                                 scl_real_t* symmetric_out);

/* transition.h:49-55 / transition.cpp:117-150 (transition.hpp:571-623):
 * power iteration pi <- T^T pi / |T^T pi|_1 from the uniform vector, at most
 * max_iter times, stopping when max |change| < tolerance or on the
 * reference's Aitken test. */
SCL_EXPORT scl_error_t scl_transition_stationary_distribution(scl_sparse_t transition_mat,
                                                              scl_index_t n,
                                                              scl_real_t* stationary,
                                                              scl_index_t max_iter,
                                                              scl_real_t tolerance);

#ifdef __cplusplus
}
#endif

#endif /* SCL_TRANSITION_H */
