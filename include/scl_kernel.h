/*
 * scl_kernel.h -- kernel weights of kNN distances on the GPU.
 *
 * Replaces scl/binding/c_api/kernel.h:18-29 (kernel kinds), 66-83
 * (implemented in scl/binding/c_api/kernel.cpp:145-197 over
 * scl/kernel/kernel.hpp:530-645).  Same signatures, kinds, argument checks
 * and error codes.  The kernels evaluate exp / cos on the GPU, which may
 * differ from the host's libm (and from the reference's vectorised Exp) in
 * the last bit; the algebraic kernels are exact.
 *
 * Checks the reference lacks, added here: kernel_row_sums needs a sum slot
 * per primary slot of the handle -> SCL_ERROR_DIMENSION_MISMATCH; device
 * allocation failure -> 3; no usable GPU -> 41.
 */
#ifndef SCL_KERNEL_H
#define SCL_KERNEL_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* kernel.h:18-29; linear, polynomial and sigmoid evaluate the Gaussian, as
 * in the reference (kernel.hpp:125-144). */
typedef enum {
    SCL_KERNEL_GAUSSIAN = 0,
    SCL_KERNEL_EPANECHNIKOV = 1,
    SCL_KERNEL_COSINE = 2,
    SCL_KERNEL_LINEAR = 3,
    SCL_KERNEL_POLYNOMIAL = 4,
    SCL_KERNEL_LAPLACIAN = 5,
    SCL_KERNEL_CAUCHY = 6,
    SCL_KERNEL_SIGMOID = 7,
    SCL_KERNEL_UNIFORM = 8,
    SCL_KERNEL_TRIANGULAR = 9
} scl_kernel_type_t;

/* kernel.h:66-71 / kernel.cpp:145-163 (compute_kernel_matrix,
 * kernel.hpp:530-590): kernel_values[k] = K(d_k^2) for every stored distance
 * d_k of the handle, in stored order (nnz values); the bandwidth is clamped
 * below at 1e-10. */
SCL_EXPORT scl_error_t scl_kernel_compute_matrix(scl_sparse_t distances,
                                                 scl_real_t* kernel_values,
                                                 scl_real_t bandwidth,
                                                 scl_kernel_type_t kernel_type);

/* kernel.h:77-83 / kernel.cpp:169-197 (kernel_row_sums,
 * kernel.hpp:596-645): per primary slot, the kernel values summed in stored
 * order.  n_points must equal the handle's rows. */
SCL_EXPORT scl_error_t scl_kernel_row_sums(scl_sparse_t distances, scl_real_t* sums,
                                           scl_size_t n_points, scl_real_t bandwidth,
                                           scl_kernel_type_t kernel_type);

#ifdef __cplusplus
}
#endif

#endif /* SCL_KERNEL_H */
