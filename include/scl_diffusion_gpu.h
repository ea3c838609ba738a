/*
 * scl_diffusion_gpu.h -- device-pointer extensions of the diffusion ABI.
 *
 * These entry points have no counterpart in the reference (which is CPU
 * only).  They exist so that callers whose data already lives in GPU memory
 * (a torch.autograd.Function, a benchmark, a pipeline stage) can run the same
 * kernels as scl_diffusion_diffuse_matrix without the host round trip the
 * drop-in's host-pointer signature forces (scl/binding/c_api/diffusion.h:29-35).
 *
 * Conventions:
 *   - every pointer named *_dev is device memory on the current HIP device;
 *   - `stream` is a hipStream_t (NULL = the legacy default stream); the call
 *     enqueues work on it and returns without waiting;
 *   - dense matrices are row-major, n_nodes x n_features, fp32, contiguous;
 *   - the operator is the handle's primary-slot CSR (a CSC handle applies the
 *     transpose), exactly as in scl_diffusion.h;
 *   - errors are reported through the same codes and thread-local message.
 */
#ifndef SCL_DIFFUSION_GPU_H
#define SCL_DIFFUSION_GPU_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of usable HIP devices (0 when none; never an error). */
SCL_EXPORT scl_error_t scl_gpu_device_count(int* out);

/* Drops the handle's device copies; the next call re-uploads them.  Needed
 * after the caller edits the arrays of a handle made by scl_sparse_wrap. */
SCL_EXPORT scl_error_t scl_gpu_sparse_invalidate(scl_sparse_t matrix);

/* Uploads the handle's arrays to the current device now (otherwise done on
 * first use), so that later stream-ordered calls do no host work. */
SCL_EXPORT scl_error_t scl_gpu_sparse_prepare(scl_sparse_t matrix);

/* Y_dev = P^t X_dev.  X_dev is not modified; Y_dev must not overlap X_dev.
 * work_dev (n_nodes x n_features floats, overlapping neither) is required
 * when n_steps >= 2 and ignored otherwise.  n_steps >= 1. */
SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_real_t* work_dev, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t n_steps, void* stream);

/* One step Y_dev = op(P) X_dev with op(P) = P (transpose == 0) or P^T
 * (transpose != 0, applied through a cached transposed pattern). */
SCL_EXPORT scl_error_t scl_gpu_spmm_device(scl_sparse_t transition,
                                           const scl_real_t* X_dev,
                                           scl_real_t* Y_dev,
                                           scl_index_t n_nodes,
                                           scl_index_t n_features,
                                           int transpose, void* stream);

/* Like scl_diffusion_compute_transition_matrix (the values of `adjacency`
 * are normalised in place, FP64 on the GPU), and additionally returns the
 * FP64 degrees d_i = sum_j a_ij (degrees_out: host array of rows doubles, or
 * NULL) and the FP64 stationarity residual r = || P^T pi - pi ||_1 with
 * pi = d / sum(d) evaluated with the normalised fp32 values (residual_out:
 * host double, or NULL).  r vanishes, up to rounding, for symmetric A with
 * symmetric == 0. */
SCL_EXPORT scl_error_t scl_gpu_compute_transition_matrix_ex(
    scl_sparse_t adjacency, int symmetric, double* degrees_out,
    double* residual_out);

#ifdef __cplusplus
}
#endif
#endif
