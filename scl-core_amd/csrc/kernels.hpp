// kernels.hpp -- host-callable launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace scl_amd {

// Operator view handed to the launchers: CSR with int32 indices.
struct CsrView {
    int32_t n;            // rows of the applied operator
    int32_t n_sec;        // columns
    int64_t nnz;
    const int32_t* row_ptr;
    const int32_t* col;
    const float* val;
    const int32_t* sched;  // row visiting order or null
};

// Launch shape of the row-split SpMM (see kernels.hip).
struct SpmmShape {
    int vec = 4;        // floats per lane per access (1, 2 or 4)
    int group = 64;     // lanes sharing one row (power of two, 1..64)
    int unroll = 8;     // nonzeros in flight per lane group
    bool nontemporal = false;
};

// Picks vec/group for a feature count and the alignment of X and Y.
SpmmShape spmm_default_shape(int64_t F, const void* X, const void* Y);

// One step Y = P X (Y[i,:] = sum over row i, stored order, FMA in fp32).
void launch_spmm_f32(const CsrView& P, const float* X, float* Y, int64_t F,
                     const SpmmShape& shape, hipStream_t s);

// FP64 degree normalisation of values in place: deg[i] = sum_k a_ik (FP64,
// stored order); symmetric == 0: a_ik <- (float)(a_ik / deg[i]) (0 if
// deg <= 1e-15); symmetric != 0: a_ik <- (float)(a_ik * r_i * r_j),
// r = 1/sqrt(deg) (0 if deg <= 1e-15).  deg (n doubles) is written.
void launch_row_sums_f64(const int32_t* row_ptr, const float* val, int32_t n,
                         double* deg, hipStream_t s);
void launch_normalize_rows(const int32_t* row_ptr, const int32_t* col,
                           float* val, int32_t n, const double* deg,
                           int symmetric, hipStream_t s);

// FP64 stationarity residual r = sum_j | sum_{i->j} p_ij pi_i - pi_j |,
// pi = deg / sum(deg), evaluated through the transposed pattern (t_row_ptr,
// t_col = source row, t_val).  partial must hold 2 * 1024 doubles; the
// result is written to out (one double, device memory).
void launch_stationarity_residual(const int32_t* t_row_ptr, const int32_t* t_col,
                                  const float* t_val, int32_t n,
                                  const double* deg, double* partial,
                                  double* out, hipStream_t s);

}  // namespace scl_amd
