// callers.hip -- gfx950 kernels of the other callers of the diffusion SpMM:
// the dense start of the imputations, MAGIC's row normalisation, the
// pseudotime distances and the general Y = alpha A X + beta Y.  The t-step
// SpMMs themselves run on the row-split kernels of kernels.hip.
// Contraction is off: every fused multiply-add is an explicit fmaf, matching
// the CPU restatement operation for operation.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace scl_amd {

using namespace dev;

namespace {

constexpr int kParts = 1024;

// flag = 1 when some primary slot of (ip, ix) is not strictly increasing;
// one wave per slot, each lane compares neighbouring entries
__global__ __launch_bounds__(kBlock) void k_unsorted(const int64_t* __restrict__ ip,
                                                     const int64_t* __restrict__ ix,
                                                     int64_t slots, int* __restrict__ flag) {
    const int64_t slot = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (slot >= slots) return;
    const int lane = threadIdx.x & 63;
    const int64_t end = ip[slot + 1];
    for (int64_t k = ip[slot] + lane; k + 1 < end; k += 64)
        if (ix[k + 1] <= ix[k]) *flag = 1;
}

// Dense n_cells x G start of the imputations: stored entry k of primary slot
// `slot` goes to (slot, ix[k]) for cells-major input (CSR), to (ix[k], slot)
// for genes-major input (CSC), when in range.  One wave per slot; with
// `ordered` lane 0 alone walks the slot in stored order so that a repeated
// index keeps its last value.
__global__ __launch_bounds__(kBlock) void k_densify(const int64_t* __restrict__ ip,
                                                    const int64_t* __restrict__ ix,
                                                    const float* __restrict__ v, int64_t slots,
                                                    int64_t n_cells, int64_t G, int is_csr,
                                                    int ordered, float* __restrict__ out) {
    const int64_t slot = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (slot >= slots) return;
    const int lane = threadIdx.x & 63;
    if (ordered && lane != 0) return;
    const int64_t lim = is_csr ? G : n_cells;
    const int step = ordered ? 1 : 64;
    const int64_t end = ip[slot + 1];
    for (int64_t k = ip[slot] + (ordered ? 0 : lane); k < end; k += step) {
        const int64_t j = ix[k];
        if (j < 0 || j >= lim) continue;
        out[is_csr ? slot * G + j : j * G + slot] = v[k];
    }
}

// MAGIC (impute.hpp:700-741): the FP32 row sum of the affinity in stored order
// (plain adds), then w_k = a_k * (1 / sum) for every entry of the row; a row
// whose sum is below 1e-12 gets zero weights and is listed in `copy` (its
// output row is its input row).  One thread per row.
__global__ __launch_bounds__(kBlock) void k_magic_weights(const int32_t* __restrict__ row_ptr,
                                                          const float* __restrict__ val,
                                                          int32_t n, float* __restrict__ w,
                                                          int32_t* __restrict__ copy,
                                                          int* __restrict__ n_copy) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t beg = row_ptr[i], end = row_ptr[i + 1];
    float sum = 0.0f;
    for (int32_t k = beg; k < end; ++k) sum += val[k];
    if (sum < 1e-12f) {
        for (int32_t k = beg; k < end; ++k) w[k] = 0.0f;
        copy[atomicAdd(n_copy, 1)] = i;
        return;
    }
    const float inv = 1.0f / sum;
    for (int32_t k = beg; k < end; ++k) w[k] = val[k] * inv;
}

// out[r, :] = in[r, :] for the m listed rows
__global__ __launch_bounds__(kBlock) void k_copy_rows(const int32_t* __restrict__ rows,
                                                      int32_t m, const float* __restrict__ in,
                                                      float* __restrict__ out, int64_t G) {
    const int64_t total = (int64_t)m * G;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * kBlock) {
        const int64_t r = rows[e / G], f = e % G;
        out[r * G + f] = in[r * G + f];
    }
}

__device__ __forceinline__ float block_max(float v) {
    __shared__ float red[kBlock / 64];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = red[0];
    for (int w = 1; w < kBlock / 64; ++w) r = fmaxf(r, red[w]);
    __syncthreads();
    return r;
}

// pt_i = sqrt(sum_c (Q[i,c] - Q[root,c])^2) over the row-major n x K
// components (pseudotime.hpp:691-701), and per-block maxima
__global__ __launch_bounds__(kBlock) void k_pt_distance(const float* __restrict__ Q, int64_t n,
                                                        int64_t K, int64_t root,
                                                        float* __restrict__ pt,
                                                        float* __restrict__ part) {
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        float d2 = 0.0f;
        for (int64_t c = 0; c < K; ++c) {
            const float d = Q[i * K + c] - Q[root * K + c];
            d2 = __builtin_fmaf(d, d, d2);
        }
        const float p = sqrtf(d2);
        pt[i] = p;
        m = fmaxf(m, p);
    }
    m = block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// pt *= 1 / max pt when max pt > 0 (pseudotime.hpp:703-711); every block
// folds the m partial maxima itself
__global__ __launch_bounds__(kBlock) void k_pt_scale(float* __restrict__ pt, int64_t n,
                                                     const float* __restrict__ part, int m) {
    float mx = 0.0f;
    for (int i = threadIdx.x; i < m; i += kBlock) mx = fmaxf(mx, part[i]);
    mx = block_max(mx);
    if (!(mx > 0.0f)) return;
    const float inv = 1.0f / mx;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        pt[i] = pt[i] * inv;
}

// Y = alpha A X + beta Y (algebra.hpp:546-609) for the m x k operator A
// (row-major X: k x F, Y: m x F).  One wave per row; VEC consecutive features
// per lane, slabs of 64 * VEC features.  Each element starts from beta * Y
// (0 when beta == 0, Y itself when beta == 1) and takes one FMA per stored
// entry in order with weight a_ij * alpha, skipping zero weights; alpha == 0
// leaves the scaled Y.
template <int VEC>
__global__ __launch_bounds__(kBlock) void k_spmm_axpby(const int32_t* __restrict__ row_ptr,
                                                       const int32_t* __restrict__ col,
                                                       const float* __restrict__ val,
                                                       const int32_t* __restrict__ sched,
                                                       const float* __restrict__ X,
                                                       float* __restrict__ Y, int32_t m,
                                                       int64_t F, float alpha, float beta) {
    using VT = typename Vec<VEC>::T;
    const int64_t wave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (wave >= m) return;
    const int lane = threadIdx.x & 63;
    int row = sched ? sched[wave] : (int)wave;
    row = __builtin_amdgcn_readfirstlane(row);
    const int32_t beg = row_ptr[row], end = row_ptr[row + 1];
    for (int64_t f0 = 0; f0 < F; f0 += 64 * VEC) {
        const int64_t f = f0 + (int64_t)lane * VEC;
        if (f >= F) break;
        float* y = Y + (int64_t)row * F + f;
        float acc[VEC];
        if (beta == 0.0f) {
#pragma unroll
            for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
        } else {
            const VT y0 = *reinterpret_cast<const VT*>(y);
            const float* yp = reinterpret_cast<const float*>(&y0);
#pragma unroll
            for (int c = 0; c < VEC; ++c) acc[c] = beta == 1.0f ? yp[c] : yp[c] * beta;
        }
        if (alpha != 0.0f) {
            for (int32_t k = beg; k < end; ++k) {
                const float w = val[k] * alpha;
                if (w == 0.0f) continue;
                const VT x = *reinterpret_cast<const VT*>(X + (int64_t)col[k] * F + f);
                fma_vec(w, x, acc);
            }
        }
        store_vec<VEC, false>(y, acc);
    }
}

unsigned waves_grid(int64_t waves) {
    return (unsigned)((waves * 64 + kBlock - 1) / kBlock);
}

}  // namespace

bool launch_find_unsorted(const int64_t* ip, const int64_t* ix, int64_t slots, int* flag,
                          hipStream_t s) {
    if (slots <= 0) return false;
    hipLaunchKernelGGL(k_unsorted, dim3(waves_grid(slots)), dim3(kBlock), 0, s, ip, ix, slots,
                       flag);
    check_launch("order check launch");
    return true;
}

void launch_densify(const int64_t* ip, const int64_t* ix, const float* v, int64_t slots,
                    int64_t n_cells, int64_t G, bool is_csr, bool ordered, float* out,
                    hipStream_t s) {
    if (slots <= 0) return;
    hipLaunchKernelGGL(k_densify, dim3(waves_grid(slots)), dim3(kBlock), 0, s, ip, ix, v, slots,
                       n_cells, G, is_csr ? 1 : 0, ordered ? 1 : 0, out);
    check_launch("densify launch");
}

void launch_magic_weights(const int32_t* row_ptr, const float* val, int32_t n, float* w,
                          int32_t* copy, int* n_copy, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_magic_weights, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       row_ptr, val, n, w, copy, n_copy);
    check_launch("magic weights launch");
}

void launch_copy_rows(const int32_t* rows, int32_t m, const float* in, float* out, int64_t G,
                      hipStream_t s) {
    if (m <= 0 || G <= 0) return;
    const int64_t total = (int64_t)m * G;
    const int64_t blocks = std::min<int64_t>((total + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)blocks), dim3(kBlock), 0, s, rows, m, in,
                       out, G);
    check_launch("row copy launch");
}

void launch_pt_distance(const float* Q, int64_t n, int64_t K, int64_t root, float* pt,
                        float* part, hipStream_t s) {
    if (n <= 0) return;
    const int64_t b = std::min<int64_t>((n + kBlock - 1) / kBlock, kParts);
    hipLaunchKernelGGL(k_pt_distance, dim3((unsigned)b), dim3(kBlock), 0, s, Q, n, K, root, pt,
                       part);
    hipLaunchKernelGGL(k_pt_scale, dim3((unsigned)b), dim3(kBlock), 0, s, pt, n, part, (int)b);
    check_launch("pseudotime distance launch");
}

void launch_spmm_axpby(const CsrView& A, const float* X, float* Y, int64_t F, float alpha,
                       float beta, hipStream_t s) {
    if (A.n <= 0 || F <= 0) return;
    const bool v4 = F % 4 == 0 && ((uintptr_t)X % 16) == 0 && ((uintptr_t)Y % 16) == 0;
    const unsigned g = waves_grid(A.n);
    if (v4)
        hipLaunchKernelGGL(k_spmm_axpby<4>, dim3(g), dim3(kBlock), 0, s, A.row_ptr, A.col,
                           A.val, A.sched, X, Y, A.n, F, alpha, beta);
    else
        hipLaunchKernelGGL(k_spmm_axpby<1>, dim3(g), dim3(kBlock), 0, s, A.row_ptr, A.col,
                           A.val, A.sched, X, Y, A.n, F, alpha, beta);
    check_launch("spmm axpby launch");
}

}  // namespace scl_amd
