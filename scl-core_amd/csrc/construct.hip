// construct.hip -- gfx950 kernels that build and inspect the diffusion
// operator: kernel weights of kNN distances (scl/kernel/kernel.hpp:530-645),
// row normalisation, (T + T^T)/2 symmetrisation and the velocity transition
// (scl/kernel/transition.hpp:449-565), and the pieces of the stationary
// distribution's power iteration (transition.hpp:571-623).
// Contraction is off: every fused multiply-add is an explicit fmaf, as the
// reference's compiler contracts them.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace scl_amd {

using namespace dev;

namespace {

constexpr float kEps = 1e-15f;  // transition.hpp:44
constexpr int kParts = 1024;

// kernel.hpp:108-144 and the kernels of 58-104; the kinds without a formula
// (linear, polynomial, sigmoid: 3, 4, 7) evaluate the Gaussian
__device__ __forceinline__ float kernel_value(float dsq, const KernelWeights& p) {
    switch (p.kind) {
        case 1: {  // Epanechnikov
            const float u2 = dsq * p.inv_h2;
            return u2 > 1.0f ? 0.0f : 0.75f * (1.0f - u2);
        }
        case 2: {  // cosine
            const float u = sqrtf(dsq) * p.inv_h;
            const float pi = 3.14159265358979323846f;
            return u > 1.0f ? 0.0f : (pi / 4.0f) * cosf(pi * u / 2.0f);
        }
        case 5:  // Laplacian
            return expf(-sqrtf(dsq) * p.inv_h);
        case 6:  // Cauchy
            return 1.0f / __builtin_fmaf(dsq, p.inv_h2, 1.0f);
        case 8:  // uniform
            return dsq < p.h2 ? 0.5f : 0.0f;
        case 9: {  // triangular
            const float u = sqrtf(dsq) * p.inv_h;
            return u >= 1.0f ? 0.0f : 1.0f - u;
        }
        default:  // Gaussian
            return expf(-dsq * p.inv_2h2);
    }
}

__global__ __launch_bounds__(kBlock) void k_kernel_values(const float* __restrict__ d,
                                                          int64_t nnz, KernelWeights p,
                                                          float* __restrict__ out) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * kBlock)
        out[k] = kernel_value(d[k] * d[k], p);
}

// per row: the kernel values summed in stored order (one thread per row)
__global__ __launch_bounds__(kBlock) void k_kernel_row_sums(const int64_t* __restrict__ ip,
                                                            const float* __restrict__ d,
                                                            int64_t n, KernelWeights p,
                                                            float* __restrict__ sums) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s += kernel_value(d[k] * d[k], p);
    sums[i] = s;
}

// row_normalize_to_stochastic: out = v * (sum > 1e-15 ? 1 / sum : 0), the
// sum in stored order (one thread per row)
__global__ __launch_bounds__(kBlock) void k_row_normalize(const int64_t* __restrict__ ip,
                                                          const float* __restrict__ v,
                                                          int64_t n, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s += v[k];
    const float inv = s > kEps ? 1.0f / s : 0.0f;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k) out[k] = v[k] * inv;
}

// first position of `key` in the strictly increasing slot [beg, end), or -1
__device__ __forceinline__ int64_t find_sorted(const int64_t* __restrict__ ix, int64_t beg,
                                               int64_t end, int64_t key) {
    int64_t lo = beg, hi = end;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (ix[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && ix[lo] == key) ? lo : -1;
}

// (T + T^T)/2 into the zeroed dense n x n S for rows with strictly increasing
// indices.  A pair {a, b}, a < b, receives v_ab/2 (row a) then v_ba/2 (row b)
// in the reference's order, i.e. 0 + h_ab + h_ba, written to both cells by
// the thread of row a, or by row b's thread when row a has no (a, b) entry;
// a diagonal entry adds its half twice.  One thread per row.
__global__ __launch_bounds__(kBlock) void k_symmetrize(const int64_t* __restrict__ ip,
                                                       const int64_t* __restrict__ ix,
                                                       const float* __restrict__ v, int64_t n,
                                                       float* __restrict__ S) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
        const int64_t j = ix[k];
        const float h = v[k] * 0.5f;
        if (j == i) {
            S[i * n + i] = (0.0f + h) + h;
            continue;
        }
        const int64_t m = find_sorted(ix, ip[j], ip[j + 1], i);
        float acc;
        if (j > i) {
            acc = 0.0f + h;
            if (m >= 0) acc = acc + v[m] * 0.5f;
        } else {
            if (m >= 0) continue;  // row j's thread writes the pair
            acc = 0.0f + h;
        }
        S[i * n + j] = acc;
        S[j * n + i] = acc;
    }
}

// the same, entry by entry in stored order on one thread (any index order)
__global__ void k_symmetrize_serial(const int64_t* __restrict__ ip,
                                    const int64_t* __restrict__ ix, const float* __restrict__ v,
                                    int64_t n, float* __restrict__ S) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
            const int64_t j = ix[k];
            const float h = v[k] * 0.5f;
            S[i * n + j] += h;
            S[j * n + i] += h;
        }
}

// vector_sum of transition.hpp:165-179 over x[0, n) on one wave: lane l < 8
// keeps the partial over l, l + 8, l + 16, ... of the full groups of eight
// (chunks of 64 loaded by all lanes, handed over with shuffles), the eight
// partials are added left to right, then the tail in order.  The partials
// are serial chains, so the loads run a batch of kSumBatch chunks ahead of
// the adds (two register buffers).  Every lane returns the result.
constexpr int kSumBatch = 16;

__device__ float wave_sum8(const float* __restrict__ x, int64_t n) {
    const int lane = threadIdx.x & 63;
    const int64_t full = n / 8 * 8;
    float part = 0.0f;
    float cur[kSumBatch], nxt[kSumBatch];
#pragma unroll
    for (int u = 0; u < kSumBatch; ++u) {
        const int64_t i = 64 * u + lane;
        cur[u] = i < full ? x[i] : 0.0f;
    }
    for (int64_t c = 0; c < full; c += 64 * kSumBatch) {
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) {
            const int64_t i = c + 64 * (kSumBatch + u) + lane;
            nxt[u] = i < full ? x[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) {
            const int64_t cb = c + 64 * u;
            const int64_t lim = min((int64_t)64, full - cb);
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const float o = __shfl(cur[u], (lane & 7) + 8 * s, 64);
                if (8 * s < lim) part = part + o;
            }
        }
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) cur[u] = nxt[u];
    }
    float r = 0.0f;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const float pl = __shfl(part, l, 64);
        r = l == 0 ? pl : r + pl;
    }
    for (int64_t i = full; i < n; ++i) r = r + x[i];
    return r;
}

// rows of the dense n x n S with vector_sum > 1e-15 scaled by 1 / sum; one
// wave per row
__global__ __launch_bounds__(kBlock) void k_dense_row_normalize(float* __restrict__ S,
                                                                int64_t n) {
    const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (row >= n) return;
    const int lane = threadIdx.x & 63;
    float* r = S + row * n;
    const float s = wave_sum8(r, n);
    if (!(s > kEps)) return;
    const float inv = 1.0f / s;
    for (int64_t j = lane; j < n; j += 64) r[j] = r[j] * inv;
}

// transition_matrix_from_velocity into the zeroed dense n x n S: per row
// the sum of max(v, 0) in stored order; positive entries stored as v / sum
// (later ones overwrite), or S[i][i] = 1 when the sum is at most 1e-15.
// One thread per row.
__global__ __launch_bounds__(kBlock) void k_from_velocity(const int64_t* __restrict__ ip,
                                                          const int64_t* __restrict__ ix,
                                                          const float* __restrict__ v,
                                                          int64_t n, float* __restrict__ S) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s += v[k] > 0.0f ? v[k] : 0.0f;
    float* row = S + i * n;
    if (!(s > kEps)) {
        row[i] = 1.0f;
        return;
    }
    const float inv = 1.0f / s;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k)
        if (v[k] > 0.0f) row[ix[k]] = v[k] * inv;
}

// stationary distribution, one iteration after y = T^T pi: wave 0 of block 0
// forms vector_sum(y) and stores the scale (1 / sum when it exceeds 1e-15,
// else 1) in scale[0]
__global__ void k_stationary_sum(const float* __restrict__ y, int64_t n,
                                 float* __restrict__ scale) {
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    const float s = wave_sum8(y, n);
    if (threadIdx.x == 0) scale[0] = s > kEps ? 1.0f / s : 1.0f;
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

// y *= scale (1 when the sum was at most 1e-15: exact), partial maxima of
// |y - pi|
__global__ __launch_bounds__(kBlock) void k_stationary_scale(float* __restrict__ y,
                                                             const float* __restrict__ pi,
                                                             int64_t n,
                                                             const float* __restrict__ scale,
                                                             float* __restrict__ part) {
    const float inv = scale[0];
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const float yi = y[i] * inv;
        y[i] = yi;
        m = fmaxf(m, fabsf(yi - pi[i]));
    }
    m = block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(kBlock) void k_max_fold(const float* __restrict__ part, int m,
                                                     float* __restrict__ out) {
    float v = 0.0f;
    for (int i = threadIdx.x; i < m; i += kBlock) v = fmaxf(v, part[i]);
    v = block_max(v);
    if (threadIdx.x == 0) out[0] = v;
}

unsigned grid_for(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 20));
}

}  // namespace

KernelWeights kernel_weights(float bandwidth, int kind) {
    KernelWeights p;
    const float h = bandwidth > 1e-10f ? bandwidth : 1e-10f;
    p.inv_h = 1.0f / h;
    p.inv_h2 = p.inv_h * p.inv_h;
    p.inv_2h2 = 0.5f * p.inv_h2;
    p.h2 = h * h;
    p.kind = kind;
    return p;
}

void launch_kernel_values(const float* d, int64_t nnz, const KernelWeights& p, float* out,
                          hipStream_t s) {
    if (nnz <= 0) return;
    hipLaunchKernelGGL(k_kernel_values, dim3(grid_for(nnz)), dim3(kBlock), 0, s, d, nnz, p, out);
    check_launch("kernel values launch");
}

void launch_kernel_row_sums(const int64_t* ip, const float* d, int64_t n,
                            const KernelWeights& p, float* sums, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_kernel_row_sums, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, ip, d, n, p, sums);
    check_launch("kernel row sums launch");
}

void launch_row_normalize(const int64_t* ip, const float* v, int64_t n, float* out,
                          hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_row_normalize, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, ip, v, n, out);
    check_launch("row normalisation launch");
}

void launch_symmetrize(const int64_t* ip, const int64_t* ix, const float* v, int64_t n,
                       bool serial, float* S, hipStream_t s) {
    if (n <= 0) return;
    if (serial)
        hipLaunchKernelGGL(k_symmetrize_serial, dim3(1), dim3(64), 0, s, ip, ix, v, n, S);
    else
        hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, ip, ix, v, n, S);
    hipLaunchKernelGGL(k_dense_row_normalize, dim3((unsigned)((n * 64 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, S, n);
    check_launch("symmetrise launch");
}

void launch_from_velocity(const int64_t* ip, const int64_t* ix, const float* v, int64_t n,
                          float* S, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_from_velocity, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, ip, ix, v, n, S);
    check_launch("velocity transition launch");
}

void launch_stationary_update(float* y, const float* pi, int64_t n, float* scale, float* part,
                              float* diff, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_stationary_sum, dim3(1), dim3(64), 0, s, y, n, scale);
    const int64_t b = std::min<int64_t>((n + kBlock - 1) / kBlock, kParts);
    hipLaunchKernelGGL(k_stationary_scale, dim3((unsigned)b), dim3(kBlock), 0, s, y, pi, n,
                       scale, part);
    hipLaunchKernelGGL(k_max_fold, dim3(1), dim3(kBlock), 0, s, part, (int)b, diff);
    check_launch("stationary update launch");
}

}  // namespace scl_amd
