// walks.hip -- gfx950 kernels of the vector-side diffusion consumers (lazy
// walk, heat kernel signature, random walk with restart, diffusion
// pseudotime, diffusion maps, diffusion distance).  The SpMV / SpMM steps run
// on the row-split kernels of kernels.hip; these are the element-wise updates
// between steps, their reductions, the Gram-Schmidt pass and the pairwise
// distances.  Contraction is off: every fused multiply-add below is an
// explicit fmaf, matching the CPU restatement operation for operation.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace scl_amd {

using namespace dev;

namespace {

constexpr int kParts = 1024;  // first-level partials of a reduction

__device__ __forceinline__ float block_reduce(float v, bool is_max) {
    __shared__ float red[kBlock / 64];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        const float o = __shfl_xor(v, m, 64);
        v = is_max ? fmaxf(v, o) : v + o;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < kBlock / 64; ++w) r = is_max ? fmaxf(r, red[w]) : r + red[w];
        red[0] = r;
    }
    __syncthreads();
    const float r = red[0];
    __syncthreads();
    return r;
}

// x[i] = fma(a, x[i], b * t[i])
__global__ __launch_bounds__(kBlock) void k_blend(float* __restrict__ x,
                                                  const float* __restrict__ t, int64_t n,
                                                  float a, float b) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const float bt = b * t[i];
        x[i] = __builtin_fmaf(a, x[i], bt);
    }
}

// y[i] = fma(alpha, y[i], beta * r[i]); partial sums of |y[i] - x[i]|
// Once *done is set the walk has converged: y takes x unchanged, so the
// converged vector stays in both buffers whatever the steps after it do.
__global__ __launch_bounds__(kBlock) void k_rwr_update(float* __restrict__ y,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ r,
                                                       int64_t n, float alpha, float beta,
                                                       float* __restrict__ part,
                                                       const float* __restrict__ done) {
    if (*done != 0.0f) {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * kBlock)
            y[i] = x[i];
        return;
    }
    float s = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const float br = beta * r[i];
        const float yi = __builtin_fmaf(alpha, y[i], br);
        y[i] = yi;
        s += fabsf(yi - x[i]);
    }
    s = block_reduce(s, false);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// One DPT step's bookkeeping (diffusion.hpp:710-733) and partial maxima of
// 1 - visited; nothing once *done is set.
__global__ __launch_bounds__(kBlock) void k_dpt_update(const float* __restrict__ pn,
                                                       float* __restrict__ vis,
                                                       float* __restrict__ hit, int64_t n,
                                                       float it, float* __restrict__ part,
                                                       const float* __restrict__ done) {
    if (*done != 0.0f) return;
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        float v = vis[i];
        const float p = pn[i];
        if (v < 0.99f && p > 1e-15f) {
            const float rem = 1.0f - v;
            const float np = p * rem;
            hit[i] = __builtin_fmaf(it, np, hit[i]);
            v = v + np;
            v = v < 1.0f ? v : 1.0f;
            vis[i] = v;
        }
        m = fmaxf(m, 1.0f - v);
    }
    m = block_reduce(m, true);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// hit /= vis where vis > 0.5; partial maxima of those hitting times
__global__ __launch_bounds__(kBlock) void k_dpt_normalise(const float* __restrict__ vis,
                                                          float* __restrict__ hit, int64_t n,
                                                          float* __restrict__ part) {
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        if (vis[i] > 0.5f) {
            const float h = hit[i] / vis[i];
            hit[i] = h;
            m = fmaxf(m, h);
        }
    }
    m = block_reduce(m, true);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// pt = vis > 0.5 ? hit * inv_max : 1, with inv_max from the device maximum
__global__ __launch_bounds__(kBlock) void k_dpt_finish(const float* __restrict__ vis,
                                                       const float* __restrict__ hit,
                                                       const float* __restrict__ max_time,
                                                       float* __restrict__ pt, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float mt = *max_time;
    const float inv = mt > 0.0f ? 1.0f / mt : 1.0f;
    pt[i] = vis[i] > 0.5f ? hit[i] * inv : 1.0f;
}

// Second level: one block folds the partials into out[0].
__global__ __launch_bounds__(kBlock) void k_fold(const float* __restrict__ part, int m,
                                                 int is_max, float* __restrict__ out) {
    float v = 0.0f;
    for (int i = threadIdx.x; i < m; i += kBlock)
        v = is_max ? fmaxf(v, part[i]) : v + part[i];
    v = block_reduce(v, is_max != 0);
    if (threadIdx.x == 0) out[0] = v;
}

// k_fold, then *done = 1 when the result is below tol; nothing once *done
// is set (the step that converged keeps its value in out).
__global__ __launch_bounds__(kBlock) void k_fold_until(const float* __restrict__ part, int m,
                                                       int is_max, float* __restrict__ out,
                                                       float tol, float* __restrict__ done) {
    if (*done != 0.0f) return;
    float v = 0.0f;
    for (int i = threadIdx.x; i < m; i += kBlock)
        v = is_max ? fmaxf(v, part[i]) : v + part[i];
    v = block_reduce(v, is_max != 0);
    if (threadIdx.x == 0) {
        out[0] = v;
        if (v < tol) *done = 1.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_set_unit(float* __restrict__ x, int64_t n,
                                                     int64_t at, float value) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = (i == at) ? value : 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_fill(float* __restrict__ x, int64_t n,
                                                 float value) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        x[i] = value;
}

__global__ __launch_bounds__(kBlock) void k_add(float* __restrict__ acc,
                                                const float* __restrict__ x, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) acc[i] = acc[i] + x[i];
}

__global__ __launch_bounds__(kBlock) void k_scale(float* __restrict__ x, int64_t n, float a) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = x[i] * a;
}

// Q[i][c] = cos(ph t) + 0.1 sin(ph t 2.7), t = i / n, ph = (c + 1) pi
// (diffusion.hpp:922-929)
__global__ __launch_bounds__(kBlock) void k_dmap_init(float* __restrict__ Q, int64_t n,
                                                      int64_t K) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n * K) return;
    const int64_t i = e / K, c = e - i * K;
    const float t = (float)i / (float)n;
    const float ph = (float)(c + 1) * 3.14159265358979323846f;
    const float s = 0.1f * sinf(ph * t * 2.7f);
    Q[e] = cosf(ph * t) + s;
}

// Modified Gram-Schmidt of the K columns of row-major n x K Q in one
// workgroup (orthogonalize_vectors, diffusion.hpp:370-446): projections onto
// earlier columns, one extra pass when some |<w, q_p>| / |w| >= 0.7,
// normalisation, unit vector e_(c mod n) for a degenerate column.  w is an
// n-float workspace.  Sums are block tree sums.
__global__ __launch_bounds__(kBlock) void k_mgs(float* __restrict__ Q, float* __restrict__ w,
                                                int64_t n, int64_t K) {
    const int tid = threadIdx.x;
    for (int64_t c = 0; c < K; ++c) {
        for (int64_t i = tid; i < n; i += kBlock) w[i] = Q[i * K + c];
        __syncthreads();
        for (int reorth = 0; reorth <= 2; ++reorth) {
            for (int64_t p = 0; p < c; ++p) {
                float dot = 0.0f, nrm = 0.0f;
                for (int64_t i = tid; i < n; i += kBlock) {
                    const float q = Q[i * K + p];
                    dot = __builtin_fmaf(w[i], q, dot);
                    nrm = __builtin_fmaf(q, q, nrm);
                }
                dot = block_reduce(dot, false);
                nrm = block_reduce(nrm, false);
                if (nrm > 1e-15f) {
                    const float coeff = dot / nrm;
                    for (int64_t i = tid; i < n; i += kBlock)
                        w[i] = __builtin_fmaf(-coeff, Q[i * K + p], w[i]);
                }
                __syncthreads();
            }
            float nv = 0.0f;
            for (int64_t i = tid; i < n; i += kBlock) nv = __builtin_fmaf(w[i], w[i], nv);
            nv = block_reduce(nv, false);
            if (reorth == 0 && c > 0) {
                float max_cos = 0.0f;
                for (int64_t p = 0; p < c; ++p) {
                    float dot = 0.0f;
                    for (int64_t i = tid; i < n; i += kBlock)
                        dot = __builtin_fmaf(w[i], Q[i * K + p], dot);
                    dot = block_reduce(dot, false);
                    const float cs = nv > 1e-15f ? fabsf(dot) / sqrtf(nv) : 0.0f;
                    max_cos = fmaxf(max_cos, cs);
                }
                if (max_cos < 0.7f) break;
            } else {
                break;
            }
        }
        float nrm = 0.0f;
        for (int64_t i = tid; i < n; i += kBlock) nrm = __builtin_fmaf(w[i], w[i], nrm);
        nrm = block_reduce(nrm, false);
        if (nrm > 1e-15f) {
            const float inv = 1.0f / sqrtf(nrm);
            for (int64_t i = tid; i < n; i += kBlock) Q[i * K + c] = w[i] * inv;
        } else {
            const int64_t u = c % n;
            for (int64_t i = tid; i < n; i += kBlock) Q[i * K + c] = (i == u) ? 1.0f : 0.0f;
        }
        __syncthreads();
    }
}

// A = I (n x n, row-major)
__global__ __launch_bounds__(kBlock) void k_identity(float* __restrict__ A, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < n * n) A[e] = (e / n == e % n) ? 1.0f : 0.0f;
}

// --- multi-workgroup Gram-Schmidt over a column-major K x n matrix C ------
// (columns contiguous).  Each step launch ends with one (dot, norm) partial
// pair per workgroup; the next launch folds those partials again in every
// workgroup, in the same order, so every workgroup derives the same
// coefficient and kernel boundaries are the only synchronisation (no
// cross-workgroup fences, which on this GPU write back an XCD's whole L2).
// Partials alternate between two buffers of gridDim.x pairs.  S holds
//   S[2] max |<w, c_p>| / |w| of the re-orthogonalisation test
//   S[3] 1 if the second pass runs

// (sum of the first, sum of the second entries) of m partial pairs, the
// same in every workgroup
__device__ __forceinline__ void gs_fold(const float* __restrict__ part, int m, float& d,
                                        float& q) {
    d = 0.0f;
    q = 0.0f;
    for (int i = threadIdx.x; i < m; i += kBlock) {
        d += part[2 * i];
        q += part[2 * i + 1];
    }
    d = block_reduce(d, false);
    q = block_reduce(q, false);
}

enum : int { kPreNone = 0, kPreCopy = 1, kPreAxpy = 2 };

// One step of the pass.  First the update of w that the previous step
// decided: pre = kPreCopy copies src (the column) into w; kPreAxpy applies
// w -= (d / q) * src, (d, q) folded from prev, when q > 1e-15 and pre_gate
// is open.  Then the partial sums of <w, b> and <b, b> (b = w when b is
// null) into cur.  A closed `gate` skips the whole step (it is only passed
// where it also closes pre_gate).  w, src and b are distinct columns, or b
// is null; prev and cur are distinct.
__global__ __launch_bounds__(kBlock) void k_gs_step(float* w, const float* src, int pre,
                                                    const float* pre_gate, const float* b,
                                                    int64_t n, const float* __restrict__ prev,
                                                    float* __restrict__ cur,
                                                    const float* __restrict__ gate) {
    if (gate && *gate == 0.0f) return;
    bool upd = pre == kPreCopy;
    float coeff = 0.0f;
    if (pre == kPreAxpy && !(pre_gate && *pre_gate == 0.0f)) {
        float d, q;
        gs_fold(prev, (int)gridDim.x, d, q);
        upd = q > 1e-15f;
        coeff = upd ? d / q : 0.0f;
    }
    float d = 0.0f, q = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        float wi;
        if (pre == kPreCopy) {
            wi = src[i];
            w[i] = wi;
        } else {
            wi = w[i];
            if (upd) {
                wi = __builtin_fmaf(-coeff, src[i], wi);
                w[i] = wi;
            }
        }
        const float bi = b ? b[i] : wi;
        d = __builtin_fmaf(wi, bi, d);
        q = __builtin_fmaf(bi, bi, q);
    }
    d = block_reduce(d, false);
    q = block_reduce(q, false);
    if (threadIdx.x == 0) {
        cur[2 * blockIdx.x] = d;
        cur[2 * blockIdx.x + 1] = q;
    }
}

// Partials of <w, c_p> for p < v: part[p * gridDim.x + workgroup].
__global__ __launch_bounds__(kBlock) void k_gs_check(const float* __restrict__ w,
                                                     const float* __restrict__ C, int64_t n,
                                                     int64_t v, float* __restrict__ part) {
    for (int64_t p = 0; p < v; ++p) {
        const float* c = C + p * n;
        float d = 0.0f;
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * kBlock)
            d = __builtin_fmaf(w[i], c[i], d);
        d = block_reduce(d, false);
        if (threadIdx.x == 0) part[p * gridDim.x + blockIdx.x] = d;
    }
}

// The re-orthogonalisation test (one workgroup): |w|^2 folded from the m
// partial pairs in fin, then S[2] = max over p < v of |<w, c_p>| / |w|
// (0 when |w|^2 <= 1e-15), folded from check, and S[3] = (S[2] >= 0.7).
__global__ __launch_bounds__(kBlock) void k_gs_test(const float* __restrict__ fin,
                                                    const float* __restrict__ check, int m,
                                                    int64_t v, float* __restrict__ S) {
    float nv, unused;
    gs_fold(fin, m, nv, unused);
    float mx = 0.0f;
    for (int64_t p = 0; p < v; ++p) {
        float d = 0.0f;
        for (int i = threadIdx.x; i < m; i += kBlock) d += check[p * m + i];
        d = block_reduce(d, false);
        mx = fmaxf(mx, nv > 1e-15f ? fabsf(d) / sqrtf(nv) : 0.0f);
    }
    if (threadIdx.x == 0) {
        S[2] = mx;
        S[3] = mx >= 0.7f ? 1.0f : 0.0f;
    }
}

// FastRNG of pseudotime.hpp:153-182: four chained rounds of the first half of
// a SplitMix64 output seed the state; each step is the bit-mirrored form of
// Marsaglia's xorshift128; uniform() = next() * 2^-32.
struct FastRng {
    uint32_t s[4];
    __device__ explicit FastRng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9e3779b97f4a7c15ULL;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            s[i] = (uint32_t)(z >> 32);
        }
    }
    __device__ float uniform() {
        uint32_t t = s[3];
        const uint32_t x = s[0];
        s[3] = s[2];
        s[2] = s[1];
        s[1] = x;
        t ^= t >> 11;
        t ^= t << 8;
        s[0] = t ^ x ^ (x << 19);
        return (float)s[0] * 2.3283064365386963e-10f;
    }
};

// Column v = w / |w|, |w|^2 folded from the m partial pairs in fin.  A
// degenerate column (|w|^2 <= 1e-15) becomes e_(v mod n) when rng_mult is 0
// (diffusion maps), else the normalised vector uniform() - 0.5 of
// FastRng(v * rng_mult), drawn and summed in index order by one thread
// (pseudotime.hpp:460-471), with w as its scratch.
__global__ __launch_bounds__(kBlock) void k_gs_store(float* __restrict__ c,
                                                     float* __restrict__ w, int64_t n,
                                                     int64_t v, const float* __restrict__ fin,
                                                     int m, uint64_t rng_mult) {
    float nv, unused;
    gs_fold(fin, m, nv, unused);
    const float inv = nv > 1e-15f ? 1.0f / sqrtf(nv) : 0.0f;
    if (inv == 0.0f && rng_mult != 0) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        FastRng r((uint64_t)v * rng_mult);
        float nrm = 0.0f;
        for (int64_t i = 0; i < n; ++i) {
            const float x = r.uniform() - 0.5f;
            w[i] = x;
            nrm = __builtin_fmaf(x, x, nrm);
        }
        const float s = 1.0f / sqrtf(nrm);
        for (int64_t i = 0; i < n; ++i) c[i] = w[i] * s;
        return;
    }
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (inv != 0.0f) c[i] = w[i] * inv;
    else c[i] = i == v % n ? 1.0f : 0.0f;
}

// B (K x n) = A (n x K)^T and back, through LDS tiles of 16 x 16; one
// workgroup per tile on a 1-D grid (tile rows outer)
__global__ __launch_bounds__(kBlock) void k_transpose_rect(const float* __restrict__ A,
                                                           float* __restrict__ B,
                                                           int64_t rows, int64_t cols) {
    __shared__ float tile[16][17];
    const int64_t tiles_c = (cols + 15) / 16;
    const int64_t bi = ((int64_t)blockIdx.x / tiles_c) * 16;
    const int64_t bj = ((int64_t)blockIdx.x % tiles_c) * 16;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    if (bi + ty < rows && bj + tx < cols) tile[ty][tx] = A[(bi + ty) * cols + bj + tx];
    __syncthreads();
    if (bj + ty < cols && bi + tx < rows) B[(bj + ty) * rows + bi + tx] = tile[tx][ty];
}

// B = A^T for n x n row-major A (16 x 16 tiles through LDS).
__global__ __launch_bounds__(kBlock) void k_transpose(const float* __restrict__ A,
                                                      float* __restrict__ B, int64_t n) {
    __shared__ float tile[16][17];
    const int64_t bi = (int64_t)blockIdx.y * 16, bj = (int64_t)blockIdx.x * 16;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    if (bi + ty < n && bj + tx < n) tile[ty][tx] = A[(bi + ty) * n + bj + tx];
    __syncthreads();
    if (bj + ty < n && bi + tx < n) B[(bj + ty) * n + bi + tx] = tile[tx][ty];
}

// D[i][j] = D[j][i] = sqrt(sum_k (A[i][k] - A[j][k])^2) for i < j, summed in
// k order, D[i][i] = 0.  A workgroup takes 64 x 64 pairs (rows bi.., bj..)
// and a thread 4 x 4 of them, each a sequential chain of FMAs; the two row
// blocks pass through LDS 16 features at a time, stored feature-major so a
// thread reads its four rows' values with one 16-byte load.  Only blocks on
// or above the diagonal work.
__global__ __launch_bounds__(kBlock) void k_pairwise(const float* __restrict__ A,
                                                     float* __restrict__ D, int64_t n) {
    constexpr int T = 64, KC = 16, PAD = 4;
    __shared__ float ai[KC][T + PAD], aj[KC][T + PAD];
    const int64_t bi = (int64_t)blockIdx.y * T, bj = (int64_t)blockIdx.x * T;
    if (bj + T <= bi) return;  // wholly below the diagonal
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    float s[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) s[r][c] = 0.0f;
    for (int64_t k0 = 0; k0 < n; k0 += KC) {
        for (int e = threadIdx.x; e < T * KC; e += kBlock) {
            const int r = e / KC, c = e % KC;
            const int64_t k = k0 + c;
            ai[c][r] = (bi + r < n && k < n) ? A[(bi + r) * n + k] : 0.0f;
            aj[c][r] = (bj + r < n && k < n) ? A[(bj + r) * n + k] : 0.0f;
        }
        __syncthreads();
        const int kmax = (int)(n - k0 < KC ? n - k0 : KC);
        for (int c = 0; c < kmax; ++c) {
            const f4 av = *reinterpret_cast<const f4*>(&ai[c][ty * 4]);
            const f4 bv = *reinterpret_cast<const f4*>(&aj[c][tx * 4]);
            const float ar[4] = {av.x, av.y, av.z, av.w};
            const float bc[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float d = ar[r] - bc[q];
                    s[r][q] = __builtin_fmaf(d, d, s[r][q]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t i = bi + ty * 4 + r, j = bj + tx * 4 + q;
            if (i >= n || j >= n) continue;
            if (i == j) {
                D[i * n + i] = 0.0f;
            } else if (i < j) {
                const float dist = sqrtf(s[r][q]);
                D[i * n + j] = dist;
                D[j * n + i] = dist;
            }
        }
}

// ---------------------------------------------------- Cholesky QR ------
// Q = M R^-1 with R^T R = M^T M (FP64), the Q factor of M with a positive
// diagonal of R, the basis modified Gram-Schmidt produces in exact
// arithmetic.  Four launches per orthonormalisation: Gram partials, their
// sum per entry, the Cholesky factor on one workgroup, the row update.

// Gram entry t (a <= b, row-major upper triangle) of a K x K matrix.
__device__ __forceinline__ void pair_of(int t, int K, int& a, int& b) {
    a = 0;
    int rem = t;
    while (rem >= K - a) {
        rem -= K - a;
        ++a;
    }
    b = a + rem;
}

// Per-workgroup partials of M^T M over tiles of kBlock rows (row-major n x K
// M), one pair per thread, summed in row order with FP64 FMAs.
__global__ __launch_bounds__(kBlock) void k_gram_partial(const float* __restrict__ M, int64_t n,
                                                         int K, double* __restrict__ part) {
    __shared__ float tile[kBlock][kCholQrMaxK + 1];
    const int t = threadIdx.x, pairs = K * (K + 1) / 2;
    int a = 0, b = 0;
    if (t < pairs) pair_of(t, K, a, b);
    double acc = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * kBlock; r0 < n; r0 += (int64_t)gridDim.x * kBlock) {
        const int64_t rows = n - r0 < kBlock ? n - r0 : kBlock;
        __syncthreads();
        for (int i = t; i < kBlock * K; i += kBlock) {
            const int r = i / K, c = i - r * K;
            tile[r][c] = r < rows ? M[(r0 + r) * K + c] : 0.0f;
        }
        __syncthreads();
        if (t < pairs)
            for (int r = 0; r < kBlock; ++r)
                acc = __builtin_fma((double)tile[r][a], (double)tile[r][b], acc);
    }
    if (t < pairs) part[(int64_t)blockIdx.x * pairs + t] = acc;
}

// Gram entry blockIdx.x summed over the partials: thread j adds partials
// j, j + kBlock, ... in order, then a fixed pairwise tree over the threads.
__global__ __launch_bounds__(kBlock) void k_gram_reduce(const double* __restrict__ part,
                                                        int n_parts, int pairs,
                                                        double* __restrict__ gram) {
    __shared__ double sh[kBlock];
    const int p = blockIdx.x, j = threadIdx.x;
    double s = 0.0;
    for (int w = j; w < n_parts; w += kBlock) s += part[(int64_t)w * pairs + p];
    sh[j] = s;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (j < h) sh[j] += sh[j + h];
        __syncthreads();
    }
    if (j == 0) gram[p] = sh[0];
}

// Factors G = R^T R and writes R^-1 (K x K, upper triangular, FP64).  A
// pivot at or below kCholQrMinPivot, or not finite, sets *degenerate (never
// cleared here) and writes the identity.  One thread factors in LDS; each
// column of R^-1 is then solved by its own thread.
__global__ __launch_bounds__(kBlock) void k_cholqr_factor(const double* __restrict__ gram, int K,
                                                          double* __restrict__ rinv,
                                                          int* __restrict__ degenerate) {
    __shared__ double G[kCholQrMaxK][kCholQrMaxK];
    __shared__ double R[kCholQrMaxK][kCholQrMaxK];
    __shared__ double Ri[kCholQrMaxK][kCholQrMaxK];
    __shared__ int bad;
    const int t = threadIdx.x, pairs = K * (K + 1) / 2;
    if (t < pairs) {
        int a, b;
        pair_of(t, K, a, b);
        G[a][b] = gram[t];
    }
    if (t == 0) bad = 0;
    __syncthreads();
    if (t == 0) {
        for (int j = 0; j < K; ++j) {
            double d = G[j][j];
            for (int k = 0; k < j; ++k) d -= R[k][j] * R[k][j];
            if (!(d > kCholQrMinPivot) || !(d < 1e300)) {
                bad = 1;
                break;
            }
            const double rjj = sqrt(d);
            R[j][j] = rjj;
            for (int i = j + 1; i < K; ++i) {
                double v = G[j][i];
                for (int k = 0; k < j; ++k) v -= R[k][j] * R[k][i];
                R[j][i] = v / rjj;
            }
        }
    }
    __syncthreads();
    if (bad) {
        if (t == 0) *degenerate = 1;
        for (int i = t; i < K * K; i += kBlock) rinv[i] = (i / K == i % K) ? 1.0 : 0.0;
        return;
    }
    // column c of R^-1: back substitution of R x = e_c
    if (t < K) {
        const int c = t;
        for (int i = K - 1; i >= 0; --i) {
            if (i > c) {
                Ri[i][c] = 0.0;
                continue;
            }
            double v = i == c ? 1.0 : 0.0;
            for (int k = i + 1; k <= c; ++k) v -= R[i][k] * Ri[k][c];
            Ri[i][c] = v / R[i][i];
        }
    }
    __syncthreads();
    for (int i = t; i < K * K; i += kBlock) rinv[i] = Ri[i / K][i % K];
}

// M <- M R^-1, row by row in FP64, rounded once to fp32.
__global__ __launch_bounds__(kBlock) void k_cholqr_apply(float* __restrict__ M, int64_t n, int K,
                                                         const double* __restrict__ rinv) {
    __shared__ double Ri[kCholQrMaxK * kCholQrMaxK];
    for (int i = threadIdx.x; i < K * K; i += kBlock) Ri[i] = rinv[i];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    float q[kCholQrMaxK];
    for (int c = 0; c < K; ++c) q[c] = M[r * K + c];
    for (int c = 0; c < K; ++c) {
        double v = 0.0;
        for (int p = 0; p <= c; ++p) v = __builtin_fma((double)q[p], Ri[p * K + c], v);
        M[r * K + c] = (float)v;
    }
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
unsigned parts_for(int64_t n) {
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)(b < kParts ? (b > 0 ? b : 1) : kParts);
}

}  // namespace

void launch_blend(float* x, const float* t, int64_t n, float a, float b, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_blend, dim3(blocks_for(n)), dim3(kBlock), 0, s, x, t, n, a, b);
    check_launch("blend launch");
}

void launch_rwr_update(float* y, const float* x, const float* r, int64_t n, float alpha,
                       float beta, float* part, float* total, float tol, float* done,
                       hipStream_t s) {
    const unsigned p = parts_for(n);
    hipLaunchKernelGGL(k_rwr_update, dim3(p), dim3(kBlock), 0, s, y, x, r, n, alpha, beta,
                       part, done);
    hipLaunchKernelGGL(k_fold_until, dim3(1), dim3(kBlock), 0, s, part, (int)p, 0, total, tol,
                       done);
    check_launch("rwr update launch");
}

void launch_dpt_update(const float* pn, float* vis, float* hit, int64_t n, float it,
                       float* part, float* max_unvisited, float tol, float* done,
                       hipStream_t s) {
    const unsigned p = parts_for(n);
    hipLaunchKernelGGL(k_dpt_update, dim3(p), dim3(kBlock), 0, s, pn, vis, hit, n, it, part,
                       done);
    hipLaunchKernelGGL(k_fold_until, dim3(1), dim3(kBlock), 0, s, part, (int)p, 1,
                       max_unvisited, tol, done);
    check_launch("dpt update launch");
}

void launch_dpt_finish(const float* vis, float* hit, float* pt, int64_t n, float* part,
                       float* max_time, hipStream_t s) {
    const unsigned p = parts_for(n);
    hipLaunchKernelGGL(k_dpt_normalise, dim3(p), dim3(kBlock), 0, s, vis, hit, n, part);
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(kBlock), 0, s, part, (int)p, 1, max_time);
    hipLaunchKernelGGL(k_dpt_finish, dim3(blocks_for(n)), dim3(kBlock), 0, s, vis, hit,
                       max_time, pt, n);
    check_launch("dpt finish launch");
}

void launch_set_unit(float* x, int64_t n, int64_t at, float value, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_set_unit, dim3(blocks_for(n)), dim3(kBlock), 0, s, x, n, at, value);
    check_launch("fill launch");
}

void launch_fill(float* x, int64_t n, float value, hipStream_t s) {
    if (n <= 0) return;
    const int64_t b = std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 20);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)b), dim3(kBlock), 0, s, x, n, value);
    check_launch("fill launch");
}

void launch_add(float* acc, const float* x, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add, dim3(blocks_for(n)), dim3(kBlock), 0, s, acc, x, n);
    check_launch("add launch");
}

void launch_scale(float* x, int64_t n, float a, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(kBlock), 0, s, x, n, a);
    check_launch("scale launch");
}

void launch_dmap_init(float* Q, int64_t n, int64_t K, hipStream_t s) {
    if (n * K <= 0) return;
    hipLaunchKernelGGL(k_dmap_init, dim3(blocks_for(n * K)), dim3(kBlock), 0, s, Q, n, K);
    check_launch("embedding init launch");
}

void launch_mgs(float* Q, float* w, int64_t n, int64_t K, hipStream_t s) {
    if (n <= 0 || K <= 0) return;
    hipLaunchKernelGGL(k_mgs, dim3(1), dim3(kBlock), 0, s, Q, w, n, K);
    check_launch("gram-schmidt launch");
}

void launch_identity(float* A, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_identity, dim3(blocks_for(n * n)), dim3(kBlock), 0, s, A, n);
    check_launch("identity launch");
}

void launch_transpose_rect(const float* A, float* B, int64_t rows, int64_t cols,
                           hipStream_t s) {
    if (rows <= 0 || cols <= 0) return;
    const int64_t tiles = ((rows + 15) / 16) * ((cols + 15) / 16);
    if (tiles > 0x7fffffff) throw std::runtime_error("transpose too large");
    hipLaunchKernelGGL(k_transpose_rect, dim3((unsigned)tiles), dim3(kBlock), 0, s, A, B,
                       rows, cols);
    check_launch("transpose launch");
}

// Modified Gram-Schmidt of the K contiguous columns of C (K x n), across the
// whole GPU: w is an n-float workspace, part holds
// gram_schmidt_part_floats(n, K) partials, S 8 floats (the scalars above).
// reorth: the diffusion-map form (diffusion.hpp:370-446: a second pass when
// some |<w, c_p>| / |w| >= 0.7, degenerate columns become unit vectors);
// otherwise the pseudotime form (pseudotime.hpp:418-474: one pass,
// degenerate columns drawn from FastRng(v * rng_mult)).  Column v costs
// v + 2 launches in one pass, 2v + 5 with the test and the second pass.
void launch_gram_schmidt(float* C, float* w, int64_t n, int64_t K, float* part, float* S,
                         bool reorth, uint64_t rng_mult, hipStream_t s) {
    if (n <= 0 || K <= 0) return;
    const unsigned p = parts_for(n), b = blocks_for(n);
    float* pairs[2] = {part, part + 2 * (int64_t)p};
    float* check = part + 4 * (int64_t)p;
    int cur = 0;  // the buffer the next step writes
    auto step = [&](int pre, const float* src, const float* pre_gate, const float* bv,
                    const float* gate) {
        hipLaunchKernelGGL(k_gs_step, dim3(p), dim3(kBlock), 0, s, w, src, pre, pre_gate, bv, n,
                           pairs[cur ^ 1], pairs[cur], gate);
        cur ^= 1;
    };
    for (int64_t v = 0; v < K; ++v) {
        float* cv = C + v * n;
        for (int pass = 0; pass < (reorth ? 2 : 1) && (pass == 0 || v > 0); ++pass) {
            // the second pass runs only when the first left some
            // |<w, c_p>| / |w| >= 0.7 (gate S[3]); while it is closed no step
            // of the pass reads or writes partials
            const float* gate = pass == 0 ? nullptr : S + 3;
            // step q applies the coefficient of c_(q-1) and sums the partials
            // of c_q's; the first step of the column copies it into w
            for (int64_t q = 0; q < v; ++q) {
                if (q == 0) step(pass == 0 ? kPreCopy : kPreNone, cv, gate, C, gate);
                else step(kPreAxpy, C + (q - 1) * n, gate, C + q * n, gate);
            }
            // the last coefficient applied, then the partials of |w|^2
            if (v == 0) step(kPreCopy, cv, nullptr, nullptr, nullptr);
            else step(kPreAxpy, C + (v - 1) * n, gate, nullptr, nullptr);
            if (reorth && pass == 0 && v > 0) {
                hipLaunchKernelGGL(k_gs_check, dim3(p), dim3(kBlock), 0, s, w, C, n, v, check);
                hipLaunchKernelGGL(k_gs_test, dim3(1), dim3(kBlock), 0, s, pairs[cur ^ 1], check,
                                   (int)p, v, S);
            }
        }
        hipLaunchKernelGGL(k_gs_store, dim3(b), dim3(kBlock), 0, s, cv, w, n, v, pairs[cur ^ 1],
                           (int)p, rng_mult);
    }
    check_launch("gram-schmidt launch");
}

int64_t cholqr_parts(int64_t n) {
    return std::max<int64_t>(1, std::min<int64_t>(1024, (n + kBlock - 1) / kBlock));
}

int64_t cholqr_workspace_doubles(int64_t n, int64_t K) {
    return (cholqr_parts(n) + 1) * (K * (K + 1) / 2) + K * K;
}

void launch_cholqr(float* M, int64_t n, int64_t K, double* work, int* degenerate,
                   hipStream_t s) {
    if (n <= 0 || K <= 0) return;
    if (K > kCholQrMaxK) throw std::runtime_error("cholesky QR: too many columns");
    const int64_t parts = cholqr_parts(n), pairs = K * (K + 1) / 2;
    double* part = work;
    double* gram = work + parts * pairs;
    double* rinv = gram + pairs;
    hipLaunchKernelGGL(k_gram_partial, dim3((unsigned)parts), dim3(kBlock), 0, s, M, n, (int)K,
                       part);
    hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)pairs), dim3(kBlock), 0, s, part,
                       (int)parts, (int)pairs, gram);
    hipLaunchKernelGGL(k_cholqr_factor, dim3(1), dim3(kBlock), 0, s, gram, (int)K, rinv,
                       degenerate);
    hipLaunchKernelGGL(k_cholqr_apply, dim3(blocks_for(n)), dim3(kBlock), 0, s, M, n, (int)K,
                       rinv);
    check_launch("cholesky QR launch");
}

int64_t gram_schmidt_part_floats(int64_t n, int64_t K) {
    return (int64_t)parts_for(n) * (4 + K);
}

void launch_transpose_square(const float* A, float* B, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)((n + 15) / 16);
    hipLaunchKernelGGL(k_transpose, dim3(g, g), dim3(kBlock), 0, s, A, B, n);
    check_launch("transpose launch");
}

void launch_pairwise_rows(const float* A, float* D, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)((n + 63) / 64);
    hipLaunchKernelGGL(k_pairwise, dim3(g, g), dim3(kBlock), 0, s, A, D, n);
    check_launch("pairwise distance launch");
}

}  // namespace scl_amd
