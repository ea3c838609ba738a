// neighbors.hip -- gfx950 kernels of the exact k-nearest-neighbour search
// (scl/kernel/neighbors.hpp:300-476): row norms and the all-pairs distance
// scan with a running top-k per query.
//
// The reference's sparse dot products add the products of matching indices
// in increasing index order, one FMA each.  Over dense rows the same chain
// in feature order adds only exact zeros besides them, so the tiled product
// below, which keeps one sequential FMA chain per (query, candidate) over
// the features, yields the reference's dot products bit for bit.  The
// distance is fma(-2, dot, n_i + n_j), clamped at 0, as the reference's
// contracted "n_i + n_j - 2 * dot".  No MFMA: its reduction order is not a
// sequential chain.
// Contraction is off: every fused multiply-add is an explicit fmaf.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace scl_amd {

using namespace dev;

namespace {

// vectorize::sum_squared for 8-lane float vectors (vectorize.hpp:782-822)
// over one row's stored values, one wave per row: lane t < 32 keeps the
// accumulator of position t in every block of 32 (accumulator t / 8, lane
// t % 8); lanes 0..7 fold (a0 + a1) + (a2 + a3), run the remaining whole
// vectors, and the eight lanes are summed ((l0 + l4) + (l2 + l6)) +
// ((l1 + l5) + (l3 + l7)) before the tail.
__global__ __launch_bounds__(kBlock) void k_sum_squared(const int64_t* __restrict__ ip,
                                                        const float* __restrict__ v, int64_t n,
                                                        float* __restrict__ norms) {
    const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (row >= n) return;
    const int lane = threadIdx.x & 63;
    const int64_t beg = ip[row], len = ip[row + 1] - beg;
    const float* x = v + beg;
    float acc = 0.0f;
    int64_t i = 0;
    for (; i + 32 <= len; i += 32) {
        const float e = lane < 32 ? x[i + lane] : 0.0f;
        if (lane < 32) acc = __builtin_fmaf(e, e, acc);
    }
    const float a1 = __shfl(acc, (lane & 7) + 8, 64);
    const float a2 = __shfl(acc, (lane & 7) + 16, 64);
    const float a3 = __shfl(acc, (lane & 7) + 24, 64);
    float a = (acc + a1) + (a2 + a3);
    for (; i + 8 <= len; i += 8) {
        const float e = lane < 8 ? x[i + lane] : 0.0f;
        a = __builtin_fmaf(e, e, a);
    }
    float l[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) l[t] = __shfl(a, t, 64);
    float r = ((l[0] + l[4]) + (l[2] + l[6])) + ((l[1] + l[5]) + (l[3] + l[7]));
    for (; i < len; ++i) r = __builtin_fmaf(x[i], x[i], r);
    if (lane == 0) norms[row] = r;
}

constexpr int kQ = 64;    // queries per workgroup
constexpr int kC = 128;   // candidates per tile
constexpr int kK = 16;    // features per LDS stage
constexpr int kPad = 4;   // LDS row padding (floats)

// The reference's k-nearest search for one query (neighbors.hpp:330-476),
// restated for a thread that sees candidates in increasing index order:
// Cauchy-Schwarz pruning against the current k-th distance, a max-heap of k
// entries (sift-up while filling, replace the top only when strictly
// closer), and at the end the heap sort (min-heap build, extraction to the
// back, reversal).  h_d / h_j hold the heap (k entries, global memory).
struct KnnHeap {
    float* h_d;
    int32_t* h_j;
    int64_t k;
    int64_t cnt;
    float maxd;

    __device__ void swap(int64_t a, int64_t b) {
        const float td = h_d[a];
        const int32_t tj = h_j[a];
        h_d[a] = h_d[b];
        h_j[a] = h_j[b];
        h_d[b] = td;
        h_j[b] = tj;
    }
    __device__ void offer(float d2, int32_t j) {
        if (cnt < k) {
            h_d[cnt] = d2;
            h_j[cnt] = j;
            int64_t p = cnt;
            while (p > 0) {
                const int64_t par = (p - 1) / 2;
                if (h_d[p] <= h_d[par]) break;
                swap(p, par);
                p = par;
            }
            ++cnt;
            if (cnt == k) maxd = h_d[0];
        } else if (d2 < h_d[0]) {
            h_d[0] = d2;
            h_j[0] = j;
            int64_t p = 0;
            for (;;) {
                int64_t big = p;
                const int64_t l = 2 * p + 1, r = 2 * p + 2;
                if (l < cnt && h_d[l] > h_d[big]) big = l;
                if (r < cnt && h_d[r] > h_d[big]) big = r;
                if (big == p) break;
                swap(p, big);
                p = big;
            }
            maxd = h_d[0];
        }
    }
    __device__ void min_sift(int64_t p, int64_t size) {
        for (;;) {
            int64_t s = p;
            const int64_t l = 2 * p + 1, r = 2 * p + 2;
            if (l < size && h_d[l] < h_d[s]) s = l;
            if (r < size && h_d[r] < h_d[s]) s = r;
            if (s == p) break;
            swap(p, s);
            p = s;
        }
    }
    // ascending order of the reference's heap sort
    __device__ void sort() {
        if (cnt >= 2)
            for (int64_t m = cnt / 2; m > 0; --m) min_sift(m - 1, cnt);
        for (int64_t m = cnt; m > 1; --m) {
            swap(0, m - 1);
            min_sift(0, m - 1);
        }
        for (int64_t m = 0; m < cnt / 2; ++m) swap(m, cnt - 1 - m);
    }
};

// One workgroup owns kQ query rows of the dense, zero-padded n_pad x f_pad
// matrix X (n_pad a multiple of kC) and scans every candidate tile.  Thread
// (ty, tx) computes the 4 x 8 dot products of queries 4ty..4ty+3 with
// candidates 4tx..4tx+3 and 64+4tx..64+4tx+3, each a sequential FMA chain
// over the features.  A candidate can change a query's heap only while the
// heap is filling or when it is strictly closer than the heap's top at the
// start of the tile (the top never grows), so only those go to an LDS
// buffer; after each tile one thread per query sorts its buffer by index
// and offers the candidates to the query's heap in that order, with the
// reference's pruning test against the heap's current top.  The heap then
// evolves exactly as in the reference's scan over j = 0..n-1, and the final
// heap sort gives its output order, ties included.
__global__ __launch_bounds__(kBlock) void k_knn(const float* __restrict__ X,
                                                const float* __restrict__ norms, int64_t n,
                                                int64_t f_pad, int64_t k,
                                                float* __restrict__ ld, int32_t* __restrict__ lj,
                                                float* __restrict__ out_dist,
                                                int64_t* __restrict__ out_idx) {
    __shared__ float As[kK][kQ + kPad];
    __shared__ float Bs[kK][kC + kPad];
    __shared__ float bd[kQ][kC];
    __shared__ int bj[kQ][kC];
    __shared__ int cnt[kQ];
    __shared__ long long hcnt[kQ];
    __shared__ float top[kQ], qn[kQ];

    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    const int64_t q0 = (int64_t)blockIdx.x * kQ;
    if (tid < kQ) {
        cnt[tid] = 0;
        hcnt[tid] = 0;
        top[tid] = 3.40282347e38f;   // the reference's max_dist_sq before the heap fills
        qn[tid] = q0 + tid < n ? norms[q0 + tid] : 0.0f;
    }
    // loader mapping: thread tid moves 4 consecutive features of one query
    // row and of two candidate rows (lr and lr + 64)
    const int lr = tid >> 2, lk = (tid & 3) * 4;
    __syncthreads();

    for (int64_t c0 = 0; c0 < n; c0 += kC) {
        float acc[4][8];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[r][c] = 0.0f;
        for (int64_t k0 = 0; k0 < f_pad; k0 += kK) {
            const f4 a = *reinterpret_cast<const f4*>(X + (q0 + lr) * f_pad + k0 + lk);
            const f4 b0 = *reinterpret_cast<const f4*>(X + (c0 + lr) * f_pad + k0 + lk);
            const f4 b1 = *reinterpret_cast<const f4*>(X + (c0 + 64 + lr) * f_pad + k0 + lk);
            As[lk + 0][lr] = a.x;
            As[lk + 1][lr] = a.y;
            As[lk + 2][lr] = a.z;
            As[lk + 3][lr] = a.w;
            Bs[lk + 0][lr] = b0.x;
            Bs[lk + 1][lr] = b0.y;
            Bs[lk + 2][lr] = b0.z;
            Bs[lk + 3][lr] = b0.w;
            Bs[lk + 0][64 + lr] = b1.x;
            Bs[lk + 1][64 + lr] = b1.y;
            Bs[lk + 2][64 + lr] = b1.z;
            Bs[lk + 3][64 + lr] = b1.w;
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < kK; ++kk) {
                const f4 av = *reinterpret_cast<const f4*>(&As[kk][ty * 4]);
                const f4 bv0 = *reinterpret_cast<const f4*>(&Bs[kk][tx * 4]);
                const f4 bv1 = *reinterpret_cast<const f4*>(&Bs[kk][64 + tx * 4]);
                const float ar[4] = {av.x, av.y, av.z, av.w};
                const float bc[8] = {bv0.x, bv0.y, bv0.z, bv0.w, bv1.x, bv1.y, bv1.z, bv1.w};
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 8; ++c) acc[r][c] = __builtin_fmaf(ar[r], bc[c], acc[r][c]);
            }
            __syncthreads();
        }
        // distances of this tile; candidates that can still enter the heap
        // go to the buffer
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int ql = ty * 4 + r;
            const int64_t q = q0 + ql;
            if (q >= n) continue;
            const bool filling = hcnt[ql] < k;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int64_t j = c0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + (c - 4));
                if (j >= n || j == q) continue;
                float d2 = __builtin_fmaf(-2.0f, acc[r][c], qn[ql] + norms[j]);
                if (d2 < 0.0f) d2 = 0.0f;
                if (filling || d2 < top[ql]) {
                    const int slot = atomicAdd(&cnt[ql], 1);
                    bd[ql][slot] = d2;
                    bj[ql][slot] = (int)j;
                }
            }
        }
        __syncthreads();
        if (tid < kQ && cnt[tid] > 0) {
            const int ql = tid;
            const int m = cnt[ql];
            // insertion sort of the buffer by candidate index
            for (int a = 1; a < m; ++a) {
                const float kd = bd[ql][a];
                const int kj = bj[ql][a];
                int b = a - 1;
                while (b >= 0 && bj[ql][b] > kj) {
                    bd[ql][b + 1] = bd[ql][b];
                    bj[ql][b + 1] = bj[ql][b];
                    --b;
                }
                bd[ql][b + 1] = kd;
                bj[ql][b + 1] = kj;
            }
            KnnHeap h{ld + (q0 + ql) * k, lj + (q0 + ql) * k, k, (int64_t)hcnt[ql], top[ql]};
            const float si = sqrtf(qn[ql]);
            for (int a = 0; a < m; ++a) {
                const float df = si - sqrtf(norms[bj[ql][a]]);
                if (df * df >= h.maxd) continue;   // neighbors.hpp:420-425
                h.offer(bd[ql][a], bj[ql][a]);
            }
            hcnt[ql] = h.cnt;
            top[ql] = h.maxd;
            cnt[ql] = 0;
        }
        __syncthreads();
    }
    if (tid < kQ && q0 + tid < n) {
        const int64_t q = q0 + tid;
        KnnHeap h{ld + q * k, lj + q * k, k, (int64_t)hcnt[tid], top[tid]};
        h.sort();
        for (int64_t m = 0; m < k; ++m) {
            if (m < h.cnt) {
                out_dist[q * k + m] = sqrtf(h.h_d[m]);
                out_idx[q * k + m] = h.h_j[m];
            } else {
                out_dist[q * k + m] = __builtin_inff();
                out_idx[q * k + m] = -1;
            }
        }
    }
}

}  // namespace

void launch_sum_squared(const int64_t* ip, const float* v, int64_t n, float* norms,
                        hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sum_squared, dim3((unsigned)((n * 64 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, ip, v, n, norms);
    check_launch("norms launch");
}

int64_t knn_rows_padded(int64_t n) { return (n + kC - 1) / kC * kC; }
int64_t knn_features_padded(int64_t f) { return (f + kK - 1) / kK * kK; }

void launch_knn(const float* X, const float* norms, int64_t n, int64_t f_pad, int64_t k,
                float* ld, int32_t* lj, float* out_dist, int64_t* out_idx, hipStream_t s) {
    if (n <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_knn, dim3((unsigned)((n + kQ - 1) / kQ)), dim3(kBlock), 0, s, X, norms,
                       n, f_pad, k, ld, lj, out_dist, out_idx);
    check_launch("knn launch");
}

}  // namespace scl_amd
