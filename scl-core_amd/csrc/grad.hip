// grad.hip -- gfx950 kernels of the reverse-mode diffusion.
//
// For L(Y), Y = X_t, X_s = P X_{s-1}, the backward pass needs per step s:
//   G_{s-1} = P^T G_s                          (adjoint SpMM)
//   dP_ij  += <G_s[i,:], X_{s-1}[j,:]>         (edge gradient, per nonzero)
// Both read G_s[i,:] for every nonzero (i, j).  Walking the transposed
// pattern (one slot per column j, entries = source rows i in increasing
// order), one gather of G_s[i, slab] feeds both: the adjoint accumulates
// p_ij * G_s[i, :] into G_{s-1}[j, :] (fp32 FMA, stored order), and the edge
// term is the dot product of that same gathered row with X_{s-1}[j, :],
// which is loaded once per column.  The dot product is reduced across the
// lanes of the row's group with xor shuffles and written once per nonzero,
// in transposed order (dPT[p]); the host driver permutes dPT to CSR order
// once at the end of the backward pass.
#include <hip/hip_runtime.h>

#include "adjoint_row.hpp"
#include "device_common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace scl_amd {

using namespace dev;

namespace {

// One backward step over a feature range [f0, f0 + Fp) of rows of stride F
// (adjoint_edge_row per row of the transposed pattern, dP accumulated in
// place).
template <int VEC, int G, int U, int NC, bool EDGE>
__global__ __launch_bounds__(kBlock) void k_adjoint_edge(
    const int32_t* __restrict__ t_row_ptr, const int32_t* __restrict__ t_col,
    const float* __restrict__ t_val, const int32_t* __restrict__ sched,
    const float* __restrict__ Gs, const float* __restrict__ Xprev,
    float* __restrict__ Gp, float* __restrict__ dPT, int accumulate, int32_t n,
    int32_t F, int32_t f0, int32_t Fp) {
    constexpr int RPW = 64 / G;
    constexpr int RPB = (kBlock / 64) * RPW;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = lane / G, li = lane % G;
    const int rr = L * RPB + wave * RPW + grp;
    if (rr >= n) return;
    int row = sched ? sched[rr] : rr;
    if constexpr (G == 64) row = __builtin_amdgcn_readfirstlane(row);
    adjoint_edge_row<VEC, G, U, NC, EDGE>(t_row_ptr, t_col, t_val, Gs, Xprev, Gp, dPT, dPT,
                                          accumulate, F, f0, Fp, row, li);
}

using AdjFn = void (*)(const int32_t*, const int32_t*, const float*, const int32_t*,
                       const float*, const float*, float*, float*, int, int32_t,
                       int32_t, int32_t, int32_t);

constexpr int kAdjU = 8;

template <int VEC, int G, int NC>
AdjFn pick_edge(bool edge) {
    return edge ? k_adjoint_edge<VEC, G, kAdjU, NC, true>
                : k_adjoint_edge<VEC, G, kAdjU, NC, false>;
}

template <int VEC>
AdjFn pick_adj(int group, int nc, bool edge) {
    if (group < 64) {
        switch (group) {
            case 1: return pick_edge<VEC, 1, 1>(edge);
            case 2: return pick_edge<VEC, 2, 1>(edge);
            case 4: return pick_edge<VEC, 4, 1>(edge);
            case 8: return pick_edge<VEC, 8, 1>(edge);
            case 16: return pick_edge<VEC, 16, 1>(edge);
            default: return pick_edge<VEC, 32, 1>(edge);
        }
    }
    return nc == 1 ? pick_edge<VEC, 64, 1>(edge) : pick_edge<VEC, 64, 2>(edge);
}

// Standalone edge gradient of one step over the forward pattern: for row i,
// G[i, :] is loaded once; every nonzero k = (i, j) gathers Xprev[j, :] and
// writes dP[k] (+)= <G[i, :], Xprev[j, :]> (CSR order).
template <int VEC, int G, int NC>
__global__ __launch_bounds__(kBlock) void k_edge_grad(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ sched, const float* __restrict__ Gs,
    const float* __restrict__ Xprev, float* __restrict__ dP, int accumulate,
    int32_t n, int32_t F, int32_t f0, int32_t Fp) {
    using VT = typename Vec<VEC>::T;
    constexpr int RPW = 64 / G;
    constexpr int RPB = (kBlock / 64) * RPW;
    constexpr int U = 8;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = lane / G, li = lane % G;
    const int rr = L * RPB + wave * RPW + grp;
    if (rr >= n) return;
    int row = sched ? sched[rr] : rr;
    if constexpr (G == 64) row = __builtin_amdgcn_readfirstlane(row);
    const int beg = row_ptr[row], end = row_ptr[row + 1];
    bool act[NC];
    int off[NC];
    VT gi[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int f = c * (G * VEC) + li * VEC;
        act[c] = f < Fp;
        off[c] = act[c] ? f0 + f : 0;
        gi[c] = *reinterpret_cast<const VT*>(Gs + (size_t)row * F + off[c]);
        if (!act[c]) gi[c] = VT{};
    }
    for (int k = beg; k < end; k += U) {
        int j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = col[min(k + u, end - 1)];
        VT x[U][NC];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < NC; ++c)
                x[u][c] = *reinterpret_cast<const VT*>(Xprev + (size_t)j[u] * F + off[c]);
        float d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            float sacc = dotv(gi[0], x[u][0]);
#pragma unroll
            for (int c = 1; c < NC; ++c) sacc += dotv(gi[c], x[u][c]);
            d[u] = group_sum<G>(sacc);
        }
        if (li == 0) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k + u < end) dP[k + u] = accumulate ? dP[k + u] + d[u] : d[u];
        }
    }
}

template <int VEC, int G, int NC>
void edge_launch(const CsrView& P, const float* Gs, const float* Xprev, float* dP,
                 int accumulate, int64_t F, int64_t f0, int64_t Fp, hipStream_t s) {
    const int rpb = (kBlock / 64) * (64 / G);
    const int64_t nb = (P.n + rpb - 1) / rpb;
    hipLaunchKernelGGL((k_edge_grad<VEC, G, NC>), dim3((unsigned)nb), dim3(kBlock), 0, s,
                       P.row_ptr, P.col, P.sched, Gs, Xprev, dP, accumulate, P.n,
                       (int32_t)F, (int32_t)f0, (int32_t)Fp);
}

template <int VEC>
void edge_pick(int group, int nc, const CsrView& P, const float* Gs, const float* Xprev,
               float* dP, int acc, int64_t F, int64_t f0, int64_t Fp, hipStream_t s) {
    switch (group) {
        case 1: edge_launch<VEC, 1, 1>(P, Gs, Xprev, dP, acc, F, f0, Fp, s); return;
        case 2: edge_launch<VEC, 2, 1>(P, Gs, Xprev, dP, acc, F, f0, Fp, s); return;
        case 4: edge_launch<VEC, 4, 1>(P, Gs, Xprev, dP, acc, F, f0, Fp, s); return;
        case 8: edge_launch<VEC, 8, 1>(P, Gs, Xprev, dP, acc, F, f0, Fp, s); return;
        case 16: edge_launch<VEC, 16, 1>(P, Gs, Xprev, dP, acc, F, f0, Fp, s); return;
        case 32: edge_launch<VEC, 32, 1>(P, Gs, Xprev, dP, acc, F, f0, Fp, s); return;
        default: break;
    }
    if (nc == 1) edge_launch<VEC, 64, 1>(P, Gs, Xprev, dP, acc, F, f0, Fp, s);
    else edge_launch<VEC, 64, 2>(P, Gs, Xprev, dP, acc, F, f0, Fp, s);
}

__global__ __launch_bounds__(kBlock) void k_sub(const float* __restrict__ a,
                                                const float* __restrict__ b,
                                                float* __restrict__ c, int64_t m) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < m) c[p] = a[p] - b[p];
}

__global__ __launch_bounds__(kBlock) void k_scatter(const int32_t* __restrict__ idx,
                                                    const float* __restrict__ src,
                                                    float* __restrict__ dst, int64_t m) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < m) dst[idx[p]] = src[p];
}

__global__ __launch_bounds__(kBlock) void k_gather(const int32_t* __restrict__ idx,
                                                   const float* __restrict__ src,
                                                   float* __restrict__ dst, int64_t m) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < m) dst[p] = src[idx[p]];
}

// Rows of an n x K row-major matrix: dst[i, :] = src[idx[i], :] (gather) or
// dst[idx[i], :] = src[i, :] (scatter), one element per thread.
__global__ __launch_bounds__(kBlock) void k_rows(const int32_t* __restrict__ idx,
                                                 const float* __restrict__ src,
                                                 float* __restrict__ dst, int64_t n, int64_t K,
                                                 int scatter) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n * K) return;
    const int64_t i = p / K, c = p - i * K, j = idx[i];
    if (scatter) dst[j * K + c] = src[p];
    else dst[p] = src[j * K + c];
}

__global__ __launch_bounds__(kBlock) void k_gather_pad(const int32_t* __restrict__ idx,
                                                       const float* __restrict__ src,
                                                       float* __restrict__ dst, int64_t m) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < m) {
        const int32_t i = idx[p];
        dst[p] = i >= 0 ? src[i] : -0.0f;
    }
}

}  // namespace

namespace {

int widest_vec(int64_t F, uintptr_t a) {
    return (F % 4 == 0 && a % 16 == 0) ? 4 : (F % 2 == 0 && a % 8 == 0) ? 2 : 1;
}

bool pow2_in(int g, int lo, int hi) { return g >= lo && g <= hi && (g & (g - 1)) == 0; }

}  // namespace

bool adjoint_shape_valid(const SpmmShape& sh, int64_t F, const void* a, const void* b,
                         const void* c) {
    if (sh.vec != 1 && sh.vec != 2 && sh.vec != 4) return false;
    if (!pow2_in(sh.group, 1, 64) || F % sh.vec != 0) return false;
    if (sh.wg_per_cu < 0 || sh.wg_per_cu > kShapeOccupancyMask) return false;
    const auto al = reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) |
                    reinterpret_cast<uintptr_t>(c);
    return al % (4 * (uintptr_t)sh.vec) == 0;
}

std::vector<SpmmShape> adjoint_candidate_shapes(int64_t F, const void* a, const void* b,
                                                const void* c, bool has_schedule) {
    std::vector<SpmmShape> out;
    for (int vec : {4, 2, 1}) {
        const int64_t lanes = (F + vec - 1) / vec;
        int gmax = 1;
        while (gmax < lanes && gmax < 64) gmax <<= 1;
        for (int g = 4; g <= gmax; g <<= 1) {
            if (g < gmax / 4) continue;             // many passes over the pattern
            if (vec < 4 && g != gmax) continue;     // narrow loads only as one slab
            for (int sc = has_schedule ? 1 : 0; sc >= 0; --sc)
                // fewer resident workgroups keep fewer rows' gathers in L2
                for (int occ : {0, 4}) {
                    SpmmShape sh;
                    sh.vec = vec;
                    sh.group = g;
                    sh.schedule = sc != 0;
                    sh.wg_per_cu = occ;
                    if (adjoint_shape_valid(sh, F, a, b, c)) out.push_back(sh);
                }
        }
    }
    return out;
}

void launch_adjoint_edge(const CsrView& PT, const float* Gs, const float* Xprev,
                         float* Gp, float* dPT, bool accumulate, int64_t F,
                         hipStream_t s, const SpmmShape* shape) {
    if (PT.n <= 0 || F <= 0) return;
    const bool edge = Xprev != nullptr && dPT != nullptr;
    int vec, gsel;
    unsigned lds = 0;
    const int32_t* sched = PT.sched;
    if (shape && adjoint_shape_valid(*shape, F, Gs, Gp, Xprev)) {
        vec = shape->vec;
        gsel = shape->group;
        lds = (unsigned)occupancy_lds_bytes(shape->wg_per_cu);
        if (!shape->schedule) sched = nullptr;
    } else {
        const auto a = reinterpret_cast<uintptr_t>(Gs) | reinterpret_cast<uintptr_t>(Gp) |
                       reinterpret_cast<uintptr_t>(Xprev);
        vec = widest_vec(F, a);
        gsel = 64;
    }
    // feature passes of at most 2 chunks of 64 lanes (more chunks per lane
    // cost occupancy: 4 chunks need 180 VGPRs, 8 chunks 256), or of one
    // group-wide slab for narrower groups; the edge term of the later passes
    // accumulates onto the first pass's value
    const int64_t per_pass = gsel < 64 ? (int64_t)gsel * vec : (int64_t)64 * vec * 2;
    for (int64_t f0 = 0; f0 < F; f0 += per_pass) {
        const int64_t Fp = std::min<int64_t>(per_pass, F - f0);
        const int64_t lanes = (Fp + vec - 1) / vec;
        int group = 1;
        while (group < lanes && group < gsel) group <<= 1;
        int nc = 1;
        if (group == 64)
            while ((int64_t)nc * 64 < lanes) nc <<= 1;
        const int rpb = (kBlock / 64) * (64 / group);
        const int64_t nb = (PT.n + rpb - 1) / rpb;
        AdjFn fn = vec == 4 ? pick_adj<4>(group, nc, edge)
                 : vec == 2 ? pick_adj<2>(group, nc, edge)
                            : pick_adj<1>(group, nc, edge);
        hipLaunchKernelGGL(fn, dim3((unsigned)nb), dim3(kBlock), lds, s, PT.row_ptr,
                           PT.col, PT.val, sched, Gs, Xprev, Gp, dPT,
                           (accumulate || f0 > 0) ? 1 : 0, PT.n, (int32_t)F,
                           (int32_t)f0, (int32_t)Fp);
        check_launch("adjoint launch");
    }
}

void launch_scatter_f32(const int32_t* idx, const float* src, float* dst, int64_t m,
                        hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)((m + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, idx, src, dst, m);
    check_launch("scatter launch");
}

void launch_edge_grad(const CsrView& P, const float* Gs, const float* Xprev, float* dP,
                      bool accumulate, int64_t F, hipStream_t s) {
    if (P.n <= 0 || F <= 0) return;
    const auto a = reinterpret_cast<uintptr_t>(Gs) | reinterpret_cast<uintptr_t>(Xprev);
    const int vec = (F % 4 == 0 && a % 16 == 0) ? 4 : (F % 2 == 0 && a % 8 == 0) ? 2 : 1;
    const int64_t per_pass = (int64_t)64 * vec * 2;
    for (int64_t f0 = 0; f0 < F; f0 += per_pass) {
        const int64_t Fp = std::min<int64_t>(per_pass, F - f0);
        const int64_t lanes = (Fp + vec - 1) / vec;
        int group = 1;
        while (group < lanes && group < 64) group <<= 1;
        int nc = 1;
        if (group == 64)
            while ((int64_t)nc * 64 < lanes) nc <<= 1;
        const int acc = (accumulate || f0 > 0) ? 1 : 0;
        if (vec == 4) edge_pick<4>(group, nc, P, Gs, Xprev, dP, acc, F, f0, Fp, s);
        else if (vec == 2) edge_pick<2>(group, nc, P, Gs, Xprev, dP, acc, F, f0, Fp, s);
        else edge_pick<1>(group, nc, P, Gs, Xprev, dP, acc, F, f0, Fp, s);
        check_launch("edge gradient launch");
    }
}

void launch_sub_f32(const float* a, const float* b, float* c, int64_t m, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_sub, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, a, b, c, m);
    check_launch("sub launch");
}

void launch_gather_f32(const int32_t* idx, const float* src, float* dst, int64_t m,
                       hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((m + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, idx, src, dst, m);
    check_launch("gather launch");
}

void launch_rows_f32(const int32_t* idx, const float* src, float* dst, int64_t n, int64_t K,
                     bool scatter, hipStream_t s) {
    if (n <= 0 || K <= 0) return;
    hipLaunchKernelGGL(k_rows, dim3((unsigned)((n * K + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, idx, src, dst, n, K, scatter ? 1 : 0);
    check_launch("row permutation launch");
}

void launch_gather_pad_f32(const int32_t* idx, const float* src, float* dst, int64_t m,
                           hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_gather_pad, dim3((unsigned)((m + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, idx, src, dst, m);
    check_launch("gather launch");
}

}  // namespace scl_amd
