// adjoint_row.hpp -- the fused adjoint + edge-gradient step of one row of
// the transposed pattern, shared by the per-step kernel (grad.hip) and the
// persistent batched kernel (persist.hip), so both produce the same bits.
//
// For transposed row j (a column of P; entries = source rows i in
// increasing order, values p_ij) and a feature range [f0, f0 + Fp):
//   Gp[j, :]  = sum_i p_ij Gs[i, :]                 (fp32 FMA, stored order)
//   ddst[p]   = (accumulate ? dsrc[p] : 0) + <Gs[i, :], Xprev[j, :]>
// (with SCAT, ddst[scat[p]]: the value lands at the entry's CSR position)
// One gather of Gs[i, slab] feeds both; Xprev[j, :] is loaded once.  The U
// partial dot products of a chunk are reduce-scattered across the row's G
// lanes, and each writer lane stores one entry.
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace scl_amd {
namespace dev {

__device__ __forceinline__ float dot4(f4 a, f4 b) {
    return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}
__device__ __forceinline__ float dotv(f4 a, f4 b) { return dot4(a, b); }
__device__ __forceinline__ float dotv(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float dotv(float a, float b) { return a * b; }

// s + <a, b> as an explicit FMA chain over the components: the rounding of
// a lane's partial dot product is then fixed by the source, whatever the
// compiler would contract in a * b + c elsewhere, so kernels that share
// this row code with other chunk lengths produce the same bits.
__device__ __forceinline__ float fma_dot(f4 a, f4 b, float s) {
    return fmaf(a.w, b.w, fmaf(a.z, b.z, fmaf(a.y, b.y, fmaf(a.x, b.x, s))));
}
__device__ __forceinline__ float fma_dot(f2 a, f2 b, float s) {
    return fmaf(a.y, b.y, fmaf(a.x, b.x, s));
}
__device__ __forceinline__ float fma_dot(float a, float b, float s) { return fmaf(a, b, s); }

// Cross-lane sums.  Inside a 16-lane row, stage b pairs lane li with
// li ^ (2^(b+1) - 1) through DPP (quad_perm, row_half_mirror, row_mirror),
// which costs no LDS traffic; lanes 16 and 32 apart are paired with
// __shfl_xor.  Successive stages use linearly independent masks, so a chain
// over all of a group's bits sums every lane of the group exactly once.
template <int B>
__device__ __forceinline__ float dpp_partner(float v) {
    static_assert(B >= 0 && B < 4, "DPP stages cover bits 0..3");
    constexpr int ctrl = B == 0 ? 0xB1 : B == 1 ? 0x1B : B == 2 ? 0x141 : 0x140;
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, 0xF, 0xF,
                                           false));
}

// Lanes with bit B clear end with a(self) + a(partner), lanes with bit B set
// with b(self) + b(partner).
template <int B>
__device__ __forceinline__ float pair_reduce(float a, float b, bool hi) {
    const float send = hi ? a : b;
    const float keep = hi ? b : a;
    return keep + dpp_partner<B>(send);
}

// Sum over bits B-1 .. 0 of the lane index (B <= 4).
template <int B>
__device__ __forceinline__ float sum_bits_below(float v) {
    if constexpr (B > 0) return sum_bits_below<B - 1>(v + dpp_partner<B - 1>(v));
    else return v;
}

template <int G>
__device__ __forceinline__ constexpr int log2i() {
    return G <= 1 ? 0 : 1 + log2i<G / 2>();
}

// Sum over lanes 16 and 32 apart, for groups wider than one row.
template <int G>
__device__ __forceinline__ float sum_across_rows(float v) {
    if constexpr (G >= 32) v += __shfl_xor(v, 16, 64);
    if constexpr (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// Sum over the G lanes of a group (G a power of two, groups aligned).
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    constexpr int rb = log2i<G>() < 4 ? log2i<G>() : 4;
    return sum_across_rows<G>(sum_bits_below<rb>(v));
}

template <int B, int C, int U>
__device__ __forceinline__ void scatter_stages(float (&v)[U], int li) {
    if constexpr (C > 1) {
        const bool hi = ((li >> B) & 1) != 0;
#pragma unroll
        for (int q = 0; q < C / 2; ++q) v[q] = pair_reduce<B>(v[q], v[q + C / 2], hi);
        scatter_stages<B - 1, C / 2, U>(v, li);
    }
}

// Reduce-scatter of U partial dot products across a group of G lanes
// (G >= U, both powers of two).  Within a row, the group's top log2(U) row
// bits each halve the set of entries a lane carries; the remaining low bits
// and the cross-row bits are then summed outright.  Lane li ends with the
// total of entry scatter_entry(li).
template <int G, int U>
struct Scatter {
    static constexpr int kRowBits = log2i<G>() < 4 ? log2i<G>() : 4;
    static constexpr int kLow = kRowBits - log2i<U>();  // bits summed outright
    static_assert(kLow >= 0, "reduce-scatter needs U <= min(G, 16)");
    __device__ static int entry(int li) { return (li >> kLow) & (U - 1); }
    __device__ static bool writer(int li) {
        return (li & ((1 << kLow) - 1)) == 0 && (G <= 16 || (li >> 4) == 0);
    }
    __device__ static float reduce(float (&d)[U], int li) {
        float v[U];
#pragma unroll
        for (int q = 0; q < U; ++q) v[q] = d[q];
        scatter_stages<kRowBits - 1, U, U>(v, li);
        return sum_across_rows<G>(sum_bits_below<kLow>(v[0]));
    }
};

// One transposed row.  NC chunks of G*VEC features per lane group cover the
// range; li is the lane's index in its group.  The source rows and weights
// of the next U entries, and the running values of the current ones, are
// fetched while the current gathers are in flight.  dsrc may equal ddst
// (in-place accumulation).  beg_in / end_in, when beg_in >= 0, are the
// row's bounds t_row_ptr[row], t_row_ptr[row + 1] loaded by the caller.
// PADDED: a chunk's indices load without clamping to the row (t_col and
// t_val are padded past their end), so its gathers past the row's end read
// other rows of Gs; a kernel whose launch writes rows of Gs that it may
// also read (the persistent batch) clamps instead, so that no wave reads a
// row before it is written.
template <int VEC, int G, int U, int NC, bool EDGE, bool SCAT = false, bool PADDED = true>
__device__ __forceinline__ void adjoint_edge_row(
    const int32_t* __restrict__ t_row_ptr, const int32_t* __restrict__ t_col,
    const float* __restrict__ t_val, const float* __restrict__ Gs,
    const float* __restrict__ Xprev, float* __restrict__ Gp, const float* dsrc, float* ddst,
    int accumulate, int32_t F, int32_t f0, int32_t Fp, int row, int li,
    const int32_t* __restrict__ scat = nullptr, int beg_in = -1, int end_in = -1) {
    using VT = typename Vec<VEC>::T;
    // the row's bounds, unless the caller loaded them already
    const int beg = beg_in >= 0 ? beg_in : t_row_ptr[row];
    const int end = beg_in >= 0 ? end_in : t_row_ptr[row + 1];

    bool act[NC];
    int off[NC];  // feature offset of chunk c for this lane (0 when inactive)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int f = c * (G * VEC) + li * VEC;
        act[c] = f < Fp;
        off[c] = act[c] ? f0 + f : 0;
    }
    VT xj[NC];
    if constexpr (EDGE) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            xj[c] = *reinterpret_cast<const VT*>(Xprev + (size_t)row * F + off[c]);
            if (!act[c]) xj[c] = VT{};
        }
    }
    float acc[NC][VEC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[c][e] = 0.0f;

    // entry of the U-chunk this lane writes, and whether it writes at all
    constexpr bool kScatter = G >= U;
    using SC = Scatter<kScatter ? G : U, U>;
    const int my_e = kScatter ? SC::entry(li) : 0;
    const bool writer = kScatter ? SC::writer(li) : li == 0;

    // PADDED: t_col and t_val hold kTransposePad (>= U) entries past the
    // last row, so a chunk's indices and weights load without clamping
    // (entries past the row's end name valid rows and are not added)
    static_assert(U <= kTransposePad, "the transposed arrays' padding covers a chunk");
    auto at = [&](int kk) { return PADDED ? kk : min(kk, end - 1); };
    int i[U];
    float w[U];
    if (beg < end) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            i[u] = t_col[at(beg + u)];
            w[u] = t_val[at(beg + u)];
        }
    }
    for (int k = beg; k < end; k += U) {
        VT g[U][NC];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < NC; ++c)
                g[u][c] = *reinterpret_cast<const VT*>(Gs + (size_t)i[u] * F + off[c]);
        // running dP values of this chunk (kScatter: one per writer lane)
        float old[kScatter ? 1 : U] = {};
        if constexpr (EDGE) {
            if (accumulate && writer) {
                if constexpr (kScatter) {
                    old[0] = dsrc[min(k + my_e, end - 1)];
                } else {
#pragma unroll
                    for (int u = 0; u < U; ++u) old[u] = dsrc[min(k + u, end - 1)];
                }
            }
        }
        float wc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) wc[u] = w[u];
        // next chunk's indices (unused past the row's end)
        const int kn = k + U;
        if (kn < end) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                i[u] = t_col[at(kn + u)];
                w[u] = t_val[at(kn + u)];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k + u < end) {
#pragma unroll
                for (int c = 0; c < NC; ++c) fma_vec(wc[u], g[u][c], acc[c]);
            }
        }
        if constexpr (EDGE) {
            float d[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                float sdot = 0.0f;
#pragma unroll
                for (int c = 0; c < NC; ++c) sdot = fma_dot(g[u][c], xj[c], sdot);
                d[u] = sdot;
            }
            if constexpr (kScatter) {
                const float r = SC::reduce(d, li);
                if (writer && k + my_e < end) {
                    const int at = SCAT ? scat[k + my_e] : k + my_e;
                    ddst[at] = accumulate ? old[0] + r : r;
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) d[u] = group_sum<G>(d[u]);
                if (writer) {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (k + u < end) {
                            const int at = SCAT ? scat[k + u] : k + u;
                            ddst[at] = accumulate ? old[u] + d[u] : d[u];
                        }
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
        if (act[c]) store_vec<VEC, false>(Gp + (size_t)row * F + off[c], acc[c]);
}

}  // namespace dev
}  // namespace scl_amd
