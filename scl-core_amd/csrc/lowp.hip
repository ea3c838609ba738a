// lowp.hip -- reduced-precision storage for the diffusion (gfx950).
//
// Y = P^t X with the intermediate iterates X_1 .. X_{t-1} stored as bf16
// (half the bytes of every gather and of every iterate write), fp32
// accumulation, and fp32 at both ends: the first step gathers the fp32 input
// and the last step writes fp32.  P's values are read either as fp32 or as
// bf16.  Every gathered element is widened to fp32 exactly (bf16 -> fp32 is
// a 16-bit shift), accumulated with fp32 FMAs in stored order, and the row
// result is rounded to bf16 once (round to nearest even) when the step's
// output is an iterate.
//
// Launch layout is the row-split one of kernels.hip: a group of G lanes owns
// a row, each lane E consecutive features (one 4/8/16-byte load per
// nonzero), rows visited in the operator's locality order, workgroups
// remapped so each XCD walks a contiguous slab range.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace scl_amd {

using namespace dev;

namespace {

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}

// round to nearest even; NaN stays NaN (quiet bit kept)
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
    return (uint16_t)((u + r) >> 16);
}

// 16-bit storage formats of the iterates
struct SBf16 {
    __device__ static float widen(uint16_t h) { return bf16_to_f32(h); }
    __device__ static uint16_t narrow(float f) { return f32_to_bf16(f); }
};
struct SFp16 {
    __device__ static float widen(uint16_t h) {
        return (float)__builtin_bit_cast(_Float16, h);
    }
    // round to nearest even (the hardware conversion), overflow -> inf
    __device__ static uint16_t narrow(float f) {
        return __builtin_bit_cast(uint16_t, (_Float16)f);
    }
};

template <int BYTES> struct Raw;
template <> struct Raw<4> { using T = uint32_t; };
template <> struct Raw<8> { using T = uint32_t __attribute__((ext_vector_type(2))); };
template <> struct Raw<16> { using T = uint32_t __attribute__((ext_vector_type(4))); };

// Loads E elements of type IT starting at p and widens them to fp32
// (16-bit elements in storage format ST, multiplied by scale).
template <typename IT, typename ST, int E>
__device__ __forceinline__ void load_wide(const IT* p, float (&x)[E], float scale) {
    constexpr int B = E * (int)sizeof(IT);
    using R = typename Raw<B>::T;
    const R r = *reinterpret_cast<const R*>(p);
    uint32_t w[B / 4];
    if constexpr (B == 4) {
        w[0] = r;
    } else {
#pragma unroll
        for (int q = 0; q < B / 4; ++q) w[q] = r[q];
    }
    if constexpr (sizeof(IT) == 4) {
#pragma unroll
        for (int e = 0; e < E; ++e) x[e] = __uint_as_float(w[e]);
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e)
            x[e] = ST::widen((uint16_t)(w[e / 2] >> (16 * (e % 2)))) * scale;
    }
}

template <typename OT, typename ST, int E>
__device__ __forceinline__ void store_narrow(OT* p, const float (&a)[E], float scale) {
    constexpr int B = E * (int)sizeof(OT);
    using R = typename Raw<B>::T;
    uint32_t w[B / 4];
    if constexpr (sizeof(OT) == 4) {
#pragma unroll
        for (int e = 0; e < E; ++e) w[e] = __float_as_uint(a[e]);
    } else {
#pragma unroll
        for (int q = 0; q < B / 4; ++q)
            w[q] = (uint32_t)ST::narrow(a[2 * q] * scale) |
                   ((uint32_t)ST::narrow(a[2 * q + 1] * scale) << 16);
    }
    R r;
    if constexpr (B == 4) {
        r = w[0];
    } else {
#pragma unroll
        for (int q = 0; q < B / 4; ++q) r[q] = w[q];
    }
    *reinterpret_cast<R*>(p) = r;
}

template <typename VT>
__device__ __forceinline__ float value_of(const VT* v, int k) {
    if constexpr (sizeof(VT) == 4) return v[k];
    else return bf16_to_f32(v[k]);
}

template <typename IT, typename OT, typename VT, typename ST, int E, int G, int U>
__global__ __launch_bounds__(kBlock) void k_spmm_mixed(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const VT* __restrict__ val, const int32_t* __restrict__ sched,
    const IT* __restrict__ X, OT* __restrict__ Y, int32_t n, int32_t F,
    int32_t n_rowblocks, float in_scale, float out_scale) {
    constexpr int RPW = 64 / G;
    constexpr int RPB = (kBlock / 64) * RPW;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int slab = L / n_rowblocks;
    const int rb = L - slab * n_rowblocks;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = lane / G, li = lane % G;
    const int rr = rb * RPB + wave * RPW + grp;
    if (rr >= n) return;
    int row = sched ? sched[rr] : rr;
    if constexpr (G == 64) row = __builtin_amdgcn_readfirstlane(row);
    const int f = slab * (G * E) + li * E;
    const bool active = f < F;
    const IT* Xf = X + (active ? f : 0);
    const int beg = row_ptr[row], end = row_ptr[row + 1];

    float acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = 0.0f;

    int k = beg;
    if (k + U <= end) {
        int j[U];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            j[u] = col[k + u];
            v[u] = value_of(val, k + u);
        }
        while (true) {
            float x[U][E];
#pragma unroll
            for (int u = 0; u < U; ++u) load_wide<IT, ST, E>(Xf + (size_t)j[u] * F, x[u], in_scale);
            const int kn = k + U;
            const bool more = kn + U <= end;
            int jn[U];
            float vn[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = min(kn + u, end - 1);
                jn[u] = col[kk];
                vn[u] = value_of(val, kk);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int e = 0; e < E; ++e) acc[e] = __builtin_fmaf(v[u], x[u][e], acc[e]);
            k = kn;
            if (!more) break;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                j[u] = jn[u];
                v[u] = vn[u];
            }
        }
    }
    if (k < end) {
        int j[U];
        float v[U];
        float x[U][E];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k + u, end - 1);
            j[u] = col[kk];
            v[u] = value_of(val, kk);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) load_wide<IT, ST, E>(Xf + (size_t)j[u] * F, x[u], in_scale);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k + u < end)
#pragma unroll
                for (int e = 0; e < E; ++e) acc[e] = __builtin_fmaf(v[u], x[u][e], acc[e]);
    }
    if (active) store_narrow<OT, ST, E>(Y + (size_t)row * F + f, acc, out_scale);
}

__global__ __launch_bounds__(kBlock) void k_to_bf16(const float* __restrict__ src,
                                                    uint16_t* __restrict__ dst, int64_t m) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < m) dst[p] = f32_to_bf16(src[p]);
}

// dst = narrow(src * scale), 4 elements per lane (m % 4 == 0, 16-byte
// aligned src, 8-byte aligned dst) or one per lane.
template <typename ST, int E>
__global__ __launch_bounds__(kBlock) void k_narrow(const float* __restrict__ src,
                                                   uint16_t* __restrict__ dst, int64_t m,
                                                   float scale) {
    const int64_t p = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * E;
    if (p >= m) return;
    if constexpr (E == 4) {
        const f4 x = *reinterpret_cast<const f4*>(src + p);
        const uint32_t lo = (uint32_t)ST::narrow(x.x * scale) |
                            ((uint32_t)ST::narrow(x.y * scale) << 16);
        const uint32_t hi = (uint32_t)ST::narrow(x.z * scale) |
                            ((uint32_t)ST::narrow(x.w * scale) << 16);
        *reinterpret_cast<uint2*>(dst + p) = make_uint2(lo, hi);
    } else {
        dst[p] = ST::narrow(src[p] * scale);
    }
}

__global__ __launch_bounds__(kBlock) void k_from_bf16(const uint16_t* __restrict__ src,
                                                      float* __restrict__ dst, int64_t m) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < m) dst[p] = bf16_to_f32(src[p]);
}

// --------------------------------------------------- compensated bf16 --
// A value q is rounded once to 24 bits (sign, 8 exponent bits, 15 mantissa
// bits: round to nearest even at bit 8 of the fp32 word; nan stays a quiet
// nan) and stored as two planes: hi, the upper 16 bits (a bf16), and lo, the
// next 8 bits, which carry the mantissa bits bf16 drops.  Read back as the
// fp32 word hi << 16 | lo << 8: exact, one byte permute per value, so a
// gathered value costs what a bf16 one does in VALU work and 3 bytes
// instead of 2 in memory.
__device__ __forceinline__ float comp_decode(uint16_t h, uint8_t l) {
    return __uint_as_float(((uint32_t)h << 16) | ((uint32_t)l << 8));
}

__device__ __forceinline__ void comp_encode(float q, uint16_t& h, uint8_t& l) {
    const uint32_t u = __float_as_uint(q);
    const uint32_t w = (u & 0x7fffffffu) > 0x7f800000u
                           ? (u >> 8) | 0x4000u
                           : (u + 0x7fu + ((u >> 8) & 1u)) >> 8;
    h = (uint16_t)(w >> 8);
    l = (uint8_t)(w & 0xffu);
}

// Iterates in compensated form are packed in groups of 4 consecutive
// values (row-major): 12 bytes per group, [hi0 hi1 | hi2 hi3 | lo0..lo3], so
// one 12-byte load brings 4 values, as one 8-byte load does for bf16.
__device__ __forceinline__ void unpack4(const uint3 g, float (&x)[4]) {
    x[0] = comp_decode((uint16_t)g.x, (uint8_t)g.z);
    x[1] = comp_decode((uint16_t)(g.x >> 16), (uint8_t)(g.z >> 8));
    x[2] = comp_decode((uint16_t)g.y, (uint8_t)(g.z >> 16));
    x[3] = comp_decode((uint16_t)(g.y >> 16), (uint8_t)(g.z >> 24));
}

__device__ __forceinline__ uint3 pack4(const float (&a)[4]) {
    uint16_t h[4];
    uint8_t l[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) comp_encode(a[e], h[e], l[e]);
    return make_uint3((uint32_t)h[0] | ((uint32_t)h[1] << 16),
                      (uint32_t)h[2] | ((uint32_t)h[3] << 16),
                      (uint32_t)l[0] | ((uint32_t)l[1] << 8) | ((uint32_t)l[2] << 16) |
                          ((uint32_t)l[3] << 24));
}

// One step with compensated input and/or output and/or values; the row
// chain is the row-split one (fp32 FMAs in stored order) over the decoded
// values.  G lanes per row, 4 values per lane (G = 64: one row per wave,
// the row's indices on the scalar unit), the (col, val) pairs of the next
// chunk fetched while the current chunk's gathers are in flight.
template <bool IN_C, bool OUT_C, bool VAL_C, int G, int U>
__global__ __launch_bounds__(kBlock) void k_spmm_comp(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, const uint16_t* __restrict__ vh,
    const uint8_t* __restrict__ vl, const int32_t* __restrict__ sched,
    const float* __restrict__ X, const uint3* __restrict__ Xc, float* __restrict__ Y,
    uint3* __restrict__ Yc, int32_t n, int32_t F, int32_t n_rowblocks) {
    constexpr int RPW = 64 / G;
    constexpr int RPB = (kBlock / 64) * RPW;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int slab = L / n_rowblocks;
    const int rb = L - slab * n_rowblocks;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = lane / G, li = lane % G;
    const int rr = rb * RPB + wave * RPW + grp;
    if (rr >= n) return;
    int row = sched ? sched[rr] : rr;
    if constexpr (G == 64) row = __builtin_amdgcn_readfirstlane(row);
    const int f = slab * (G * 4) + li * 4;
    const bool active = f < F;
    const int fo = active ? f : 0;  // lanes past F read feature 0, store nothing
    const int beg = row_ptr[row], end = row_ptr[row + 1];
    const int F4 = F / 4;
    auto value = [&](int k) -> float {
        if constexpr (VAL_C) return comp_decode(vh[k], vl[k]);
        else return val[k];
    };
    auto gather = [&](int j, float (&x)[4]) {
        if constexpr (IN_C) {
            unpack4(Xc[(size_t)j * F4 + fo / 4], x);
        } else {
            const f4 v = *reinterpret_cast<const f4*>(X + (size_t)j * F + fo);
            x[0] = v.x;
            x[1] = v.y;
            x[2] = v.z;
            x[3] = v.w;
        }
    };
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (beg < end) {
        int j[U];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(beg + u, end - 1);
            j[u] = col[kk];
            v[u] = value(kk);
        }
        for (int k = beg;; k += U) {
            float x[U][4];
#pragma unroll
            for (int u = 0; u < U; ++u) gather(j[u], x[u]);
            const int kn = k + U;
            int jn[U];
            float vn[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = min(kn + u, end - 1);
                jn[u] = col[kk];
                vn[u] = value(kk);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k + u < end)
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[e] = __builtin_fmaf(v[u], x[u][e], acc[e]);
            if (kn >= end) break;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                j[u] = jn[u];
                v[u] = vn[u];
            }
        }
    }
    if (!active) return;
    if constexpr (OUT_C) Yc[(size_t)row * F4 + f / 4] = pack4(acc);
    else *reinterpret_cast<f4*>(Y + (size_t)row * F + f) = f4{acc[0], acc[1], acc[2], acc[3]};
}

// values: one thread each, two planes
__global__ __launch_bounds__(kBlock) void k_encode_comp(const float* __restrict__ src,
                                                        uint16_t* __restrict__ h,
                                                        uint8_t* __restrict__ l, int64_t m) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    uint16_t hh;
    uint8_t ll;
    comp_encode(src[p], hh, ll);
    h[p] = hh;
    l[p] = ll;
}

// iterates: one group of 4 per thread (m % 4 == 0, src 16-byte aligned)
__global__ __launch_bounds__(kBlock) void k_pack_comp(const float* __restrict__ src,
                                                      uint3* __restrict__ dst, int64_t groups) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= groups) return;
    const f4 v = *reinterpret_cast<const f4*>(src + 4 * g);
    const float a[4] = {v.x, v.y, v.z, v.w};
    dst[g] = pack4(a);
}

template <bool IN_C, bool OUT_C, bool VAL_C>
void launch_comp_io(const CsrView& P, const CompPlanes& vals, const float* X,
                    const uint32_t* Xc, float* Y, uint32_t* Yc, int64_t F, bool use_sched,
                    hipStream_t s) {
    int G = 1;
    while (G < (F + 3) / 4 && G < 64) G <<= 1;
    G = std::max(G, 16);
    const int rpb = (kBlock / 64) * (64 / G);
    const int64_t n_rowblocks = (P.n + rpb - 1) / rpb;
    const int64_t n_slabs = (F + (int64_t)G * 4 - 1) / ((int64_t)G * 4);
    const int64_t nb = n_rowblocks * n_slabs;
    if (nb > INT32_MAX) throw std::runtime_error("compensated spmm grid too large");
    const int32_t* sched = use_sched ? P.sched : nullptr;
#define SCL_COMP(GG)                                                                         \
    hipLaunchKernelGGL((k_spmm_comp<IN_C, OUT_C, VAL_C, GG, 8>), dim3((unsigned)nb),         \
                       dim3(kBlock), 0, s, P.row_ptr, P.col, P.val, vals.hi, vals.lo, sched,  \
                       X, reinterpret_cast<const uint3*>(Xc), Y, reinterpret_cast<uint3*>(Yc), \
                       P.n, (int32_t)F, (int32_t)n_rowblocks)
    switch (G) {
        case 16: SCL_COMP(16); break;
        case 32: SCL_COMP(32); break;
        default: SCL_COMP(64); break;
    }
#undef SCL_COMP
}

constexpr int kLowpU = 8;

template <typename IT, typename OT, typename VT, typename ST>
void launch_mixed(const CsrView& P, const VT* vals, const IT* X, OT* Y, int64_t F,
                  bool use_sched, float in_scale, float out_scale, hipStream_t s) {
    // elements per lane: one 16-byte load when the row is wide enough for
    // 64 lanes, else narrower loads that still give every lane of the wave
    // its own features
    constexpr int wmax = sizeof(IT) > sizeof(OT) ? (int)sizeof(IT) : (int)sizeof(OT);
    const int e_max = 16 / wmax;
    const uintptr_t a = reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y);
    int E = e_max;
    while (E > 2 && (F % E != 0 || (a % (E * sizeof(IT))) != 0 ||
                     (a % (E * sizeof(OT))) != 0 || (F / E) < 64))
        E /= 2;
    if (F % E != 0 || a % (E * sizeof(IT)) != 0 || a % (E * sizeof(OT)) != 0)
        throw std::runtime_error("reduced-precision SpMM needs an even feature count "
                                 "and aligned buffers");
    int G = 1;
    while (G < (F + E - 1) / E && G < 64) G <<= 1;
    G = std::max(G, 16);
    const int rpb = (kBlock / 64) * (64 / G);
    const int64_t n_rowblocks = (P.n + rpb - 1) / rpb;
    const int64_t n_slabs = (F + (int64_t)G * E - 1) / ((int64_t)G * E);
    const int64_t nb = n_rowblocks * n_slabs;
    const int32_t* sched = use_sched ? P.sched : nullptr;
#define SCL_LAUNCH_MIXED(EE, GG)                                                       \
    hipLaunchKernelGGL((k_spmm_mixed<IT, OT, VT, ST, EE, GG, kLowpU>), dim3((unsigned)nb), \
                       dim3(kBlock), 0, s, P.row_ptr, P.col, vals, sched, X, Y, P.n,       \
                       (int32_t)F, (int32_t)n_rowblocks, in_scale, out_scale)
#define SCL_LAUNCH_MIXED_G(EE)                       \
    switch (G) {                                     \
        case 16: SCL_LAUNCH_MIXED(EE, 16); break;    \
        case 32: SCL_LAUNCH_MIXED(EE, 32); break;    \
        default: SCL_LAUNCH_MIXED(EE, 64); break;    \
    }
    if (E == 8) {
        if constexpr (16 / wmax >= 8) {
            SCL_LAUNCH_MIXED_G(8);
        }
    } else if (E == 4) {
        SCL_LAUNCH_MIXED_G(4);
    } else {
        SCL_LAUNCH_MIXED_G(2);
    }
#undef SCL_LAUNCH_MIXED_G
#undef SCL_LAUNCH_MIXED
    check_launch("bf16 spmm launch");
}

}  // namespace

void launch_to_bf16(const float* src, uint16_t* dst, int64_t m, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, src, dst, m);
    check_launch("bf16 convert launch");
}

void launch_narrow_f32(const float* src, uint16_t* dst, int64_t m, bool fp16, float scale,
                       hipStream_t s) {
    if (m <= 0) return;
    const bool wide = m % 4 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(dst) % 8 == 0;
    const int E = wide ? 4 : 1;
    const int64_t lanes = (m + E - 1) / E;
    const dim3 grid((unsigned)((lanes + kBlock - 1) / kBlock));
    if (fp16) {
        if (wide) hipLaunchKernelGGL((k_narrow<SFp16, 4>), grid, dim3(kBlock), 0, s, src, dst, m, scale);
        else hipLaunchKernelGGL((k_narrow<SFp16, 1>), grid, dim3(kBlock), 0, s, src, dst, m, scale);
    } else {
        if (wide) hipLaunchKernelGGL((k_narrow<SBf16, 4>), grid, dim3(kBlock), 0, s, src, dst, m, scale);
        else hipLaunchKernelGGL((k_narrow<SBf16, 1>), grid, dim3(kBlock), 0, s, src, dst, m, scale);
    }
    check_launch("16-bit convert launch");
}

void launch_from_bf16(const uint16_t* src, float* dst, int64_t m, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_from_bf16, dim3((unsigned)((m + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, src, dst, m);
    check_launch("bf16 convert launch");
}

template <typename ST>
void launch_lowp_fmt(const CsrView& P, const uint16_t* vb, LowpIo io, const void* X,
                     void* Y, int64_t F, bool sched, float in_scale, float out_scale,
                     hipStream_t s) {
    using U16 = uint16_t;
    switch (io) {
        case LowpIo::F32_TO_BF16:
            if (vb) launch_mixed<float, U16, U16, ST>(P, vb, (const float*)X, (U16*)Y, F, sched, in_scale, out_scale, s);
            else launch_mixed<float, U16, float, ST>(P, P.val, (const float*)X, (U16*)Y, F, sched, in_scale, out_scale, s);
            break;
        case LowpIo::BF16_TO_BF16:
            if (vb) launch_mixed<U16, U16, U16, ST>(P, vb, (const U16*)X, (U16*)Y, F, sched, in_scale, out_scale, s);
            else launch_mixed<U16, U16, float, ST>(P, P.val, (const U16*)X, (U16*)Y, F, sched, in_scale, out_scale, s);
            break;
        case LowpIo::BF16_TO_F32:
            if (vb) launch_mixed<U16, float, U16, ST>(P, vb, (const U16*)X, (float*)Y, F, sched, in_scale, out_scale, s);
            else launch_mixed<U16, float, float, ST>(P, P.val, (const U16*)X, (float*)Y, F, sched, in_scale, out_scale, s);
            break;
        case LowpIo::F32_TO_F32:
            if (vb) launch_mixed<float, float, U16, ST>(P, vb, (const float*)X, (float*)Y, F, sched, in_scale, out_scale, s);
            else launch_mixed<float, float, float, ST>(P, P.val, (const float*)X, (float*)Y, F, sched, in_scale, out_scale, s);
            break;
    }
}

void launch_spmm_lowp(const CsrView& P, const uint16_t* vals_bf16, LowpIo io,
                      const void* X, void* Y, int64_t F, bool use_sched, bool fp16,
                      float in_scale, float out_scale, hipStream_t s) {
    if (P.n <= 0 || F <= 0) return;
    if (fp16)
        launch_lowp_fmt<SFp16>(P, vals_bf16, io, X, Y, F, use_sched, in_scale, out_scale, s);
    else
        launch_lowp_fmt<SBf16>(P, vals_bf16, io, X, Y, F, use_sched, in_scale, out_scale, s);
}

void launch_encode_comp(const float* src, const CompPlanes& dst, int64_t m, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_encode_comp, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, src, dst.hi, dst.lo, m);
    check_launch("compensated encode launch");
}

void launch_pack_comp(const float* src, uint32_t* dst, int64_t m, hipStream_t s) {
    if (m <= 0) return;
    if (m % 4 != 0 || reinterpret_cast<uintptr_t>(src) % 16 != 0)
        throw std::runtime_error("compensated packing needs groups of 4 aligned values");
    const int64_t groups = m / 4;
    hipLaunchKernelGGL(k_pack_comp, dim3((unsigned)((groups + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, src, reinterpret_cast<uint3*>(dst), groups);
    check_launch("compensated pack launch");
}

void launch_spmm_comp(const CsrView& P, const CompPlanes* vals, const float* X,
                      const uint32_t* Xc, float* Y, uint32_t* Yc, int64_t F, bool use_sched,
                      hipStream_t s) {
    if (P.n <= 0 || F <= 0) return;
    if (F % 4 != 0) throw std::runtime_error("compensated storage needs n_features % 4 == 0");
    const CompPlanes none{};
    const CompPlanes& v = vals ? *vals : none;
    const int code = (Xc ? 4 : 0) | (Yc ? 2 : 0) | (vals ? 1 : 0);
    switch (code) {
        case 0: launch_comp_io<false, false, false>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
        case 1: launch_comp_io<false, false, true>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
        case 2: launch_comp_io<false, true, false>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
        case 3: launch_comp_io<false, true, true>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
        case 4: launch_comp_io<true, false, false>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
        case 5: launch_comp_io<true, false, true>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
        case 6: launch_comp_io<true, true, false>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
        default: launch_comp_io<true, true, true>(P, v, X, Xc, Y, Yc, F, use_sched, s); break;
    }
    check_launch("compensated spmm launch");
}

namespace {

__global__ __launch_bounds__(kBlock) void k_absmax_partial(const float* __restrict__ x,
                                                           int64_t m,
                                                           float* __restrict__ partial) {
    __shared__ float sh[kBlock / 64];
    float v = 0.0f;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
        // four 16-byte loads in flight per thread, then the tail
        const f4* x4 = reinterpret_cast<const f4*>(x);
        const int64_t m4 = m / 4;
        for (; p + 3 * stride < m4; p += 4 * stride) {
            f4 a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = x4[p + u * stride];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                v = fmaxf(v, fmaxf(fmaxf(fabsf(a[u].x), fabsf(a[u].y)),
                                   fmaxf(fabsf(a[u].z), fabsf(a[u].w))));
        }
        for (; p < m4; p += stride) {
            const f4 a = x4[p];
            v = fmaxf(v, fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))));
        }
        p = m4 * 4 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    }
    for (; p < m; p += stride) v = fmaxf(v, fabsf(x[p]));
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = sh[0];
        for (int q = 1; q < kBlock / 64; ++q) t = fmaxf(t, sh[q]);
        partial[blockIdx.x] = t;
    }
}

}  // namespace

void launch_absmax(const float* x, int64_t m, float* partial, int n_partial, hipStream_t s) {
    hipLaunchKernelGGL(k_absmax_partial, dim3(n_partial), dim3(kBlock), 0, s, x, m, partial);
    check_launch("absmax launch");
}

}  // namespace scl_amd
