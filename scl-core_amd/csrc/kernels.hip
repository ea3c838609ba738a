// kernels.hip -- gfx950 kernels of the diffusion path.
//
// Row-split CSR SpMM (one step Y = P X):
//   * a group of G lanes (G a power of two, 1..64) owns one output row and a
//     slab of G*VEC consecutive features; a 256-thread workgroup (4 wave64s)
//     owns 4 * 64/G rows of one slab;
//   * every lane walks its row's nonzeros in stored order and accumulates
//     acc = fma(p_ik, X[j_k, f], acc) in fp32 registers, so each output
//     element is the same FMA chain as the reference's axpy loop
//     (scl/kernel/diffusion.hpp:249-256 over axpy_simd 100-134);
//   * X rows are gathered with VEC-wide (up to 16 B) loads, U nonzeros in
//     flight per lane; for G = 64 the row is wave-uniform and its column
//     indices and values come through the scalar unit;
//   * workgroups are numbered slab-major (all rows of slab 0 first) and the
//     hardware workgroup id is remapped so each XCD walks one contiguous run
//     of rows (workgroups are dealt to the 8 XCDs round-robin), keeping the
//     rows gathered by one XCD inside its own L2.
// No MFMA: the contraction is sparse.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <stdexcept>

namespace scl_amd {

using namespace dev;

namespace {

// Workgroup -> (feature slab, row block).  rb_per_xcd > 0: XCD x owns row
// blocks [x * rb_per_xcd, (x+1) * rb_per_xcd); with rb_per_block == 0 it
// walks them slab by slab (all XCDs on the same slab at a time), with
// rb_per_block > 0 it takes them rb_per_block at a time and runs every slab
// of those before the next ones, so the X rows the block gathers stay in
// the XCD's L2 across its rows; else the logical order is slab-major and
// each XCD owns a contiguous run of it.
__device__ __forceinline__ bool block_coords(int n_rowblocks, int rb_per_xcd, int rb_per_block,
                                             int& slab, int& rb) {
    if (rb_per_block > 0) {
        const int x = blockIdx.x % kNumXcd, idx = blockIdx.x / kNumXcd;
        const int n_blk = (rb_per_xcd + rb_per_block - 1) / rb_per_block;
        const int n_slabs = (int)(gridDim.x / kNumXcd) / (n_blk * rb_per_block);
        const int per = rb_per_block * n_slabs;
        const int blk = idx / per, rem = idx - blk * per;
        slab = rem / rb_per_block;
        const int loc = blk * rb_per_block + (rem - slab * rb_per_block);
        rb = x * rb_per_xcd + loc;
        return loc < rb_per_xcd && rb < n_rowblocks;
    }
    if (rb_per_xcd > 0) {
        const int x = blockIdx.x % kNumXcd, idx = blockIdx.x / kNumXcd;
        slab = idx / rb_per_xcd;
        rb = x * rb_per_xcd + (idx - slab * rb_per_xcd);
        return rb < n_rowblocks;
    }
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    slab = L / n_rowblocks;
    rb = L - slab * n_rowblocks;
    return true;
}

// F = 1: y = P x with one lane per row (rows in the visiting order).  A
// lane reads its row's indices and values with 16-byte loads once the
// position is a multiple of four (col and val 16-byte aligned), sixteen
// (then eight, then four) at a time, and gathers their x values together;
// the products are added in stored order, one FMA each, the row-split
// kernel's chain.
__global__ __launch_bounds__(kBlock) void k_spmv_lane(const int32_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col,
                                                      const float* __restrict__ val,
                                                      const int32_t* __restrict__ sched,
                                                      const float* __restrict__ x,
                                                      float* __restrict__ y, int32_t n) {
    const int rr = xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (rr >= n) return;
    const int row = sched ? sched[rr] : rr;
    int k = row_ptr[row];
    const int e = row_ptr[row + 1];
    float acc = 0.0f;
    for (; k < e && (k & 3); ++k) acc = __builtin_fmaf(val[k], x[col[k]], acc);
    // sixteen at a time: the four groups' x values are gathered together
    for (; k + 16 <= e; k += 16) {
        int4 c[4];
        f4 v[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            c[g] = *reinterpret_cast<const int4*>(col + k + 4 * g);
            v[g] = *reinterpret_cast<const f4*>(val + k + 4 * g);
        }
        float xs[16];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            xs[4 * g + 0] = x[c[g].x];
            xs[4 * g + 1] = x[c[g].y];
            xs[4 * g + 2] = x[c[g].z];
            xs[4 * g + 3] = x[c[g].w];
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            acc = __builtin_fmaf(v[g].x, xs[4 * g + 0], acc);
            acc = __builtin_fmaf(v[g].y, xs[4 * g + 1], acc);
            acc = __builtin_fmaf(v[g].z, xs[4 * g + 2], acc);
            acc = __builtin_fmaf(v[g].w, xs[4 * g + 3], acc);
        }
    }
    // eight at a time: both groups' x values are gathered together
    for (; k + 8 <= e; k += 8) {
        const int4 c0 = *reinterpret_cast<const int4*>(col + k);
        const int4 c1 = *reinterpret_cast<const int4*>(col + k + 4);
        const f4 v0 = *reinterpret_cast<const f4*>(val + k);
        const f4 v1 = *reinterpret_cast<const f4*>(val + k + 4);
        const float x0 = x[c0.x], x1 = x[c0.y], x2 = x[c0.z], x3 = x[c0.w];
        const float x4 = x[c1.x], x5 = x[c1.y], x6 = x[c1.z], x7 = x[c1.w];
        acc = __builtin_fmaf(v0.x, x0, acc);
        acc = __builtin_fmaf(v0.y, x1, acc);
        acc = __builtin_fmaf(v0.z, x2, acc);
        acc = __builtin_fmaf(v0.w, x3, acc);
        acc = __builtin_fmaf(v1.x, x4, acc);
        acc = __builtin_fmaf(v1.y, x5, acc);
        acc = __builtin_fmaf(v1.z, x6, acc);
        acc = __builtin_fmaf(v1.w, x7, acc);
    }
    if (k + 4 <= e) {
        const int4 c = *reinterpret_cast<const int4*>(col + k);
        const f4 v = *reinterpret_cast<const f4*>(val + k);
        const float x0 = x[c.x], x1 = x[c.y], x2 = x[c.z], x3 = x[c.w];
        acc = __builtin_fmaf(v.x, x0, acc);
        acc = __builtin_fmaf(v.y, x1, acc);
        acc = __builtin_fmaf(v.z, x2, acc);
        acc = __builtin_fmaf(v.w, x3, acc);
        k += 4;
    }
    for (; k < e; ++k) acc = __builtin_fmaf(val[k], x[col[k]], acc);
    y[row] = acc;
}

// F = 1: y = P x with one lane per row and 64 consecutive rows per wave.
// The wave stages its rows' entries in windows of C: coalesced loads of
// indices and values, the x values gathered for them, and (value, x) pairs
// written to its own LDS region.  Each lane then adds its row's products in
// stored order from 0, one FMA each (the row-split kernel's chain), reading
// the pairs eight at a time.  Waves never share LDS, so a wave-level fence
// orders the writes before the reads.  With out_row set, row r is stored at
// y[out_row[r]] (a renumbered operator writing the original numbering).
template <int C>
__global__ __launch_bounds__(kBlock) void k_spmv_staged(const int32_t* __restrict__ row_ptr,
                                                        const int32_t* __restrict__ col,
                                                        const float* __restrict__ val,
                                                        const float* __restrict__ x,
                                                        float* __restrict__ y, int32_t n,
                                                        const int32_t* __restrict__ out_row) {
    constexpr int kWaves = kBlock / 64;
    constexpr int kBatch = C / 64 < 16 ? C / 64 : 16;  // entries per lane per batch
    constexpr int kSpan = 64 * kBatch;                  // entries per batch
    __shared__ float2 pairs[kWaves][C];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = (blockIdx.x * kWaves + wid) * 64;
    if (r0 >= n) return;
    const int row = r0 + lane;
    const bool valid = row < n;
    const int rs = valid ? row_ptr[row] : 0;
    const int re = valid ? row_ptr[row + 1] : 0;
    const int S = row_ptr[r0];
    const int E = row_ptr[min(r0 + 64, n)];
    float2* buf = pairs[wid];
    int k = rs;
    float acc = 0.0f;
    // batch b covers entries [b, b + kSpan); positions past E are clamped to
    // E - 1, so every load stays inside the operator's arrays and nothing is
    // branched on.  Offsets are taken from the batch start (uniform), 32-bit.
    int c[kBatch];
    float v[kBatch];
    auto load = [&](int b) {
        const char* cb = reinterpret_cast<const char*>(col + b);
        const char* vb = reinterpret_cast<const char*>(val + b);
        const unsigned last = (unsigned)(E - 1 - b);
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const unsigned o = min((unsigned)(lane + 64 * j), last) << 2;
            c[j] = *reinterpret_cast<const int32_t*>(cb + o);
            v[j] = *reinterpret_cast<const float*>(vb + o);
        }
    };
    if (S < E) load(S);
    for (int ws = S; ws < E; ws += C) {
        const int we = min(ws + C, E);
        for (int b = ws; b < we; b += kSpan) {
            float xg[kBatch], vc[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                // byte offsets fit 32 bits: the launcher requires n_sec < 2^30
                xg[j] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(x) +
                                                        ((unsigned)c[j] << 2));
                vc[j] = v[j];
            }
            // the next batch's indices and values load under these gathers
            // and under the window's chains
            if (b + kSpan < E) load(b + kSpan);
            // batches tile the window, so every place written is inside the
            // wave's region; places past the window are never read
#pragma unroll
            for (int j = 0; j < kBatch; ++j)
                buf[b - ws + lane + 64 * j] = make_float2(vc[j], xg[j]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int hi = min(re, we);
        const int k_next = max(k, hi);        // where the chain resumes
        for (; k < hi; k += 8) {
            float2 p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = buf[min(k + u, we - 1) - ws];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < hi) acc = __builtin_fmaf(p[u].x, p[u].y, acc);
        }
        k = k_next;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (valid) y[out_row ? out_row[row] : row] = acc;
}

template <int VEC, int G, int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_spmm_rowsplit(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, const int32_t* __restrict__ sched,
    const float* __restrict__ X, float* __restrict__ Y, int32_t n, int32_t F,
    int32_t n_rowblocks, int32_t rb_per_xcd, int32_t rb_per_block) {
    using VT = typename Vec<VEC>::T;
    constexpr int RPW = 64 / G;
    constexpr int RPB = (kBlock / 64) * RPW;
    int slab, rb;
    if (!block_coords(n_rowblocks, rb_per_xcd, rb_per_block, slab, rb)) return;
    const int lane = threadIdx.x & 63;
    // wave index is uniform across the wave: keep it in a scalar register
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = lane / G, li = lane % G;
    const int rr = rb * RPB + wave * RPW + grp;
    if (rr >= n) return;
    int row = sched ? sched[rr] : rr;
    if constexpr (G == 64) row = __builtin_amdgcn_readfirstlane(row);
    const int f = slab * (G * VEC) + li * VEC;
    const bool active = f < F;
    // lanes past the last feature gather a valid address (column 0 of the
    // same rows) and store nothing, so the loop below has no branches
    const float* Xf = X + (active ? f : 0);
    const int beg = row_ptr[row], end = row_ptr[row + 1];

    float acc[VEC];
#pragma unroll
    for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;

    int k = beg;
    if (k + U <= end) {
        // software pipeline: the (col, val) pairs of chunk k+U are fetched
        // while the gathers of chunk k are in flight
        int j[U];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            j[u] = col[k + u];
            v[u] = val[k + u];
        }
        while (true) {
            VT x[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                x[u] = *reinterpret_cast<const VT*>(Xf + (size_t)j[u] * F);
            const int kn = k + U;
            const bool more = kn + U <= end;
            int jn[U];
            float vn[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // clamped: always inside this (non-empty) row
                const int kk = min(kn + u, end - 1);
                jn[u] = col[kk];
                vn[u] = val[kk];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) fma_vec(v[u], x[u], acc);
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
        // tail of fewer than U nonzeros: all gathers issued at once, then the
        // FMAs of the real entries in order
        int j[U];
        float v[U];
        VT x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k + u, end - 1);
            j[u] = col[kk];
            v[u] = val[kk];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            x[u] = *reinterpret_cast<const VT*>(Xf + (size_t)j[u] * F);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k + u < end) fma_vec(v[u], x[u], acc);
    }
    if (active) store_vec<VEC, NT>(Y + (size_t)row * F + f, acc);
}


// ---------------------------------------------------------- row groups --
// One wave computes R rows (a group of consecutive rows in visiting order)
// for one slab of 64 * VEC features.  It walks the union of the rows' column
// indices in increasing order and gathers each X[j, slab] once for all the
// rows holding column j (the mask in the entry's top 8 bits), so neighbour
// rows shared inside the group cost one gather instead of several.  Each row
// still meets its nonzeros in sorted order with one fp32 FMA each: the same
// chain as the row-split kernels, bit for bit.  The values of a chunk of U
// entries (at most U * R <= 64 of them) are loaded one per lane and handed
// to the FMAs with readlane, so no scalar load sits inside the row loop.
//
// The scalar unit carries the loop's bookkeeping, so the full chunks keep it
// short: their U words come in one contiguous scalar load (the word array is
// padded past its end), the next chunk's words load under the current
// gathers, and with OFF32 (the gathered operand spans less than 4 GiB) a
// gather's offset is one 32-bit multiply.  Only the last, partial chunk
// clamps its entries and checks each against the group's end.
template <int VEC, int R, int U, bool NT, bool OFF32>
__global__ __launch_bounds__(kBlock) void k_spmm_grouped(
    const int32_t* __restrict__ gptr, const uint32_t* __restrict__ gword,
    const int32_t* __restrict__ gvptr, const float* __restrict__ gval,
    const int32_t* __restrict__ grow, const float* __restrict__ X, float* __restrict__ Y,
    int32_t n_groups, int32_t F, int32_t n_gblocks, int32_t rb_per_xcd, int32_t rb_per_block,
    int32_t layout, int32_t n_rows) {
    static_assert(U * R <= 128, "a chunk's values must fit two per lane");
    using VT = typename Vec<VEC>::T;
    int slab, gb;
    if (!block_coords(n_gblocks, rb_per_xcd, rb_per_block, slab, gb)) return;
    const int lane = threadIdx.x & 63;
    const int g = gb * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (g >= n_groups) return;
    // slab-planar operands (layout bit 0: X, bit 1: Y) keep slab s of all
    // rows contiguous, n_rows x ws floats from s * 64 * VEC * n_rows, so a
    // full slab's row starts on a 64 * VEC * 4-byte boundary
    const int f0 = slab * (64 * VEC), fl = lane * VEC;
    const int ws = min(64 * VEC, F - f0);
    const bool active = fl < ws;
    const size_t plane = (size_t)f0 * (size_t)n_rows;
    const float* Xs = (layout & 1) ? X + plane : X + f0;   // wave-uniform
    const float* Xf = Xs + (active ? fl : 0);
    const size_t xs = (layout & 1) ? (size_t)ws : (size_t)F;
    const uint32_t xs4 = (uint32_t)xs * 4u, lo4 = (uint32_t)(active ? fl : 0) * 4u;
    auto gather = [&](uint32_t wd) -> VT {
        const uint32_t j = wd & 0xFFFFFFu;
        if constexpr (OFF32)
            return *reinterpret_cast<const VT*>(reinterpret_cast<const char*>(Xs) +
                                                (j * xs4 + lo4));
        else
            return *reinterpret_cast<const VT*>(Xf + (size_t)j * xs);
    };
    const int e1 = __builtin_amdgcn_readfirstlane(gptr[g + 1]);
    int e = __builtin_amdgcn_readfirstlane(gptr[g]);
    int vo = __builtin_amdgcn_readfirstlane(gvptr[g]);

    float acc[R][VEC];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc[r][c] = 0.0f;

    if (e + U <= e1) {
        uint32_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = gword[e + u];
        while (true) {
            // this chunk's values, one per lane, and the next 64 when a chunk
            // can hold more (gval is padded by 128); loaded ahead of the
            // gathers so that waiting for them does not wait for the gathers
            const int vl = __builtin_bit_cast(int, gval[vo + lane]);
            const int vh = U * R > 64 ? __builtin_bit_cast(int, gval[vo + 64 + lane]) : 0;
            __builtin_amdgcn_sched_barrier(0);
            VT x[U];
#pragma unroll
            for (int u = 0; u < U; ++u) x[u] = gather(w[u]);
            // the next chunk's words (the array is padded by 2 U entries)
            uint32_t wn[U];
#pragma unroll
            for (int u = 0; u < U; ++u) wn[u] = gword[e + U + u];
            __builtin_amdgcn_sched_barrier(0);
            int c = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t m = w[u] >> 24;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (m & (1u << r)) {
                        const float v = __builtin_bit_cast(
                            float, U * R > 64 && c >= 64 ? __builtin_amdgcn_readlane(vh, c - 64)
                                                         : __builtin_amdgcn_readlane(vl, c));
                        ++c;
                        fma_vec(v, x[u], acc[r]);
                    }
                }
            }
            vo += c;
            e += U;
            if (e + U > e1) break;
#pragma unroll
            for (int u = 0; u < U; ++u) w[u] = wn[u];
        }
    }
    if (e < e1) {
        // the partial last chunk: entries past the group's end repeat its
        // last one and add nothing
        uint32_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = gword[min(e + u, e1 - 1)];
        const int vl = __builtin_bit_cast(int, gval[vo + lane]);
        const int vh = U * R > 64 ? __builtin_bit_cast(int, gval[vo + 64 + lane]) : 0;
        __builtin_amdgcn_sched_barrier(0);
        VT x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = gather(w[u]);
        int c = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (e + u < e1) {
                const uint32_t m = w[u] >> 24;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (m & (1u << r)) {
                        const float v = __builtin_bit_cast(
                            float, U * R > 64 && c >= 64 ? __builtin_amdgcn_readlane(vh, c - 64)
                                                         : __builtin_amdgcn_readlane(vl, c));
                        ++c;
                        fma_vec(v, x[u], acc[r]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = __builtin_amdgcn_readfirstlane(grow[(size_t)g * R + r]);
        float* y = (layout & 2) ? Y + plane + (size_t)row * ws + fl : Y + (size_t)row * F + f0 + fl;
        if (row >= 0 && active) store_vec<VEC, NT>(y, acc[r]);
    }
}

// ------------------------------------------------------------ merge path --
// The merge path of a CSR matrix interleaves its n row ends with its nnz
// nonzeros: nonzero k is consumed while k < row_ptr[i + 1] for the current
// row i, then row i's end.  Cutting the path into pieces of D items gives
// every wave the same amount of work whatever the row lengths.  A wave walks
// its piece with all 64 lanes on one row at a time (lanes over features,
// indices wave-uniform on the scalar unit), writes every row whose end lies
// in its piece, and leaves the partial sum of the row it does not finish in
// carry_val.  A second pass adds a row's carries, in path order, to the
// value written where the row ends.  A row inside one piece is the same FMA
// chain as in the row-split kernels; a row cut into p pieces is p chains
// added in order (p - 1 extra fp32 roundings).

// (row, nonzero) where path diagonal d = w * items starts: the smallest i in
// [max(0, d - nnz), min(d, n)] with row_ptr[i + 1] > d - i - 1.
__global__ __launch_bounds__(kBlock) void k_merge_split(const int32_t* __restrict__ row_ptr,
                                                        int32_t n, int32_t nnz, int32_t items,
                                                        int32_t waves,
                                                        int32_t* __restrict__ split) {
    const int w = blockIdx.x * kBlock + threadIdx.x;
    if (w > waves) return;
    const int64_t total = (int64_t)n + nnz;
    const int d = (int)min((int64_t)w * items, total);
    int lo = max(d - nnz, 0), hi = min(d, n);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row_ptr[mid + 1] <= d - mid - 1) lo = mid + 1;
        else hi = mid;
    }
    split[2 * w] = lo;
    split[2 * w + 1] = d - lo;
}

template <int VEC, int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_spmm_merge(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, const int32_t* __restrict__ split,
    const float* __restrict__ X, float* __restrict__ Y, int32_t n, int32_t F,
    int32_t waves, int32_t wgs, int32_t* __restrict__ carry_row,
    float* __restrict__ carry_val) {
    using VT = typename Vec<VEC>::T;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int slab = L / wgs;
    const int wg = L - slab * wgs;
    const int lane = threadIdx.x & 63;
    const int w = wg * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (w >= waves) return;
    const int i0 = __builtin_amdgcn_readfirstlane(split[2 * w]);
    const int k0 = __builtin_amdgcn_readfirstlane(split[2 * w + 1]);
    const int i1 = __builtin_amdgcn_readfirstlane(split[2 * w + 2]);
    const int k1 = __builtin_amdgcn_readfirstlane(split[2 * w + 3]);
    const int f = slab * (64 * VEC) + lane * VEC;
    const bool active = f < F;
    const float* Xf = X + (active ? f : 0);

    float acc[VEC];
#pragma unroll
    for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
    int r = i0;  // current row; its end offset e (none past the last row)
    int e = r < n ? row_ptr[r + 1] : INT32_MAX;
    auto finish_row = [&]() {
        if (active) store_vec<VEC, NT>(Y + (size_t)r * F + f, acc);
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;
        ++r;
        e = r < n ? row_ptr[r + 1] : INT32_MAX;
    };
    for (int k = k0; k < k1; k += U) {
        int j[U];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k + u, k1 - 1);
            j[u] = col[kk];
            v[u] = val[kk];
        }
        VT x[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            x[u] = *reinterpret_cast<const VT*>(Xf + (size_t)j[u] * F);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k + u < k1) {
                while (k + u >= e) finish_row();
                fma_vec(v[u], x[u], acc);
            }
        }
    }
    // rows ending inside the piece after its last nonzero (empty rows too)
    while (r < i1) finish_row();
    // row i1, when some of its nonzeros lie in [k0, k1), continues in the
    // next piece
    const bool carry = i1 < n && k1 > max(k0, row_ptr[min(i1, n - 1)]);
    if (slab == 0 && lane == 0) carry_row[w] = carry ? i1 : -1;
    if (carry && active) store_vec<VEC, false>(carry_val + (size_t)w * F + f, acc);
}

// Adds the carries of each cut row to the value written where the row ends:
// Y[r] = ((c_w + c_{w+1}) + ...) + Y[r] over the consecutive waves w.. that
// carry r.  One wave per carrying wave; only the first of a run works.
template <int VEC>
__global__ __launch_bounds__(kBlock) void k_merge_fixup(const int32_t* __restrict__ carry_row,
                                                        const float* __restrict__ carry_val,
                                                        float* __restrict__ Y, int32_t waves,
                                                        int32_t F) {
    const int w = blockIdx.x * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= waves) return;
    const int r = carry_row[w];
    if (r < 0 || (w > 0 && carry_row[w - 1] == r)) return;
    int w_end = w + 1;
    while (w_end < waves && carry_row[w_end] == r) ++w_end;
    for (int f = lane * VEC; f < F; f += 64 * VEC) {
        float s[VEC];
#pragma unroll
        for (int c = 0; c < VEC; ++c) s[c] = carry_val[(size_t)w * F + f + c];
        for (int q = w + 1; q < w_end; ++q) {
#pragma unroll
            for (int c = 0; c < VEC; ++c) s[c] += carry_val[(size_t)q * F + f + c];
        }
        float* y = Y + (size_t)r * F + f;
#pragma unroll
        for (int c = 0; c < VEC; ++c) y[c] = s[c] + y[c];
    }
}

// Several rows per wave (RPW = 2 or 4, G = 64 / RPW lanes each) with the
// rows' column indices and values kept on the scalar unit: every lane picks
// its row's (j, p) with a select instead of loading them itself, so a
// narrow slab (G * VEC features, 128 or 64 for VEC = 4) costs one vector
// gather per nonzero like the one-row-per-wave kernel.  The loop runs to the
// longest of the wave's rows (rows are sorted by length inside windows of
// 64 by the schedule); shorter rows skip their FMAs.
template <int RPW, typename T>
__device__ __forceinline__ T pick_row(int grp, const T (&a)[RPW]) {
    T r = a[0];
#pragma unroll
    for (int q = 1; q < RPW; ++q) r = (grp == q) ? a[q] : r;
    return r;
}

template <int VEC, int RPW, int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_spmm_multirow(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, const int32_t* __restrict__ sched,
    const float* __restrict__ X, float* __restrict__ Y, int32_t n, int32_t F,
    int32_t n_rowblocks, int32_t rb_per_xcd, int32_t rb_per_block) {
    using VT = typename Vec<VEC>::T;
    constexpr int G = 64 / RPW;
    constexpr int RPB = (kBlock / 64) * RPW;
    int slab, rb;
    if (!block_coords(n_rowblocks, rb_per_xcd, rb_per_block, slab, rb)) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = lane / G, li = lane % G;
    const int rr0 = rb * RPB + wave * RPW;
    if (rr0 >= n) return;
    int row[RPW], beg[RPW], len[RPW];
    int maxlen = 0;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        int r = -1, b = 0, e = 0;
        if (rr0 + q < n) {
            r = __builtin_amdgcn_readfirstlane(sched ? sched[rr0 + q] : rr0 + q);
            b = __builtin_amdgcn_readfirstlane(row_ptr[r]);
            e = __builtin_amdgcn_readfirstlane(row_ptr[r + 1]);
        }
        row[q] = r;
        beg[q] = b;
        len[q] = e - b;
        maxlen = max(maxlen, e - b);
    }
    const int my_row = pick_row<RPW>(grp, row);
    const int my_len = pick_row<RPW>(grp, len);
    const int f = slab * (G * VEC) + li * VEC;
    const bool active = f < F && my_row >= 0;
    const float* Xf = X + (f < F ? f : 0);

    float acc[VEC];
#pragma unroll
    for (int c = 0; c < VEC; ++c) acc[c] = 0.0f;

    for (int k = 0; k < maxlen; k += U) {
        int jq[U][RPW];
        float vq[U][RPW];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
                // clamped inside row q; an empty row reads entry 0, which
                // exists because some row of the wave is not empty
                const int kk = len[q] ? beg[q] + min(k + u, len[q] - 1) : 0;
                jq[u][q] = col[kk];
                vq[u][q] = val[kk];
            }
        VT x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = pick_row<RPW>(grp, jq[u]);
            x[u] = *reinterpret_cast<const VT*>(Xf + (size_t)j * F);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float v = pick_row<RPW>(grp, vq[u]);
            if (k + u < my_len) fma_vec(v, x[u], acc);
        }
    }
    if (active) store_vec<VEC, NT>(Y + (size_t)my_row * F + f, acc);
}

using SpmmFn = void (*)(const int32_t*, const int32_t*, const float*,
                        const int32_t*, const float*, float*, int32_t, int32_t,
                        int32_t, int32_t, int32_t);

template <int VEC, int G, int U>
SpmmFn pick_nt(bool nt) {
    return nt ? k_spmm_rowsplit<VEC, G, U, true> : k_spmm_rowsplit<VEC, G, U, false>;
}

template <int VEC, int G>
SpmmFn pick_u(int u, bool nt) {
    if (u >= 16) return pick_nt<VEC, G, 16>(nt);
    return u >= 8 ? pick_nt<VEC, G, 8>(nt) : pick_nt<VEC, G, 4>(nt);
}

template <int VEC>
SpmmFn pick_g(int g, int u, bool nt) {
    switch (g) {
        case 1: return pick_u<VEC, 1>(u, nt);
        case 2: return pick_u<VEC, 2>(u, nt);
        case 4: return pick_u<VEC, 4>(u, nt);
        case 8: return pick_u<VEC, 8>(u, nt);
        case 16: return pick_u<VEC, 16>(u, nt);
        case 32: return pick_u<VEC, 32>(u, nt);
        default: return pick_u<VEC, 64>(u, nt);
    }
}

template <int VEC, int RPW>
SpmmFn pick_multirow(int u, bool nt) {
    if (u >= 16) return nt ? k_spmm_multirow<VEC, RPW, 16, true> : k_spmm_multirow<VEC, RPW, 16, false>;
    if (u >= 8) return nt ? k_spmm_multirow<VEC, RPW, 8, true> : k_spmm_multirow<VEC, RPW, 8, false>;
    return nt ? k_spmm_multirow<VEC, RPW, 4, true> : k_spmm_multirow<VEC, RPW, 4, false>;
}

SpmmFn pick_spmm(const SpmmShape& s) {
    if (s.scalar_rows && (s.group == 16 || s.group == 32)) {
        const int rpw = 64 / s.group;
        switch (s.vec) {
            case 1: return rpw == 2 ? pick_multirow<1, 2>(s.unroll, s.nontemporal)
                                    : pick_multirow<1, 4>(s.unroll, s.nontemporal);
            case 2: return rpw == 2 ? pick_multirow<2, 2>(s.unroll, s.nontemporal)
                                    : pick_multirow<2, 4>(s.unroll, s.nontemporal);
            default: return rpw == 2 ? pick_multirow<4, 2>(s.unroll, s.nontemporal)
                                     : pick_multirow<4, 4>(s.unroll, s.nontemporal);
        }
    }
    switch (s.vec) {
        case 1: return pick_g<1>(s.group, s.unroll, s.nontemporal);
        case 2: return pick_g<2>(s.group, s.unroll, s.nontemporal);
        default: return pick_g<4>(s.group, s.unroll, s.nontemporal);
    }
}

// Grid of a row-split or grouped launch over n_rb row blocks and n_slabs
// slabs: rb_per_xcd (0: slab-major order), rbb (row blocks per XCD block,
// 0: none) and the workgroup count.
void xcd_grid(const SpmmShape& shape, int64_t n_rb, int64_t n_slabs, int64_t& rb_per_xcd,
              int64_t& rbb, int64_t& nb) {
    rbb = 0;
    if (shape.xcd_block > 0) {
        rb_per_xcd = (n_rb + kNumXcd - 1) / kNumXcd;
        rbb = std::min<int64_t>((int64_t)16 << shape.xcd_block, rb_per_xcd);
        const int64_t n_blk = (rb_per_xcd + rbb - 1) / rbb;
        nb = kNumXcd * n_blk * rbb * n_slabs;
    } else if (shape.xcd_rows) {
        rb_per_xcd = (n_rb + kNumXcd - 1) / kNumXcd;
        nb = kNumXcd * rb_per_xcd * n_slabs;
    } else {
        rb_per_xcd = 0;
        nb = n_rb * n_slabs;
    }
    if (nb > INT32_MAX) throw std::runtime_error("spmm grid too large");
}

int pow2_ceil(int64_t x) {
    int p = 1;
    while (p < x && p < 64) p <<= 1;
    return p;
}


}  // namespace

bool spmm_shape_valid(const SpmmShape& s, int64_t F, const void* X, const void* Y) {
    const auto a = reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y);
    if (!(s.vec == 1 || s.vec == 2 || s.vec == 4)) return false;
    if (F % s.vec != 0 || a % (4 * (uintptr_t)s.vec) != 0) return false;
    if (s.group < 1 || s.group > 64 || (s.group & (s.group - 1)) != 0) return false;
    if (s.wg_per_cu < 0 || s.wg_per_cu > kShapeOccupancyMask) return false;
    if (s.xcd_block < 0 || s.xcd_block > 8) return false;
    // merge path: lanes over features, one piece of the path per wave
    if (s.merge < 0 || s.merge > kMergeMax || (s.merge > 0 && s.group != 64)) return false;
    // grouped form: lanes over features, 4 or 8 rows per wave, not with merge
    if (s.rows_per_group != 0 && s.rows_per_group != kStagedRows &&
        s.rows_per_group != kTiledRows &&
        (!(s.rows_per_group == 4 || s.rows_per_group == 8) || s.group != 64 || s.merge))
        return false;
    // tiled form: 4 or 8 lanes per row, an LDS budget of 1 or 2 workgroups
    // per CU, unroll = slabs per workgroup
    if (s.rows_per_group == kTiledRows)
        return s.vec == 4 && (s.group == 4 || s.group == 8) &&
               (s.wg_per_cu == 1 || s.wg_per_cu == 2) && !s.merge && !s.scalar_rows &&
               (s.unroll == 4 || s.unroll == 8 || s.unroll == 16);
    // staged form: 64 lanes x 4 features, its own LDS (no occupancy cap)
    if (s.rows_per_group == kStagedRows &&
        (s.vec != 4 || s.group != 64 || s.merge || s.wg_per_cu || s.scalar_rows))
        return false;
    return s.unroll == 4 || s.unroll == 8 || s.unroll == 16;
}

int occupancy_lds_bytes(int wg_per_cu) {
    if (wg_per_cu <= 0) return 0;
    // 160 KiB of LDS per gfx950 CU, allocated in 512-byte granules
    constexpr int kLdsPerCu = 160 * 1024;
    return (kLdsPerCu / wg_per_cu) / 512 * 512;
}

SpmmShape spmm_default_shape(int64_t F, const void* X, const void* Y) {
    SpmmShape s;
    const auto a = reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y);
    // the widest access the feature count and alignment allow: fewer,
    // larger gathers beat one-row-per-wave layouts with narrow loads
    if (F % 4 == 0 && a % 16 == 0) s.vec = 4;
    else if (F % 2 == 0 && a % 8 == 0) s.vec = 2;
    else s.vec = 1;
    s.group = pow2_ceil((F + s.vec - 1) / s.vec);
    // 16 gathers in flight per lane: on knn30_200k 93.8 against 99.9 us at
    // F = 10, 91.3 against 102.9 at F = 12, 99.1 against 108.0 at F = 20
    // (F = 1 keeps 8: there the unroll picks the SpMV's staging window)
    s.unroll = F > 1 ? 16 : 8;
    s.nontemporal = false;
    s.schedule = true;
    return s;
}

std::vector<SpmmShape> spmm_candidate_shapes(int64_t F, const void* X, const void* Y,
                                             bool has_schedule) {
    std::vector<SpmmShape> out;
    const auto a = reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y);
    for (int vec : {4, 2, 1}) {
        if (F % vec != 0 || a % (4 * (uintptr_t)vec) != 0) continue;
        const int gmax = pow2_ceil((F + vec - 1) / vec);
        for (int g = 4; g <= gmax; g <<= 1) {
            if (g < gmax / 8) continue;       // slabs far narrower than F
            if (vec < 4 && g < 64 && g != gmax) continue;
            for (int u : {4, 8, 16})
                for (int nt = 0; nt < 2; ++nt)
                    for (int xr = 0; xr < 2; ++xr)
                        // scalar multi-row only with two rows per wave: four
                        // rows per wave measured 5-10x slower everywhere
                        for (int sr = 0; sr < ((vec == 4 && g == 32) ? 2 : 1); ++sr)
                            for (int sc = has_schedule ? 1 : 0; sc >= 0; --sc)
                                // fewer resident waves shrink the set of rows
                                // whose gathers share L2 at once (wide rows)
                                for (int occ : {0, 4, 5}) {
                                    if (occ && !(vec == 4 && g == 64 && u <= 8)) continue;
                                    SpmmShape s;
                                    s.vec = vec;
                                    s.group = g;
                                    s.unroll = u;
                                    s.nontemporal = nt != 0;
                                    s.xcd_rows = xr != 0;
                                    s.scalar_rows = sr != 0;
                                    s.schedule = sc != 0;
                                    s.wg_per_cu = occ;
                                    out.push_back(s);
                                }
        }
    }
    if (out.empty()) out.push_back(spmm_default_shape(F, X, Y));
    return out;
}

int64_t merge_waves(int64_t n, int64_t nnz, int m) {
    const int64_t items = (int64_t)256 << (m - 1);
    return std::max<int64_t>((n + nnz + items - 1) / items, 1);
}

namespace {

template <int VEC, bool NT>
void launch_merge_vec(const CsrView& P, const float* X, float* Y, int64_t F, int64_t waves,
                      int32_t items, int lds, hipStream_t s) {
    const MergeScratch& m = *P.merge;
    hipLaunchKernelGGL(k_merge_split, dim3((unsigned)((waves + 1 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, P.row_ptr, P.n, (int32_t)P.nnz, items,
                       (int32_t)waves, m.split);
    const int64_t wgs = (waves + kBlock / 64 - 1) / (kBlock / 64);
    const int64_t slabs = (F + 64 * VEC - 1) / (64 * VEC);
    if (wgs * slabs > INT32_MAX) throw std::runtime_error("merge-path grid too large");
    hipLaunchKernelGGL((k_spmm_merge<VEC, 8, NT>), dim3((unsigned)(wgs * slabs)),
                       dim3(kBlock), (unsigned)lds, s, P.row_ptr, P.col, P.val, m.split, X,
                       Y, P.n, (int32_t)F, (int32_t)waves, (int32_t)wgs, m.carry_row,
                       m.carry_val);
    hipLaunchKernelGGL((k_merge_fixup<VEC>), dim3((unsigned)wgs), dim3(kBlock), 0, s,
                       m.carry_row, m.carry_val, Y, (int32_t)waves, (int32_t)F);
}

void launch_spmm_merge(const CsrView& P, const float* X, float* Y, int64_t F,
                       const SpmmShape& shape, hipStream_t s) {
    if ((int64_t)P.n + P.nnz >= INT32_MAX)
        throw std::runtime_error("merge path: rows + nonzeros must fit int32");
    const int64_t waves = merge_waves(P.n, P.nnz, shape.merge);
    if (!P.merge || P.merge->waves < waves || P.merge->floats < waves * F)
        throw std::runtime_error("merge-path scratch missing or too small for this launch");
    const int32_t items = 256 << (shape.merge - 1);
    const int lds = occupancy_lds_bytes(shape.wg_per_cu);
#define SCL_MERGE(V)                                                                  \
    (shape.nontemporal ? launch_merge_vec<V, true>(P, X, Y, F, waves, items, lds, s)   \
                       : launch_merge_vec<V, false>(P, X, Y, F, waves, items, lds, s))
    if (shape.vec == 4) SCL_MERGE(4);
    else if (shape.vec == 2) SCL_MERGE(2);
    else SCL_MERGE(1);
#undef SCL_MERGE
    check_launch("merge-path spmm launch");
}

}  // namespace

namespace {

template <int VEC, int R, int U, bool NT>
void launch_grouped_vec(const GroupedCsr& G, const float* X, float* Y, int64_t F,
                        const SpmmShape& shape, int lds, int layout, int32_t n_rows,
                        int64_t n_gathered, hipStream_t s) {
    const int64_t n_gblocks = (G.n_groups + kBlock / 64 - 1) / (kBlock / 64);
    const int64_t n_slabs = (F + 64 * VEC - 1) / (64 * VEC);
    int64_t rb_per_xcd, rbb, nb;
    xcd_grid(shape, n_gblocks, n_slabs, rb_per_xcd, rbb, nb);
    // 32-bit gather offsets when every byte of the gathered rows (of either
    // layout) lies within 4 GiB of the slab's base
    const bool off32 = (uint64_t)n_gathered * (uint64_t)F * 4u < ((uint64_t)1 << 32);
    auto kern = off32 ? k_spmm_grouped<VEC, R, U, NT, true> : k_spmm_grouped<VEC, R, U, NT, false>;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kBlock), (unsigned)lds, s, G.gptr,
                       G.gword, G.gvptr, G.gval, G.grow, X, Y, (int32_t)G.n_groups, (int32_t)F,
                       (int32_t)n_gblocks, (int32_t)rb_per_xcd, (int32_t)rbb, (int32_t)layout,
                       n_rows);
}

// ---------------------------------------------------- LDS-staged groups --

// One workgroup of W waves: group g of kStagedRows rows (kStagedRows / W per
// wave) and one slab of 256 features (4 per lane).  The group's union
// columns are staged through LDS in rounds of C columns (1 KiB of X each,
// one coalesced load per column, issued back to back); every row then takes
// its stored entries whose union place falls in the round, in stored order,
// reading X from LDS.  A row's (union place, value) pairs sit in two VGPRs,
// lane l holding entry base + l: a ballot counts the round's entries of each
// row, and the rows of a wave advance in lockstep, entry t of every row at
// once, with a select keeping rows that have no entry t unchanged (their
// LDS reads all in flight together); a row longer than one 64-entry window
// finishes its round entry by entry after a reload.  Rows with ascending
// columns visit their entries in stored order, so each output element is
// the same FMA chain as the row-split kernel.
template <int C, int W, bool NT>
__global__ __launch_bounds__(64 * W) void k_spmm_staged(
    const int32_t* __restrict__ row_ptr, const float* __restrict__ val,
    const int32_t* __restrict__ upos, const int32_t* __restrict__ gptr,
    const int32_t* __restrict__ ucol, const int32_t* __restrict__ grow,
    const float* __restrict__ X, float* __restrict__ Y, int32_t n_groups, int32_t F,
    int32_t rb_per_xcd) {
    constexpr int RW = kStagedRows / W;   // rows per wave
    constexpr int CW = C / W;             // columns each wave stages per round
    static_assert(CW <= 64 && RW * W == kStagedRows, "staged layout");
    __shared__ f4 stage[C][64];
    int slab, g;
    if (!block_coords(n_groups, rb_per_xcd, 0, slab, g)) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int f = slab * 256 + lane * 4;
    const bool active = f < F;
    // lanes past the last feature stage and read column 0 and store nothing
    const float* Xf = X + (active ? f : 0);
    const int32_t* rows = grow + (int64_t)g * kStagedRows + wave * RW;
    int k[RW], e[RW], wb[RW], wpos[RW];
    float wval[RW];
    float acc[RW][4];
#pragma unroll
    for (int i = 0; i < RW; ++i) {
        const int r = __builtin_amdgcn_readfirstlane(rows[i]);
        k[i] = r >= 0 ? __builtin_amdgcn_readfirstlane(row_ptr[r]) : 0;
        e[i] = r >= 0 ? __builtin_amdgcn_readfirstlane(row_ptr[r + 1]) : 0;
        wb[i] = k[i];
        const int idx = k[i] + lane;
        wpos[i] = idx < e[i] ? upos[idx] : 0x7fffffff;
        wval[i] = idx < e[i] ? val[idx] : 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][c] = 0.0f;
    }
    const int ub = __builtin_amdgcn_readfirstlane(gptr[g]);
    const int un = __builtin_amdgcn_readfirstlane(gptr[g + 1]) - ub;
    for (int cb = 0; cb < un; cb += C) {
        const int cn = min(C, un - cb);
        // this wave stages columns [CW wave, CW wave + CW) of the round; past
        // the round's last column it loads that column again (slots never
        // read), so the loads issue back to back without branches
        const int c0 = wave * CW;
        if (c0 < cn) {
            const int mine = ucol[ub + cb + min(c0 + (lane % CW), cn - 1)];
            f4 xs[CW];
#pragma unroll
            for (int t = 0; t < CW; ++t)
                xs[t] = *reinterpret_cast<const f4*>(
                    Xf + (size_t)__builtin_amdgcn_readlane(mine, t) * F);
#pragma unroll
            for (int t = 0; t < CW; ++t) stage[c0 + t][lane] = xs[t];
        }
        __syncthreads();
        const int lim = cb + cn;
        // entries of each row in this round within its current window
        int m[RW], o[RW];
        int tmax = 0;
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            o[i] = k[i] - wb[i];
            m[i] = o[i] < 64 ? (int)__popcll(__ballot(lane >= o[i] && wpos[i] < lim)) : 0;
            tmax = max(tmax, m[i]);
        }
        for (int t = 0; t < tmax; ++t) {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                const bool on = t < m[i];
                const int oo = on ? o[i] + t : 0;
                const int p = __builtin_amdgcn_readlane(wpos[i], oo);
                const float v =
                    __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wval[i]), oo));
                const f4 x = stage[on ? p - cb : 0][lane];
                float nx[4] = {acc[i][0], acc[i][1], acc[i][2], acc[i][3]};
                fma_vec(v, x, nx);
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[i][c] = on ? nx[c] : acc[i][c];
            }
        }
        // advance; a row whose window ran out keeps going entry by entry
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            k[i] += m[i];
            while (k[i] < e[i] && k[i] - wb[i] == 64) {
                wb[i] = k[i];
                const int idx = k[i] + lane;
                wpos[i] = idx < e[i] ? upos[idx] : 0x7fffffff;
                wval[i] = idx < e[i] ? val[idx] : 0.0f;
                const int mm = (int)__popcll(__ballot(wpos[i] < lim));
                for (int t = 0; t < mm; ++t) {
                    const int p = __builtin_amdgcn_readlane(wpos[i], t);
                    const float v =
                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wval[i]), t));
                    fma_vec(v, stage[p - cb][lane], acc[i]);
                }
                k[i] += mm;
            }
        }
        __syncthreads();
    }
    if (active) {
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int r = __builtin_amdgcn_readfirstlane(rows[i]);
            if (r >= 0) store_vec<4, NT>(Y + (size_t)r * F + f, acc[i]);
        }
    }
}

template <int C, int W>
void launch_staged_c(const CsrView& P, const StagedCsr& G, const float* X, float* Y,
                     int64_t F, bool nt, int64_t nb, int64_t rb_per_xcd, hipStream_t s) {
    if (nt)
        hipLaunchKernelGGL((k_spmm_staged<C, W, true>), dim3((unsigned)nb), dim3(64 * W), 0, s,
                           P.row_ptr, P.val, G.upos, G.gptr, G.ucol, G.grow, X, Y,
                           (int32_t)G.n_groups, (int32_t)F, (int32_t)rb_per_xcd);
    else
        hipLaunchKernelGGL((k_spmm_staged<C, W, false>), dim3((unsigned)nb), dim3(64 * W), 0,
                           s, P.row_ptr, P.val, G.upos, G.gptr, G.ucol, G.grow, X, Y,
                           (int32_t)G.n_groups, (int32_t)F, (int32_t)rb_per_xcd);
}

// unroll selects the layout: 4 -> 8 waves, rounds of 64 union columns
// (64 KiB of LDS, two workgroups per CU); 8 -> 4 waves, 64 columns;
// 16 -> 8 waves, 128 columns (128 KiB, one workgroup per CU)
void launch_spmm_staged(const CsrView& P, const float* X, float* Y, int64_t F,
                        const SpmmShape& shape, hipStream_t s) {
    const StagedCsr* G = P.staged;
    if (!G || !G->built || G->n_groups * kStagedRows < P.n)
        throw std::runtime_error("staged form of the operator missing for this launch");
    const int64_t n_slabs = (F + 255) / 256;
    const int64_t rb_per_xcd = shape.xcd_rows ? (G->n_groups + kNumXcd - 1) / kNumXcd : 0;
    const int64_t nb = shape.xcd_rows ? kNumXcd * rb_per_xcd * n_slabs : G->n_groups * n_slabs;
    if (nb > INT32_MAX) throw std::runtime_error("staged spmm grid too large");
    if (shape.unroll >= 16)
        launch_staged_c<128, 8>(P, *G, X, Y, F, shape.nontemporal, nb, rb_per_xcd, s);
    else if (shape.unroll >= 8)
        launch_staged_c<64, 4>(P, *G, X, Y, F, shape.nontemporal, nb, rb_per_xcd, s);
    else
        launch_staged_c<64, 8>(P, *G, X, Y, F, shape.nontemporal, nb, rb_per_xcd, s);
    check_launch("staged spmm launch");
}

// ------------------------------------------------------------- LDS tiles --
// One workgroup of kTiledThreads lanes computes one tile (a run of rows of
// the visiting order whose column union fits LDS) for a chunk of slabs of
// LPR * 4 features.  Per slab, the waves copy the union's slab of X rows into
// LDS (one 16-byte piece per lane, 64 / LPR rows per wave load), then every
// lane group of LPR lanes walks one row's (union slot, value) entries from
// LDS in stored order and reads the neighbour's 4 features there:
// acc = fma(p, x, acc), the row-split chain bit for bit.  X leaves the
// vector memory path once per tile and slab instead of once per stored
// nonzero.  The next slab's pieces are loaded into registers while the
// current slab is computed, and written to LDS after it.
//
// Rows are padded to a multiple of four entries with (zero slot, -0.0):
// fma(-0.0, +0.0, acc) == acc for every acc (-0.0 + -0.0 = -0.0 and
// +0.0 + -0.0 = +0.0), so the padding needs no guard.  The zero slot is the
// last LDS slot; unions are capped one wave load short of it.
//
// Workgroups are numbered so that XCD x walks its contiguous range of tiles
// in blocks of as many tiles as it runs at once, and each block chunk by
// chunk: the tiles in flight on one XCD are neighbours in the visiting order
// working on the same slabs, so their shared X rows meet in that XCD's L2.
template <int LPR, int B>
struct TiledLds {
    static constexpr int kSlotBytes = LPR * 16;
    static constexpr int kSlots = tiled_slot_capacity(LPR, B);
    static constexpr int kEntries = tiled_entries_max(B);
    static constexpr int kSlotsPerLoad = 64 / LPR;  // one 1 KiB wave load
    static constexpr int kLoadsPerWave =
        (tiled_union_max(LPR, B) / kSlotsPerLoad + kTiledThreads / 64 - 1) / (kTiledThreads / 64);
    static constexpr int kBytes = kSlots * kSlotBytes + kEntries * 6;
    static_assert(kBytes <= 163840 / B, "tile exceeds the LDS budget");
    static_assert(kSlots % kSlotsPerLoad == 0 && kEntries % 8 == 0, "tile layout");
    static_assert(tiled_union_max(LPR, B) + kSlotsPerLoad == kSlots, "zero slot");
    static_assert(kSlots * kSlotBytes / 16 <= 65536, "slot units fit 16 bits");
};

template <int LPR, int B, bool NT>
__global__ __launch_bounds__(kTiledThreads) void k_spmm_tiled(
    const int4* __restrict__ info, const int32_t* __restrict__ ucol,
    const int32_t* __restrict__ trow, const int32_t* __restrict__ trbeg,
    const int32_t* __restrict__ trlen, const uint16_t* __restrict__ tslot,
    const float* __restrict__ tval, const float* __restrict__ X, float* __restrict__ Y,
    int32_t F, int32_t n_tiles, int32_t tiles_per_xcd, int32_t block_tiles,
    int32_t chunk_slabs, int32_t n_chunks, int32_t n_slabs) {
    using L = TiledLds<LPR, B>;
    constexpr int kWaves = kTiledThreads / 64;
    __shared__ __attribute__((aligned(16))) char lds[L::kBytes];
    char* xs = lds;
    uint16_t* ms = reinterpret_cast<uint16_t*>(lds + L::kSlots * L::kSlotBytes);
    float* mv = reinterpret_cast<float*>(lds + L::kSlots * L::kSlotBytes + L::kEntries * 2);
    // XCD x receives hardware ids x, x + 8, ...: its k-th workgroup takes
    // block k / (block_tiles n_chunks) of its tile range, chunk by chunk
    const int xcd = blockIdx.x % kNumXcd, k = blockIdx.x / kNumXcd;
    const int per_block = block_tiles * n_chunks;
    const int blk = k / per_block, rem = k - blk * per_block;
    const int chunk = rem / block_tiles;
    const int local = blk * block_tiles + (rem - chunk * block_tiles);
    const int tile = xcd * tiles_per_xcd + local;
    if (local >= tiles_per_xcd || tile >= n_tiles) return;  // whole workgroup
    const int tid = threadIdx.x;
    const int4 ti = info[tile];
    const int4 tn = info[tile + 1];
    const int u_off = ti.x, n_u = ti.y, e_off = ti.z, r_off = ti.w;
    const int ne = tn.z - e_off, nr = tn.w - r_off;
    // the tile's entries (ne is a multiple of 8) and the zero slot
    {
        const uint4* src = reinterpret_cast<const uint4*>(tslot + e_off);
        uint4* dst = reinterpret_cast<uint4*>(ms);
        for (int p = tid; p < ne / 8; p += kTiledThreads) dst[p] = src[p];
        const f4* vs = reinterpret_cast<const f4*>(tval + e_off);
        f4* vd = reinterpret_cast<f4*>(mv);
        for (int p = tid; p < ne / 4; p += kTiledThreads) vd[p] = vs[p];
        if (tid < LPR)
            *reinterpret_cast<f4*>(xs + (L::kSlots - 1) * L::kSlotBytes + tid * 16) =
                f4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    const int g = tid / LPR, sub = tid % LPR;
    int row = 0, rb = 0, rl = 0;
    if (g < nr) {
        row = trow[r_off + g];
        rb = trbeg[r_off + g];
        rl = (trlen[r_off + g] + 3) & ~3;
    }
    // X row offsets of the slots this wave stages (32-bit: the launcher
    // checks columns * F < 2^32); slots past the union stage row 0
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n_loads = (n_u + L::kSlotsPerLoad - 1) / L::kSlotsPerLoad;
    uint32_t roff[L::kLoadsPerWave];
#pragma unroll
    for (int i = 0; i < L::kLoadsPerWave; ++i) {
        const int ld = wave + kWaves * i;
        const int slot = ld * L::kSlotsPerLoad + lane / LPR;
        // clamped read (ucol holds one spare entry past the last union), so
        // the loads carry no branch and issue together
        const uint32_t c = (uint32_t)ucol[u_off + min(slot, max(n_u - 1, 0))];
        roff[i] = (ld < n_loads && slot < n_u) ? c * (uint32_t)F : 0u;
    }
    f4 stg[L::kLoadsPerWave];
    // every wave issues all its loads (past the union they read row 0 and
    // are not written to LDS): no branch between the loads, so they are all
    // in flight together
    auto fetch = [&](int slab) {
        // lanes past the last feature fetch feature 0 (never read)
        const int fs = slab * (LPR * 4) + (lane % LPR) * 4;
        const float* Xs = X + (fs < F ? fs : 0);
#pragma unroll
        for (int i = 0; i < L::kLoadsPerWave; ++i)
            stg[i] = *reinterpret_cast<const f4*>(Xs + roff[i]);
    };
    const int s_beg = chunk * chunk_slabs;
    const int s_end = min(s_beg + chunk_slabs, n_slabs);
    const char* xl = xs + sub * 16;
    fetch(s_beg);
    for (int slab = s_beg; slab < s_end; ++slab) {
#pragma unroll
        for (int i = 0; i < L::kLoadsPerWave; ++i) {
            const int ld = wave + kWaves * i;
            if (ld < n_loads) *reinterpret_cast<f4*>(xs + ld * 1024 + lane * 16) = stg[i];
        }
        if (slab + 1 < s_end) fetch(slab + 1);
        __syncthreads();
        const int f = slab * (LPR * 4) + sub * 4;
        if (g < nr && f < F) {
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int e = 0; e < rl; e += 4) {
                // slots in 16-byte units: lane offset + 16 * unit
                const uint2 sl = *reinterpret_cast<const uint2*>(ms + rb + e);
                const f4 v = *reinterpret_cast<const f4*>(mv + rb + e);
                const f4 x0 = *reinterpret_cast<const f4*>(xl + ((sl.x & 0xFFFFu) << 4));
                const f4 x1 = *reinterpret_cast<const f4*>(xl + ((sl.x >> 16) << 4));
                const f4 x2 = *reinterpret_cast<const f4*>(xl + ((sl.y & 0xFFFFu) << 4));
                const f4 x3 = *reinterpret_cast<const f4*>(xl + ((sl.y >> 16) << 4));
                fma_vec(v.x, x0, acc);
                fma_vec(v.y, x1, acc);
                fma_vec(v.z, x2, acc);
                fma_vec(v.w, x3, acc);
            }
            store_vec<4, NT>(Y + (size_t)row * F + f, acc);
        }
        __syncthreads();
    }
}

template <int LPR, int B, bool NT>
void launch_tiled_t(const TiledCsr& T, const float* X, float* Y, int64_t F, int chunk,
                    hipStream_t s) {
    const int64_t n_slabs = (F + LPR * 4 - 1) / (LPR * 4);
    const int64_t n_chunks = (n_slabs + chunk - 1) / chunk;
    const int64_t tpx = (T.n_tiles + kNumXcd - 1) / kNumXcd;
    // tiles one XCD runs at once: its 32 CUs, B workgroups each
    const int64_t block = std::min<int64_t>(32 * B, tpx);
    const int64_t blocks = (tpx + block - 1) / block;
    const int64_t nb = kNumXcd * blocks * block * n_chunks;
    if (nb > INT32_MAX) throw std::runtime_error("tiled spmm grid too large");
    hipLaunchKernelGGL((k_spmm_tiled<LPR, B, NT>), dim3((unsigned)nb), dim3(kTiledThreads), 0, s,
                       reinterpret_cast<const int4*>(T.info), T.ucol, T.trow, T.trbeg, T.trlen,
                       T.tslot, T.tval, X, Y, (int32_t)F, (int32_t)T.n_tiles, (int32_t)tpx,
                       (int32_t)block, (int32_t)chunk, (int32_t)n_chunks, (int32_t)n_slabs);
}

void launch_spmm_tiled(const CsrView& P, const float* X, float* Y, int64_t F,
                       const SpmmShape& shape, hipStream_t s) {
    const int lpr = shape.group, budget = shape.wg_per_cu;
    const TiledCsr* T = P.tiled ? &P.tiled[tiled_index(lpr, budget)] : nullptr;
    if (!T || T->lpr != lpr || T->budget != budget)
        throw std::runtime_error("tiled form of the operator missing for this launch");
    if ((uint64_t)P.n_sec * (uint64_t)F > 0xFFFFFFFFull)
        throw std::runtime_error("tiled form: columns x features must stay below 2^32");
    if (T->n_tiles == 0) return;
    const int chunk = std::max(shape.unroll, 1);
#define SCL_TILED(LP, BB)                                                               \
    (shape.nontemporal ? launch_tiled_t<LP, BB, true>(*T, X, Y, F, chunk, s)            \
                       : launch_tiled_t<LP, BB, false>(*T, X, Y, F, chunk, s))
    if (lpr == 8) {
        if (budget == 2) SCL_TILED(8, 2);
        else SCL_TILED(8, 1);
    } else {
        if (budget == 2) SCL_TILED(4, 2);
        else SCL_TILED(4, 1);
    }
#undef SCL_TILED
    check_launch("tiled spmm launch");
}

void launch_spmm_grouped(const CsrView& P, const float* X, float* Y, int64_t F,
                         const SpmmShape& shape, int layout, hipStream_t s) {
    const int R = shape.rows_per_group;
    const GroupedCsr* G = P.grouped ? &P.grouped[R == 8 ? 1 : 0] : nullptr;
    if (!G || G->R != R || G->n_groups * R < P.n)
        throw std::runtime_error("grouped form of the operator missing for this launch");
    const int lds = occupancy_lds_bytes(shape.wg_per_cu);
    // U entries per chunk with U * R <= 128 (a chunk's values fit two per lane)
#define SCL_GROUPED(V, RR, UU)                                                          \
    (shape.nontemporal                                                                  \
         ? launch_grouped_vec<V, RR, UU, true>(*G, X, Y, F, shape, lds, layout, P.n,    \
                                               P.n_sec, s)                              \
         : launch_grouped_vec<V, RR, UU, false>(*G, X, Y, F, shape, lds, layout, P.n,   \
                                                P.n_sec, s))
#define SCL_GROUPED_V(RR, UU)                          \
    do {                                               \
        if (shape.vec == 4) SCL_GROUPED(4, RR, UU);    \
        else if (shape.vec == 2) SCL_GROUPED(2, RR, UU); \
        else SCL_GROUPED(1, RR, UU);                   \
    } while (0)
    if (R == 8) {
        if (shape.unroll >= 16) SCL_GROUPED_V(8, 16);
        else if (shape.unroll >= 8) SCL_GROUPED_V(8, 8);
        else SCL_GROUPED_V(8, 4);
    } else {
        if (shape.unroll >= 16) SCL_GROUPED_V(4, 16);
        else SCL_GROUPED_V(4, 8);
    }
#undef SCL_GROUPED_V
#undef SCL_GROUPED
    check_launch("grouped spmm launch");
}

}  // namespace

void launch_spmv_staged(const CsrView& P, const int32_t* out_row, const float* x, float* y,
                        int unroll, hipStream_t s) {
    if (P.n <= 0) return;
    // windows of 256 * unroll entries per wave (4: 1,024, 8: 2,048, 16: 4,096)
    const unsigned nb = (unsigned)((P.n + kBlock - 1) / kBlock);
    if (unroll == 4)
        hipLaunchKernelGGL(k_spmv_staged<1024>, dim3(nb), dim3(kBlock), 0, s, P.row_ptr, P.col,
                           P.val, x, y, P.n, out_row);
    else if (unroll == 16)
        hipLaunchKernelGGL(k_spmv_staged<4096>, dim3(nb), dim3(kBlock), 0, s, P.row_ptr, P.col,
                           P.val, x, y, P.n, out_row);
    else
        hipLaunchKernelGGL(k_spmv_staged<2048>, dim3(nb), dim3(kBlock), 0, s, P.row_ptr, P.col,
                           P.val, x, y, P.n, out_row);
    check_launch("spmv launch");
}

bool spmm_layout_supported(const SpmmShape& shape, const CsrView& P) {
    return shape.rows_per_group > 0 && shape.rows_per_group != kStagedRows &&
           shape.rows_per_group != kTiledRows && !shape.merge && P.n == P.n_sec;
}

void launch_spmm_f32(const CsrView& P, const float* X, float* Y, int64_t F,
                     const SpmmShape& shape, hipStream_t s, int layout) {
    if (P.n <= 0 || F <= 0) return;
    if (layout != 0) {
        if (!spmm_layout_supported(shape, P))
            throw std::runtime_error("slab-planar operands need the row-grouped shapes");
        launch_spmm_grouped(P, X, Y, F, shape, layout, s);
        return;
    }
    if (shape.merge > 0) {
        launch_spmm_merge(P, X, Y, F, shape, s);
        return;
    }
    if (shape.rows_per_group == kStagedRows) {
        launch_spmm_staged(P, X, Y, F, shape, s);
        return;
    }
    if (shape.rows_per_group == kTiledRows) {
        launch_spmm_tiled(P, X, Y, F, shape, s);
        return;
    }
    if (shape.rows_per_group > 0) {
        launch_spmm_grouped(P, X, Y, F, shape, 0, s);
        return;
    }
    if (F == 1 && shape.group == 1 && shape.vec == 1 && P.n_sec < (1 << 30)) {
        launch_spmv_staged(P, nullptr, X, Y, shape.unroll, s);
        return;
    }
    if (F == 1 && shape.group == 1 && shape.vec == 1 &&
        ((reinterpret_cast<uintptr_t>(P.col) | reinterpret_cast<uintptr_t>(P.val)) & 15) == 0) {
        hipLaunchKernelGGL(k_spmv_lane, dim3((unsigned)((P.n + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, P.row_ptr, P.col, P.val,
                           shape.schedule ? P.sched : nullptr, X, Y, P.n);
        check_launch("spmv launch");
        return;
    }
    const int rpb = (kBlock / 64) * (64 / shape.group);
    const int64_t n_rowblocks = (P.n + rpb - 1) / rpb;
    const int64_t n_slabs = (F + (int64_t)shape.group * shape.vec - 1) /
                            ((int64_t)shape.group * shape.vec);
    int64_t rb_per_xcd, rbb, nb;
    xcd_grid(shape, n_rowblocks, n_slabs, rb_per_xcd, rbb, nb);
    const int32_t* sched = shape.schedule ? P.sched : nullptr;
    hipLaunchKernelGGL(pick_spmm(shape), dim3((unsigned)nb), dim3(kBlock),
                       (unsigned)occupancy_lds_bytes(shape.wg_per_cu), s,
                       P.row_ptr, P.col, P.val, sched, X, Y, P.n, (int32_t)F,
                       (int32_t)n_rowblocks, (int32_t)rb_per_xcd, (int32_t)rbb);
    check_launch("spmm launch");
}

// ---------------------------------------------------------- FP64 degrees --

namespace {

__global__ __launch_bounds__(kBlock) void k_row_sums_f64(
    const int32_t* __restrict__ row_ptr, const float* __restrict__ val,
    int32_t n, double* __restrict__ deg) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) acc += (double)val[k];
    deg[i] = acc;
}

__global__ __launch_bounds__(kBlock) void k_normalize_rows(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    float* __restrict__ val, int32_t n, const double* __restrict__ deg,
    int symmetric) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double di = deg[i];
    if (symmetric) {
        const double ri = di > 1e-15 ? 1.0 / sqrt(di) : 0.0;
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            const double dj = deg[col[k]];
            const double rj = dj > 1e-15 ? 1.0 / sqrt(dj) : 0.0;
            val[k] = (float)(((double)val[k] * ri) * rj);
        }
    } else {
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
            val[k] = di > 1e-15 ? (float)((double)val[k] / di) : 0.0f;
    }
}

constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kBlock / 64; ++q) t += sh[q];
    __syncthreads();
    return t;
}

// partial[b] = sum of deg over a grid-stride slice
__global__ __launch_bounds__(kBlock) void k_sum_partial(
    const double* __restrict__ x, int32_t n, double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    double v = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        v += x[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// out = sum of partial[0..m)
__global__ __launch_bounds__(kBlock) void k_sum_final(
    const double* __restrict__ partial, int m, double* __restrict__ out) {
    __shared__ double sh[kBlock / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < m; i += kBlock) v += partial[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) *out = v;
}

// partial[b] = sum over transposed rows j of |(P^T pi)_j - pi_j|
__global__ __launch_bounds__(kBlock) void k_residual_partial(
    const int32_t* __restrict__ t_row_ptr, const int32_t* __restrict__ t_col,
    const float* __restrict__ t_val, int32_t n, const double* __restrict__ deg,
    const double* __restrict__ total, double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    const double D = *total;
    double v = 0.0;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        double y = 0.0;
        for (int p = t_row_ptr[j]; p < t_row_ptr[j + 1]; ++p)
            y += (double)t_val[p] * (deg[t_col[p]] / D);
        v += fabs(y - deg[j] / D);
    }
    v = block_sum(v, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

}  // namespace

void launch_row_sums_f64(const int32_t* row_ptr, const float* val, int32_t n,
                         double* deg, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_row_sums_f64, dim3((n + kBlock - 1) / kBlock), dim3(kBlock),
                       0, s, row_ptr, val, n, deg);
    check_launch("row sums launch");
}

void launch_normalize_rows(const int32_t* row_ptr, const int32_t* col,
                           float* val, int32_t n, const double* deg,
                           int symmetric, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_normalize_rows, dim3((n + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, row_ptr, col, val, n, deg, symmetric);
    check_launch("normalize launch");
}

void launch_stationarity_residual(const int32_t* t_row_ptr, const int32_t* t_col,
                                  const float* t_val, int32_t n,
                                  const double* deg, double* partial,
                                  double* out, hipStream_t s) {
    const int nb = std::max(1, std::min(kRedBlocks, (n + kBlock - 1) / kBlock));
    double* total = partial + kRedBlocks;
    hipLaunchKernelGGL(k_sum_partial, dim3(nb), dim3(kBlock), 0, s, deg, n, partial);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, s, partial, nb, total);
    hipLaunchKernelGGL(k_residual_partial, dim3(nb), dim3(kBlock), 0, s, t_row_ptr,
                       t_col, t_val, n, deg, total, partial);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, s, partial, nb, out);
    check_launch("residual launch");
}

}  // namespace scl_amd
