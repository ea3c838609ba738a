// persist.hip -- one persistent launch running every graph of a
// block-diagonal batch through its forward steps, dY = Y - T, its fused
// adjoint + edge-gradient steps and the final dP permutation (persist.hpp
// describes the queues and the hand-off between a graph's steps).
//
// Every output is computed exactly as by the per-step kernels: each forward
// element is the fp32 FMA chain over its row's nonzeros in stored order, and
// each adjoint row runs adjoint_edge_row with the per-step kernel's default
// vec / group / unroll, so Y, dY, dX and dP are bit-identical to the
// replayed launch sequence.
#include <hip/hip_runtime.h>

#include "adjoint_row.hpp"
#include "device_common.hpp"
#include "persist.hpp"

#include <algorithm>
#include <stdexcept>

namespace scl_amd {

using namespace dev;

namespace {

// XCC_ID hardware register (hwreg 20): bits [3:0] name the XCD the wave
// runs on.
__device__ __forceinline__ int xcd_of_wave() {
    constexpr int kHwRegXccId = 20;
    return (int)(__builtin_amdgcn_s_getreg(kHwRegXccId | (0 << 6) | ((4 - 1) << 11)) & 0xFu);
}

// A work item's descriptor through the scalar unit: the queues are
// read-only in the launch, so their loads go through the constant address
// space, independent of the vector memory counter that the item's own loads
// and stores hold up.
__device__ __forceinline__ PersistItem load_item(const PersistItem* p) {
    using CI = const __attribute__((address_space(4))) int32_t;
    CI* q = (CI*)(p);
    PersistItem r;
    r.g = q[0];
    r.phase = q[1];
    r.begin = q[2];
    r.end = q[3];
    r.need = q[4];
    r.slab = q[5];
    r.pad1 = 0;
    r.pad2 = 0;
    return r;
}

__device__ __forceinline__ int load_agent(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One forward row (stored entries [beg, end)): Y[row, slab] = sum_k val_k
// X[col_k, slab] in stored order, one fp32 FMA per nonzero, for feature slab `slab` of G * 4
// features (lanes past F load column 0 and store nothing); with dY set also dY[row, :] = Y[row, :] -
// T[row, :].  The (col, val) pairs of the next U nonzeros load while the
// current U gathers are in flight; gathers past the row's end repeat its
// last entry and are not added.
template <int G, int U>
__device__ __forceinline__ void forward_row(const int32_t* __restrict__ col,
                                            const float* __restrict__ val,
                                            const float* __restrict__ X, float* __restrict__ Y,
                                            const float* __restrict__ T, float* __restrict__ dY,
                                            int32_t F, int slab, int row, int li, int beg,
                                            int end) {
    {
        const int f = (slab * G + li) * 4;
        const bool active = f < F;
        const float* Xf = X + (active ? f : 0);
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int j[U];
        float v[U];
        if (beg < end) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = min(beg + u, end - 1);
                j[u] = col[kk];
                v[u] = val[kk];
            }
        }
        for (int k = beg; k < end; k += U) {
            f4 x[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                x[u] = *reinterpret_cast<const f4*>(Xf + (size_t)j[u] * F);
            float vc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) vc[u] = v[u];
            const int kn = k + U;
            if (kn < end) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int kk = min(kn + u, end - 1);
                    j[u] = col[kk];
                    v[u] = val[kk];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k + u < end) fma_vec(vc[u], x[u], acc);
        }
        if (active) {
            const size_t o = (size_t)row * F + f;
            // streaming store: the outputs leave L2 to the gathered rows
            __builtin_nontemporal_store(f4{acc[0], acc[1], acc[2], acc[3]},
                                        reinterpret_cast<f4*>(Y + o));
            if (dY) {
                const f4 tt = *reinterpret_cast<const f4*>(T + o);
                *reinterpret_cast<f4*>(dY + o) =
                    f4{acc[0] - tt.x, acc[1] - tt.y, acc[2] - tt.z, acc[3] - tt.w};
            }
        }
    }
}

// The items of graph g completed so far, times 64: the sum of its
// per-sub-queue counters, read by the whole wave (coalesced loads) and
// broadcast.
__device__ __forceinline__ int done_of(const int32_t* done_g) {
    int sum = 0;
#pragma unroll
    for (int s = 0; s < kPersistSubQueues; ++s) sum += load_agent(done_g + s * kCtlStride);
    return __builtin_amdgcn_readfirstlane(sum);
}

// Polls graph g's done counters until they reach `need` (times 64), sleeping longer
// after every empty poll (4 .. 64 units of 64 cycles); false when the XCD's
// abort flag is set or the wait outlasts wait_ticks of the wall clock (which
// sets the flag, so every other wait of the XCD gives up at its next poll).
// The whole wave polls and branches on broadcast values, so the loop is
// wave-uniform.
__device__ __forceinline__ bool wait_done(const int32_t* done_g, int32_t* abort_flag, int need,
                                          int64_t wait_ticks) {
    need *= 64;
    if (done_of(done_g) >= need) return true;
    const int64_t t0 = wall_clock64();
    int nap = 1;
    bool ok = true;
    while (true) {
        if (nap < 32) nap <<= 1;
        for (int i = 0; i < nap; ++i) __builtin_amdgcn_s_sleep(2);
        if (done_of(done_g) >= need) break;
        if (__builtin_amdgcn_readfirstlane(load_agent(abort_flag)) != 0 ||
            wall_clock64() - t0 > wait_ticks) {
            ok = false;
            break;
        }
    }
    if (!ok) __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // no load of the item's inputs moves above the wait
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return ok;
}

// BWD = false: forward-only plans, with no adjoint code (fewer registers,
// more resident waves).
template <int G, int NC, bool BWD>
__global__ __launch_bounds__(kBlock) void k_persist(const PersistArgs a) {
    constexpr int RPW = 64 / G;
    constexpr int U = 8;    // adjoint chunk (the per-step kernel's default)
    // forward gathers in flight per lane: with the adjoint's registers
    // (4 waves per SIMD either way) 12; alone, 8 keeps 5 waves per SIMD
    constexpr int UF = BWD ? 12 : 8;
    // the forward runs half the adjoint's lane group (twice the rows per
    // wave; each item one feature slab of GF * 4, persist_forward_slabs of
    // them): twice the rows' gathers in flight at the register count the
    // adjoint sets; its results do not depend on it
    constexpr int GF = G >= 16 ? G / 2 : G;
    constexpr int RPWF = 64 / GF;
    // The item loop has no lane-dependent branch around its control values:
    // every lane issues the dequeue and completion atomics (with the same
    // operand) and everything the loop branches on is broadcast.  With
    // an `if (lane == 0)` dequeue the compiler may split the loop so that
    // lane 0 alone goes back to dequeue while the other lanes iterate on a
    // stale item index, never leaving it.
    const int q = __builtin_amdgcn_readfirstlane(xcd_of_wave());
    if (q >= kPersistQueues) return;
    const int lane = threadIdx.x & 63;
    const int grp = lane / G, li = lane % G;
    const int grpf = lane / GF, lif = lane % GF;
    // sub-queue of this wave: the first kPersistSubQueues / 4 workgroups of
    // each XCD (the first dispatched) already cover every sub-queue
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sub = (int)((blockIdx.x / kPersistQueues * (kBlock / 64) + wave) % kPersistSubQueues);
    const int l = q * kPersistSubQueues + sub;
    int32_t* head = a.ctl + l * kCtlStride;
    int32_t* abort_flag = a.ctl + (kCtlAbort + q) * kCtlStride;
    int32_t* done = a.ctl + kCtlDone * kCtlStride;
    const int q0 = __builtin_amdgcn_readfirstlane(a.q_off[l]);
    const int qn = __builtin_amdgcn_readfirstlane(a.q_off[l + 1]) - q0;
    const int t = a.t;
    const int F = a.F;
    // dequeue: every lane adds 1 (one atomic of 64 once folded), lane 0's
    // result / 64 is the item index.  A result that is not a multiple of 64
    // would mean the adds were not folded (the index could repeat): the
    // wave then stops and flags the XCD, which the plan check reports.
    auto claim = [&]() {
        return __hip_atomic_fetch_add(head, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto index_of = [&](int raw) -> int {
        const int first = __builtin_amdgcn_readlane(raw, 0);
        return (first & 63) == 0 ? first >> 6 : -1;
    };
    // The next item is claimed as this one starts (a wave's items stay in
    // queue order, so the drain argument of the queues holds); the compiler
    // reads a folded atomic's result right away, so the wave waits for the
    // claim there, together with the first row's bounds loaded just before,
    // and the next descriptor's load then runs under this item's work.
    const int it = index_of(claim());
    if (it < 0) {
        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    if (it >= qn) return;
    PersistItem item = load_item(a.items + q0 + it);
    // Completions are published per run of items of one (graph, phase): the
    // stores of the run drain once, then the run's count is added.  A wave
    // publishes everything pending before it waits, so it never waits on an
    // item it holds unpublished (the drain argument of the queues holds).
    int pend_g = -1, pend_phase = -1, pend_n = 0;
    auto publish = [&]() {
        if (pend_n > 0) {
            // the run's stores reach the XCD's L2 before its count is published
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_fetch_add(
                done + ((size_t)pend_g * kPersistSubQueues + sub) * kCtlStride, pend_n,
                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pend_n = 0;
        }
    };
    // the last dependency found satisfied: later items of the same graph
    // needing no more skip the check
    int ok_g = -1, ok_need = 0;
    while (true) {
        item.g = __builtin_amdgcn_readfirstlane(item.g);
        item.phase = __builtin_amdgcn_readfirstlane(item.phase);
        item.begin = __builtin_amdgcn_readfirstlane(item.begin);
        item.end = __builtin_amdgcn_readfirstlane(item.end);
        item.need = __builtin_amdgcn_readfirstlane(item.need);
        item.slab = __builtin_amdgcn_readfirstlane(item.slab);
        // the bounds of this lane group's first row load ahead of the claim,
        // whose result the wave waits for at once: the two round trips overlap
        const bool fwd = item.phase < t;
        int b0 = 0, e0 = 0;
        {
            const int r = item.begin + (fwd ? grpf : grp);
            const int32_t* rp = fwd ? a.row_ptr : a.t_row_ptr;
            if (r < item.end) {
                b0 = rp[r];
                e0 = rp[r + 1];
            }
        }
        // the next item: claimed now, its descriptor loaded under this one
        const int next = index_of(claim());
        PersistItem nitem{};
        if (next >= 0 && next < qn) nitem = load_item(a.items + q0 + next);
        if (item.g != pend_g || item.phase != pend_phase) publish();
        if (item.need > 0 && !(item.g == ok_g && item.need <= ok_need)) {
            publish();
            if (!wait_done(done + (size_t)item.g * kPersistSubQueues * kCtlStride, abort_flag,
                           item.need, a.wait_ticks))
                break;
            ok_g = item.g;
            ok_need = item.need;
        }
        if (fwd) {
            // forward step s: X_s = P X_{s-1}
            const int s = item.phase + 1;
            const float* in = s == 1 ? a.X : a.Xw + (size_t)(s - 2) * a.plane;
            float* out = s == t ? a.Y : a.Xw + (size_t)(s - 1) * a.plane;
            const bool grad = s == t && a.dY != nullptr;
            for (int r0 = item.begin; r0 < item.end; r0 += RPWF) {
                const int r = r0 + grpf;
                if (r < item.end) {
                    const bool first = r0 == item.begin;
                    const int beg = first ? b0 : a.row_ptr[r];
                    const int end = first ? e0 : a.row_ptr[r + 1];
                    forward_row<GF, UF>(a.col, a.val, in, out, grad ? a.T : nullptr,
                                        grad ? a.dY : nullptr, F, item.slab, r, lif, beg, end);
                }
            }
        } else if (BWD && item.phase < 2 * t) {
            // adjoint step s: G_{s-1} = P^T G_s, D_s = D_{s+1} + <G_s[i], X_{s-1}[j]>;
            // step 1 writes D_1 straight to dP at the entries' CSR positions
            const int s = 2 * t - item.phase;
            const float* Gs = s == t ? a.dY : a.Gw + (size_t)(s - 1) * a.plane;
            float* Gp = s == 1 ? a.dX : a.Gw + (size_t)(s - 2) * a.plane;
            const float* Xprev = s == 1 ? a.X : a.Xw + (size_t)(s - 2) * a.plane;
            for (int r0 = item.begin; r0 < item.end; r0 += RPW) {
                const int r = r0 + grp;
                if (r >= item.end) continue;
                const bool first = r0 == item.begin;
                const int beg = first ? b0 : a.t_row_ptr[r];
                const int end = first ? e0 : a.t_row_ptr[r + 1];
                if (a.want_dp) {
                    const int64_t dp = a.dpad[item.g];
                    const float* ds = s == t ? nullptr : a.D + (size_t)(s - 1) * a.dstride + dp;
                    // step s > 1 writes D_{s-1} (slot s - 2); step 1 the totals
                    // in transposed order (slot t - 1), which a gather after
                    // this kernel puts in CSR order: stores scattered from
                    // inside the step cost more than that pass
                    float* dd = a.D + (size_t)(s > 1 ? s - 2 : t - 1) * a.dstride + dp;
                    // clamped index loads: rows of other graphs are being
                    // written in this launch (see adjoint_edge_row)
                    adjoint_edge_row<4, G, U, NC, true, false, false>(
                        a.t_row_ptr, a.t_col, a.t_val, Gs, Xprev, Gp, ds, dd,
                        s < t ? 1 : 0, F, 0, F, r, li, nullptr, beg, end);
                } else {
                    adjoint_edge_row<4, G, U, NC, false, false, false>(
                        a.t_row_ptr, a.t_col, a.t_val, Gs, nullptr, Gp, nullptr, nullptr, 0, F,
                        0, F, r, li, nullptr, beg, end);
                }
            }
        }
        pend_g = item.g;
        pend_phase = item.phase;
        ++pend_n;
        if (next < 0) {
            __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        if (next >= qn) break;
        item = nitem;
    }
    publish();
}

__global__ __launch_bounds__(kBlock) void k_zero_i32(int32_t* __restrict__ p, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < m) p[i] = 0;
}

using PersistFn = void (*)(const PersistArgs);

template <bool BWD>
PersistFn pick_persist_bwd(int64_t F) {
    switch (persist_group(F)) {
        case 8: return k_persist<8, 1, BWD>;
        case 16: return k_persist<16, 1, BWD>;
        case 32: return k_persist<32, 1, BWD>;
        default: return F > 256 ? k_persist<64, 2, BWD> : k_persist<64, 1, BWD>;
    }
}

PersistFn pick_persist(int64_t F, bool backward) {
    return backward ? pick_persist_bwd<true>(F) : pick_persist_bwd<false>(F);
}

}  // namespace

int persist_group(int64_t F) {
    const int64_t lanes = F / 4;
    int g = 1;
    while (g < lanes && g < 64) g <<= 1;
    return g;
}

int persist_forward_slabs(int64_t F) {
    const int G = persist_group(F);
    const int GF = G >= 16 ? G / 2 : G;
    return (int)((F / 4 + GF - 1) / GF);
}

int persist_grid(int64_t F, bool backward, int device, int64_t widest) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(pick_persist(F, backward)), kBlock, 0);
    if (e != hipSuccess) throw std::runtime_error("persistent kernel occupancy query failed");
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) throw std::runtime_error("device CU count query failed");
    const int64_t full = (int64_t)std::max(1, per_cu) * std::max(1, cus);
    // workgroups are dealt to the XCDs in turn: kPersistQueues per round
    const int64_t waves_per_wg = kBlock / 64;
    // (at least enough workgroups to give every sub-queue a wave)
    const int64_t waves = std::max<int64_t>(widest, kPersistSubQueues);
    const int64_t need = (waves + waves_per_wg - 1) / waves_per_wg * kPersistQueues;
    return (int)std::min(full, need);
}

void launch_persist(const PersistArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(pick_persist(a.F, a.backward != 0), dim3((unsigned)grid), dim3(kBlock), 0, s, a);
    check_launch("persistent batch launch");
}

void launch_zero_i32(int32_t* p, int64_t m, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, p, m);
    check_launch("zero launch");
}

}  // namespace scl_amd
