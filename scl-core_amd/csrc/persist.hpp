// persist.hpp -- the persistent batched diffusion kernel (persist.hip) and
// its host-built work queues.
//
// Many graphs stacked block-diagonally run their whole forward (t steps),
// the loss gradient dY = Y - T, the reverse pass (t fused adjoint + edge
// steps) and the final dP permutation in ONE launch.  Each graph is pinned
// to one XCD: the host bin-packs the graphs over the XCDs and writes one
// queue of work items per XCD; a wave reads its XCD from the XCC_ID hardware
// register and takes items from that queue only.  A graph's data therefore
// passes between its steps through a single L2, so the hand-off needs no
// cache write-back: the producing wave waits for its stores to reach L2 and
// then bumps the graph's done counter; a consuming wave waits until the
// counter covers every item of the previous step.  Every intermediate (the
// iterates, the adjoint iterates and the per-step dP partials) has its own
// buffer, so no address is read before it is written within a launch and no
// vector L1 can hold a stale line of it.  Rows of a graph never share a
// 128-byte line with another graph's (F % 32 == 0, 128-byte aligned
// buffers, dP partials padded per graph).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

namespace scl_amd {

constexpr int kPersistQueues = 8;          // XCDs
// Each XCD's queue is dealt round-robin into kPersistSubQueues sub-queues
// (and each graph's done count into as many counters, one per sub-queue),
// so that no single address takes every dequeue or completion of an XCD:
// same-address atomics serialise at a few tens of ns each.
constexpr int kPersistSubQueues = 8;
constexpr int kPersistLanes = kPersistQueues * kPersistSubQueues;
constexpr int kPersistWaitSeconds = 2;     // a wait gives up after this long
// Control words, each alone in a 128-byte line (32 ints) so that polls of
// one counter never queue behind another's traffic.  Every lane of a wave
// adds to them (a wave-uniform operand, which the compiler folds into one
// atomic of 64 times it), so heads and done counts run in units of 64.
// head[l] of sub-queue l
// (l = XCD * kPersistSubQueues + sub) at word l, abort[q] at word
// kPersistLanes + q, done[g][sub] at word kPersistLanes + kPersistQueues +
// g * kPersistSubQueues + sub.
constexpr int kCtlStride = 32;
constexpr int64_t kCtlAbort = kPersistLanes;
constexpr int64_t kCtlDone = kPersistLanes + kPersistQueues;
constexpr int64_t persist_ctl_words(int64_t n_graphs) {
    return (kCtlDone + n_graphs * kPersistSubQueues) * kCtlStride;
}

// One work item: rows [begin, end) of graph g in phase `phase` (0 .. t-1:
// forward step phase + 1, feature slab `slab` only; t .. 2t-1: adjoint
// step 2t - phase, whole rows; the last adjoint step writes dP at its CSR
// positions).  It may start once graph g's done counters reach `need` (the
// items of its earlier phases).
struct PersistItem {
    int32_t g, phase, begin, end, need, slab, pad1, pad2;
};

struct PersistArgs {
    // the stacked operator (primary-slot CSR) and its transposed pattern
    const int32_t* row_ptr;
    const int32_t* col;
    const float* val;
    const int32_t* t_row_ptr;
    const int32_t* t_col;
    const float* t_val;
    // caller buffers (n x F row-major)
    const float* X;
    float* Y;
    const float* T;
    float* dY;
    float* dX;
    // plan buffers: X_1 .. X_{t-1}, G_1 .. G_{t-1} (planes of n x F) and t
    // slots of dstride floats, graph g at dpad[g] plus its transposed
    // positions: the dP partials D_1 .. D_{t-1} (slot s - 1 holds D_s) and,
    // in slot t - 1, the totals step 1 leaves, in transposed order
    float* Xw;
    float* Gw;
    float* D;
    const int32_t* dpad;
    // sub-queues: items[q_off[l] .. q_off[l + 1]) belong to sub-queue l
    const PersistItem* items;
    const int32_t* q_off;
    // control words (persist_ctl_words, zeroed before every launch)
    int32_t* ctl;
    int64_t plane;
    int64_t dstride;
    int64_t wait_ticks;  // wall-clock ticks a wait may take
    int32_t F;
    int32_t t;
    int32_t backward;
    int32_t want_dp;
};

// The work queues of a batched plan over graphs with row offsets off[0..G]
// and nonzero offsets nz[0..G] (nz[g] = first stored entry of graph g, in
// CSR and transposed order alike): graphs bin-packed over the queues by
// rows + nonzeros, largest first; a queue takes its graphs
// graphs_per_chunk at a time (0: 4) and, for each chunk, phase by phase,
// graph by graph, lists the items (rows_per_item rows, 0: 16; a forward
// phase slab by slab over fwd_slabs feature slabs, so that the graph's X
// rows a queue gathers at a time are one slab wide), then deals that list
// round-robin into the XCD's sub-queues.  Every item's `need` counts its
// graph's items of earlier phases; each sub-queue keeps the list's (chunk,
// phase) order, so the XCD's sub-queues drained in order never wait on an
// item not yet taken.
struct PersistQueues {
    std::vector<PersistItem> items;
    std::vector<int32_t> q_off;   // kPersistLanes + 1
    std::vector<int32_t> expect;  // per graph: its items over all phases
    std::vector<int32_t> dpad;    // per graph: dP partial offset minus nz[g]
    int64_t dstride = 0;          // floats per dP partial (padded, + 32)
    int64_t widest = 0;           // longest run of one phase's items in an XCD's
                                  // list: more waves per XCD than this only wait
};
PersistQueues build_persist_queues(const std::vector<int64_t>& off,
                                   const std::vector<int64_t>& nz, int64_t t, bool backward,
                                   int64_t rows_per_item,
                                   int64_t graphs_per_chunk, int fwd_slabs);

// Lanes per row of the persistent kernel for F features (F % 32 == 0,
// 32 <= F <= 512): the fused adjoint's default group, so dP matches the
// per-step kernels bit for bit.
int persist_group(int64_t F);
// Feature slabs of the persistent forward for F features (each a separate
// work item): the forward runs half the adjoint's lane group from 16 lanes
// up, so twice the rows per wave and twice the slabs.
int persist_forward_slabs(int64_t F);
// Workgroups of one persistent launch: as many as fit on the device at
// once, but no more per XCD than `widest` items need (one wave each).
int persist_grid(int64_t F, bool backward, int device, int64_t widest);
void launch_persist(const PersistArgs& a, int grid, hipStream_t s);
// p[0, m) = 0 with a kernel (graph-capture safe)
void launch_zero_i32(int32_t* p, int64_t m, hipStream_t s);

}  // namespace scl_amd
