// plan.cpp -- a whole diffusion step sequence recorded once into a hipGraph
// and replayed with one launch: forward with checkpoints, optionally the
// gradient of 0.5 ||Y - T||^2 and the reverse pass.  Used for many small
// graphs stacked block-diagonally (scl_sparse_block_diag), where per-kernel
// launch overhead would otherwise dominate.  The batched plan runs the same
// sequence for every graph of the stack inside one persistent kernel
// (persist.hip), each graph's steps ordered by its own dependencies.
#include "internal.hpp"
#include "kernels.hpp"
#include "persist.hpp"
#include "scl_diffusion_gpu.h"

#include <algorithm>
#include <limits>
#include <numeric>
#include <string>

using namespace scl_amd;

struct scl_gpu_plan {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    scl_sparse_t op = nullptr;
    int device = -1;
    // batched (persistent) plans: device buffers owned by the plan, the
    // control words and the done count every graph must reach
    bool persistent = false;
    std::vector<void*> owned;
    int32_t* ctl = nullptr;
    std::vector<int32_t> expect_done;
    std::vector<int32_t> queue_len;
};

namespace {

void unpin(scl_sparse_t op, int device) {
    std::lock_guard<std::mutex> lk(op->mu);
    auto it = op->dev.find(device);
    if (it != op->dev.end() && it->second && it->second->pins > 0) it->second->pins--;
}

void release_plan(scl_gpu_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (void* b : p->owned) (void)hipFree(b);
    p->owned.clear();
}

template <typename T>
T* plan_alloc(scl_gpu_plan* p, size_t count, const char* what) {
    void* b = nullptr;
    hip_check(hipMalloc(&b, std::max<size_t>(count, 1) * sizeof(T)), what);
    p->owned.push_back(b);
    return static_cast<T*>(b);
}

bool aligned128(const void* p) { return p == nullptr || reinterpret_cast<uintptr_t>(p) % 128 == 0; }

}  // namespace

extern "C" {

SCL_EXPORT scl_error_t scl_gpu_plan_diffusion(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_real_t* ckpt_dev, scl_real_t* fwd_work_dev, const scl_real_t* T_dev,
    scl_real_t* dY_dev, scl_real_t* dX_dev, scl_real_t* dP_dev,
    void* bwd_work_dev, scl_size_t bwd_work_bytes, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every,
    scl_gpu_plan_t* out) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_CHECK(T_dev == nullptr || dY_dev != nullptr, SCL_ERROR_NULL_POINTER,
                  "dY buffer is required with a target");
    SCL_API_CHECK(dY_dev == nullptr || dX_dev != nullptr, SCL_ERROR_NULL_POINTER,
                  "dX buffer is required for the reverse pass");
    SCL_API_TRY
        const int dev = current_device();
        {
            // uploads happen now, so the capture below records kernels only
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR& d = device_csr(transition, dev);
            if (dY_dev) ensure_transpose(transition, d, nullptr);
            d.pins++;
        }
        auto* plan = new scl_gpu_plan();
        plan->op = transition;
        plan->device = dev;
        hipStream_t s = nullptr;
        scl_error_t rc = SCL_OK;
        try {
            hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "plan stream");
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
            rc = scl_gpu_diffuse_forward_ckpt_device(transition, X_dev, Y_dev, ckpt_dev,
                                                     fwd_work_dev, n_nodes, n_features,
                                                     n_steps, every, s);
            if (rc == SCL_OK && T_dev)
                launch_sub_f32(Y_dev, T_dev, dY_dev, n_nodes * n_features, s);
            if (rc == SCL_OK && dY_dev)
                rc = scl_gpu_diffuse_backward_device(
                    transition, X_dev, ckpt_dev, dY_dev, dX_dev, dP_dev, bwd_work_dev,
                    bwd_work_bytes, n_nodes, n_features, n_steps, every, s);
            hipGraph_t g = nullptr;
            hipError_t ec = hipStreamEndCapture(s, &g);
            plan->graph = g;
            if (rc != SCL_OK) fail(rc, scl_get_last_error());
            hip_check(ec, "end capture");
            hip_check(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0),
                      "graph instantiate");
            hip_check(hipStreamDestroy(s), "plan stream");
            s = nullptr;
        } catch (...) {
            if (s) (void)hipStreamDestroy(s);
            if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
            if (plan->graph) (void)hipGraphDestroy(plan->graph);
            unpin(transition, dev);
            delete plan;
            throw;
        }
        *out = plan;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_plan_diffusion_batched(
    scl_sparse_t transition, const scl_index_t* block_rows, scl_size_t n_blocks,
    const scl_real_t* X_dev, scl_real_t* Y_dev, const scl_real_t* T_dev, scl_real_t* dY_dev,
    scl_real_t* dX_dev, scl_real_t* dP_dev, scl_index_t n_features, scl_index_t n_steps,
    scl_index_t rows_per_item, scl_index_t graphs_per_chunk, scl_gpu_plan_t* out) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(block_rows, "Block sizes are null");
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_CHECK(X_dev && Y_dev, SCL_ERROR_NULL_POINTER, "X or Y is null");
    SCL_API_CHECK(n_blocks > 0 && n_steps > 0 && rows_per_item >= 0 && graphs_per_chunk >= 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_CHECK(dX_dev == nullptr || (T_dev && dY_dev), SCL_ERROR_NULL_POINTER,
                  "T and dY are required for the reverse pass");
    SCL_API_CHECK(dP_dev == nullptr || dX_dev, SCL_ERROR_NULL_POINTER,
                  "dX is required with dP");
    SCL_API_CHECK(n_features >= 32 && n_features <= 512 && n_features % 32 == 0,
                  SCL_ERROR_INVALID_ARGUMENT,
                  "Batched plans take a multiple of 32 features, at most 512");
    SCL_API_CHECK(aligned128(X_dev) && aligned128(Y_dev) && aligned128(T_dev) &&
                      aligned128(dY_dev) && aligned128(dX_dev),
                  SCL_ERROR_INVALID_ARGUMENT, "X, Y, T, dY and dX must be 128-byte aligned");
    SCL_API_CHECK(n_steps <= 1 << 20 && n_blocks < (1u << 30), SCL_ERROR_INVALID_ARGUMENT,
                  "Too many steps or blocks");
    SCL_API_TRY
        scl_sparse_matrix* h = transition;
        const int64_t n = h->primary(), F = n_features, t = n_steps;
        if (h->rows != h->cols) fail(SCL_ERROR_DIMENSION_MISMATCH, "Transition matrix must be square");
        const int64_t G = (int64_t)n_blocks;
        std::vector<int64_t> off(G + 1, 0);
        for (int64_t g = 0; g < G; ++g) {
            if (block_rows[g] <= 0) fail(SCL_ERROR_INVALID_ARGUMENT, "Block sizes must be positive");
            off[g + 1] = off[g] + block_rows[g];
        }
        if (off[G] != n) fail(SCL_ERROR_DIMENSION_MISMATCH, "Block sizes do not add up to the operator");
        if (n > std::numeric_limits<int32_t>::max() / 2 ||
            n > std::numeric_limits<int64_t>::max() / F / 4)
            fail(SCL_ERROR_INVALID_ARGUMENT, "n_nodes * n_features overflows");
        // every stored entry of a block's rows stays inside the block
        std::vector<int64_t> nz(G + 1, 0);
        for (int64_t g = 0; g < G; ++g) {
            const int64_t a = h->indptr[off[g]], b = h->indptr[off[g + 1]];
            for (int64_t k = a; k < b; ++k)
                if (h->indices[k] < off[g] || h->indices[k] >= off[g + 1])
                    fail(SCL_ERROR_INVALID_ARGUMENT, "Operator is not block diagonal over these blocks");
            nz[g] = a;
        }
        nz[G] = h->nnz;
        const bool backward = dX_dev != nullptr;
        const bool want_dp = backward && dP_dev != nullptr;
        PersistQueues Q = build_persist_queues(off, nz, t, backward, rows_per_item,
                                               graphs_per_chunk, persist_forward_slabs(F));
        // heads and done counts run in units of 64 (int32)
        for (int l = 0; l < kPersistLanes; ++l)
            if ((int64_t)(Q.q_off[l + 1] - Q.q_off[l]) + 2 > std::numeric_limits<int32_t>::max() / 64)
                fail(SCL_ERROR_INVALID_ARGUMENT, "Too many work items in one queue");
        if (Q.dstride > std::numeric_limits<int32_t>::max())
            fail(SCL_ERROR_INVALID_ARGUMENT, "Too many stored entries for one batched plan");
        const int dev = current_device();
        DeviceCSR* dp = nullptr;
        {
            std::lock_guard<std::mutex> lk(h->mu);
            DeviceCSR& d = device_csr(h, dev);
            if (backward) ensure_transpose(h, d, nullptr);
            d.pins++;
            dp = &d;
        }
        auto* plan = new scl_gpu_plan();
        plan->op = transition;
        plan->device = dev;
        plan->persistent = true;
        plan->expect_done = Q.expect;
        hipStream_t s = nullptr;
        try {
            const size_t plane = (size_t)n * (size_t)F;
            PersistArgs a{};
            a.row_ptr = dp->row_ptr;
            a.col = dp->col;
            a.val = dp->val;
            a.t_row_ptr = dp->t_row_ptr;
            a.t_col = dp->t_col;
            a.t_val = dp->t_val;
            a.X = X_dev;
            a.Y = Y_dev;
            a.T = T_dev;
            a.dY = backward ? dY_dev : nullptr;
            a.dX = dX_dev;
            a.Xw = t >= 2 ? plan_alloc<float>(plan, plane * (size_t)(t - 1), "plan iterates") : nullptr;
            a.Gw = backward && t >= 2
                       ? plan_alloc<float>(plan, plane * (size_t)(t - 1), "plan adjoint iterates")
                       : nullptr;
            // partials D_1 .. D_{t-1} and the totals of step 1 (transposed
            // order, padded per graph), and where each CSR entry's total is
            a.D = want_dp ? plan_alloc<float>(plan, (size_t)Q.dstride * (size_t)t, "plan dP partials")
                          : nullptr;
            int32_t* d_dp_map = nullptr;
            if (want_dp) {
                std::vector<int32_t> t_src((size_t)h->nnz), map((size_t)h->nnz);
                hip_check(hipMemcpy(t_src.data(), dp->t_src, t_src.size() * 4,
                                    hipMemcpyDeviceToHost), "plan download");
                for (int64_t g = 0; g < G; ++g)
                    for (int64_t q = nz[g]; q < nz[g + 1]; ++q)
                        map[(size_t)t_src[(size_t)q]] = (int32_t)(Q.dpad[g] + q);
                d_dp_map = plan_alloc<int32_t>(plan, map.size(), "plan dP map");
                hip_check(hipMemcpy(d_dp_map, map.data(), map.size() * 4, hipMemcpyHostToDevice),
                          "plan upload");
            }
            int32_t* d_dpad = plan_alloc<int32_t>(plan, (size_t)G, "plan offsets");
            PersistItem* d_items = plan_alloc<PersistItem>(plan, Q.items.size(), "plan items");
            int32_t* d_qoff = plan_alloc<int32_t>(plan, Q.q_off.size(), "plan queues");
            const int64_t n_ctl = persist_ctl_words(G);
            plan->ctl = plan_alloc<int32_t>(plan, (size_t)n_ctl, "plan control");
            hip_check(hipMemcpy(d_dpad, Q.dpad.data(), Q.dpad.size() * 4, hipMemcpyHostToDevice),
                      "plan upload");
            hip_check(hipMemcpy(d_items, Q.items.data(), Q.items.size() * sizeof(PersistItem),
                                hipMemcpyHostToDevice), "plan upload");
            hip_check(hipMemcpy(d_qoff, Q.q_off.data(), Q.q_off.size() * 4, hipMemcpyHostToDevice),
                      "plan upload");
            int wall_khz = 0;
            hip_check(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev),
                      "wall clock rate");
            a.dpad = d_dpad;
            a.items = d_items;
            a.q_off = d_qoff;
            a.ctl = plan->ctl;
            a.plane = (int64_t)plane;
            a.dstride = Q.dstride;
            a.wait_ticks = (int64_t)std::max(wall_khz, 1) * 1000 * kPersistWaitSeconds;
            a.F = (int32_t)F;
            a.t = (int32_t)t;
            a.backward = backward ? 1 : 0;
            a.want_dp = want_dp ? 1 : 0;
            for (int l = 0; l < kPersistLanes; ++l)
                plan->queue_len.push_back(Q.q_off[l + 1] - Q.q_off[l]);
            const int grid = persist_grid(F, backward, dev, Q.widest);
            hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "plan stream");
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
            launch_zero_i32(plan->ctl, n_ctl, s);
            launch_persist(a, grid, s);
            if (want_dp)
                launch_gather_f32(d_dp_map, a.D + (size_t)(t - 1) * (size_t)Q.dstride, dP_dev,
                                  h->nnz, s);
            hipGraph_t g = nullptr;
            hip_check(hipStreamEndCapture(s, &g), "end capture");
            plan->graph = g;
            hip_check(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0),
                      "graph instantiate");
            hip_check(hipStreamDestroy(s), "plan stream");
            s = nullptr;
        } catch (...) {
            if (s) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                (void)hipStreamDestroy(s);
            }
            release_plan(plan);
            unpin(transition, dev);
            delete plan;
            throw;
        }
        *out = plan;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_plan_check(scl_gpu_plan_t plan, void* stream) {
    SCL_API_CHECK_NULL(plan, "Plan is null");
    SCL_API_TRY
        hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "plan check");
        if (!plan->persistent) SCL_API_RETURN_OK;
        const size_t G = plan->expect_done.size();
        std::vector<int32_t> ctl((size_t)persist_ctl_words((int64_t)G));
        hip_check(hipMemcpy(ctl.data(), plan->ctl, ctl.size() * 4, hipMemcpyDeviceToHost),
                  "plan check");
        auto word = [&](size_t i) { return ctl[i * kCtlStride]; };
        for (int q = 0; q < kPersistQueues; ++q)
            if (word(kCtlAbort + q) != 0)
                fail(SCL_ERROR_INTERNAL, "batched plan: a wait on XCD " + std::to_string(q) +
                                             " gave up");
        for (int l = 0; l < kPersistLanes; ++l)
            if ((int64_t)word(l) < (int64_t)plan->queue_len[l] * 64)
                fail(SCL_ERROR_INTERNAL, "batched plan: sub-queue " + std::to_string(l) +
                                             " was not drained (no wave ran there)");
        for (size_t g = 0; g < G; ++g) {
            int64_t sum = 0;
            for (int sq = 0; sq < kPersistSubQueues; ++sq)
                sum += word(kCtlDone + g * kPersistSubQueues + sq);
            if (sum != (int64_t)plan->expect_done[g] * 64)
                fail(SCL_ERROR_INTERNAL, "batched plan: graph " + std::to_string(g) +
                                             " did not complete");
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_plan_launch(scl_gpu_plan_t plan, void* stream) {
    SCL_API_CHECK_NULL(plan, "Plan is null");
    SCL_API_TRY
        hip_check(hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream)),
                  "graph launch");
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_plan_destroy(scl_gpu_plan_t* plan) {
    if (plan == nullptr || *plan == nullptr) SCL_API_RETURN_OK;
    SCL_API_TRY
        scl_gpu_plan* p = *plan;
        release_plan(p);
        unpin(p->op, p->device);
        delete p;
        *plan = nullptr;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
