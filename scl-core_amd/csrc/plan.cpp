// plan.cpp -- a whole diffusion step sequence recorded once into a hipGraph
// and replayed with one launch: forward with checkpoints, optionally the
// gradient of 0.5 ||Y - T||^2 and the reverse pass.  Used for many small
// graphs stacked block-diagonally (scl_sparse_block_diag), where per-kernel
// launch overhead would otherwise dominate.
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_diffusion_gpu.h"

using namespace scl_amd;

struct scl_gpu_plan {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    scl_sparse_t op = nullptr;
    int device = -1;
};

namespace {

void unpin(scl_sparse_t op, int device) {
    std::lock_guard<std::mutex> lk(op->mu);
    auto it = op->dev.find(device);
    if (it != op->dev.end() && it->second && it->second->pins > 0) it->second->pins--;
}

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
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        unpin(p->op, p->device);
        delete p;
        *plan = nullptr;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
