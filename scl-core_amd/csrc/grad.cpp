// grad.cpp -- checkpointed forward and reverse-mode diffusion (device
// pointers), plus in-place replacement of an operator's values from device
// memory.  The reference has no gradient; these calls back the
// torch.autograd.Function in scl_amd/autograd.py.
//
// Forward with checkpoints every c steps stores X_c, X_2c, ... (those before
// X_t).  The backward walks the t steps in segments of c, newest first: it
// recomputes the segment's iterates from its checkpoint with the forward
// SpMM, then runs one fused adjoint + edge-gradient kernel per step
// (grad.hip), so each step reads G_s once for both products.
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_diffusion_gpu.h"

#include <limits>

using namespace scl_amd;

namespace {

constexpr size_t kAlign = 256;

size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

void check_square(const scl_sparse_matrix* P, int64_t n, int64_t F) {
    if (P->rows != P->cols)
        fail(SCL_ERROR_DIMENSION_MISMATCH, "Transition matrix must be square");
    if (P->rows != n)
        fail(SCL_ERROR_DIMENSION_MISMATCH,
             "n_nodes does not match the transition matrix dimensions");
    if (F > std::numeric_limits<int32_t>::max() ||
        n > std::numeric_limits<int64_t>::max() / F / 4)
        fail(SCL_ERROR_INVALID_ARGUMENT, "n_nodes * n_features overflows");
}

int64_t n_checkpoints(int64_t t, int64_t every) { return (t + every - 1) / every - 1; }

CsrView fwd_view(const DeviceCSR& d) {
    return CsrView{d.n, d.n_sec, d.nnz, d.row_ptr, d.col, d.val, d.sched, &d.merge,
                   d.grouped, &d.staged, d.tiled};
}
CsrView adj_view(const DeviceCSR& d) {
    return CsrView{d.n_sec, d.n, d.nnz, d.t_row_ptr, d.t_col, d.t_val, d.t_sched,
                   &d.merge};
}

SpmmShape fwd_shape(DeviceCSR& d, int64_t F, const void* a, const void* b) {
    auto it = d.shapes.find(F);
    if (it != d.shapes.end() && spmm_shape_valid(it->second, F, a, b)) return it->second;
    return spmm_default_shape(F, a, b);
}

size_t backward_bytes(int64_t n, int64_t F, int64_t nnz, int64_t every, bool want_dp,
                      bool values = false) {
    const size_t plane = round_up((size_t)n * (size_t)F * sizeof(float));
    size_t b = plane * (size_t)(2 + (every - 1));
    if (want_dp) b += round_up((size_t)nnz * sizeof(float));
    // transposed values, with the pad the adjoint kernels read past the end
    if (values) b += round_up(((size_t)nnz + kTransposePad) * sizeof(float));
    return b;
}

// A forward view over the operator's pattern with the caller's values
// (stored order).  The grouped and tiled forms hold copies of the handle's
// values, so a shape that reads one falls back to the default.
void with_values(CsrView& v, SpmmShape& sh, const float* values, int64_t F, const void* a,
                 const void* b) {
    v.val = values;
    v.grouped = nullptr;
    v.tiled = nullptr;
    if (sh.rows_per_group && sh.rows_per_group != kStagedRows) sh = spmm_default_shape(F, a, b);
}

}  // namespace

extern "C" {

SCL_EXPORT scl_error_t scl_gpu_sparse_set_values_device(scl_sparse_t matrix,
                                                        const scl_real_t* values_dev,
                                                        void* stream) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK(values_dev != nullptr || matrix->nnz == 0, SCL_ERROR_NULL_POINTER,
                  "values is null");
    SCL_API_TRY
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        std::lock_guard<std::mutex> lk(matrix->mu);
        DeviceCSR& d = device_csr(matrix, dev);
        const bool keep_t = d.t_row_ptr && d.t_generation == d.generation;
        if (matrix->nnz > 0) {
            hip_check(hipMemcpyAsync(d.val, values_dev, (size_t)matrix->nnz * 4,
                                     hipMemcpyDeviceToDevice, s),
                      "copy values");
            if (keep_t) launch_gather_f32(d.t_src, d.val, d.t_val, matrix->nnz, s);
            refresh_grouped_values(d, s);
            download_host(matrix->data, d.val, (size_t)matrix->nnz * 4, s, dev);
            hip_check(hipStreamSynchronize(s), "set values");
        }
        matrix->generation++;
        d.generation = matrix->generation;
        if (keep_t) d.t_generation = d.generation;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_edge_gradient_device(
    scl_sparse_t transition, const scl_real_t* G_dev, const scl_real_t* Xprev_dev,
    scl_real_t* dP_dev, scl_index_t n_nodes, scl_index_t n_features, int accumulate,
    void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(G_dev && Xprev_dev && dP_dev, SCL_ERROR_NULL_POINTER,
                  "G, Xprev and dP must be set");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        check_square(transition, n_nodes, n_features);
        const int dev = current_device();
        CsrView v;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            v = fwd_view(device_csr(transition, dev));
        }
        launch_edge_grad(v, G_dev, Xprev_dev, dP_dev, accumulate != 0, n_features,
                         static_cast<hipStream_t>(stream));
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_forward_ckpt_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_real_t* ckpt_dev, scl_real_t* work_dev, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream) {
    return scl_gpu_diffuse_forward_ckpt_values_device(transition, nullptr, X_dev, Y_dev,
                                                      ckpt_dev, work_dev, n_nodes,
                                                      n_features, n_steps, every, stream);
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_forward_ckpt_values_device(
    scl_sparse_t transition, const scl_real_t* values_dev, const scl_real_t* X_dev,
    scl_real_t* Y_dev, scl_real_t* ckpt_dev, scl_real_t* work_dev, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(X_dev && Y_dev, SCL_ERROR_NULL_POINTER, "X or Y is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0 && every > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_CHECK(n_checkpoints(n_steps, every) == 0 || ckpt_dev != nullptr,
                  SCL_ERROR_NULL_POINTER, "checkpoint buffer is required");
    SCL_API_CHECK(n_steps == 1 || work_dev != nullptr, SCL_ERROR_NULL_POINTER,
                  "work buffer is required for n_steps >= 2");
    SCL_API_TRY
        check_square(transition, n_nodes, n_features);
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        CsrView v;
        SpmmShape sh;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR& d = device_csr(transition, dev);
            v = fwd_view(d);
            sh = fwd_shape(d, n_features, X_dev, Y_dev);
            if (values_dev) with_values(v, sh, values_dev, n_features, X_dev, Y_dev);
            if (!spmm_shape_valid(sh, n_features, ckpt_dev ? ckpt_dev : Y_dev, work_dev ? work_dev : Y_dev))
                sh = spmm_default_shape(n_features, X_dev, Y_dev);
        }
        const size_t plane = (size_t)n_nodes * (size_t)n_features;
        float* tmp[2] = {work_dev, work_dev ? work_dev + plane : nullptr};
        const float* cur = X_dev;
        for (int64_t st = 1; st <= n_steps; ++st) {
            float* dst;
            if (st == n_steps) dst = Y_dev;
            else if (st % every == 0) dst = ckpt_dev + (size_t)(st / every - 1) * plane;
            else dst = tmp[st % 2];
            launch_spmm_f32(v, cur, dst, n_features, sh, s);
            cur = dst;
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_workspace(
    scl_sparse_t transition, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t every, int want_dp, scl_size_t* bytes) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(bytes, "Output pointer is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && every > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_TRY
        check_square(transition, n_nodes, n_features);
        *bytes = backward_bytes(n_nodes, n_features, transition->nnz, every, want_dp != 0);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_values_workspace(
    scl_sparse_t transition, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t every, int want_dp, scl_size_t* bytes) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(bytes, "Output pointer is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && every > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_TRY
        check_square(transition, n_nodes, n_features);
        *bytes = backward_bytes(n_nodes, n_features, transition->nnz, every, want_dp != 0,
                                true);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_device(
    scl_sparse_t transition, const scl_real_t* X_dev, const scl_real_t* ckpt_dev,
    const scl_real_t* dY_dev, scl_real_t* dX_dev, scl_real_t* dP_dev,
    void* work_dev, scl_size_t work_bytes, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream) {
    return scl_gpu_diffuse_backward_values_device(transition, nullptr, X_dev, ckpt_dev, dY_dev,
                                                  dX_dev, dP_dev, work_dev, work_bytes,
                                                  n_nodes, n_features, n_steps, every,
                                                  stream);
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_values_device(
    scl_sparse_t transition, const scl_real_t* values_dev, const scl_real_t* X_dev,
    const scl_real_t* ckpt_dev, const scl_real_t* dY_dev, scl_real_t* dX_dev,
    scl_real_t* dP_dev, void* work_dev, scl_size_t work_bytes, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(X_dev && dY_dev && dX_dev && work_dev, SCL_ERROR_NULL_POINTER,
                  "X, dY, dX and work must be set");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0 && every > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_CHECK(n_checkpoints(n_steps, every) == 0 || ckpt_dev != nullptr,
                  SCL_ERROR_NULL_POINTER, "checkpoint buffer is required");
    SCL_API_CHECK((const void*)dX_dev != (const void*)dY_dev, SCL_ERROR_INVALID_ARGUMENT,
                  "dX and dY must be distinct");
    SCL_API_TRY
        check_square(transition, n_nodes, n_features);
        const int64_t n = n_nodes, F = n_features, t = n_steps;
        const bool want_dp = dP_dev != nullptr && transition->nnz > 0;
        const bool own_values = values_dev != nullptr && transition->nnz > 0;
        if (work_bytes < backward_bytes(n, F, transition->nnz, every, want_dp, own_values))
            fail(SCL_ERROR_INVALID_ARGUMENT,
                 own_values ? "work buffer smaller than scl_gpu_diffuse_backward_values_workspace"
                            : "work buffer smaller than scl_gpu_diffuse_backward_workspace");
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        CsrView fv, av;
        SpmmShape sh, adj;
        bool has_adj = false;
        const int32_t* t_src = nullptr;
        const int32_t* t_dst = nullptr;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR& d = device_csr(transition, dev);
            ensure_transpose(transition, d, s);
            fv = fwd_view(d);
            av = adj_view(d);
            t_src = d.t_src;
            t_dst = d.t_dst;
            sh = fwd_shape(d, F, X_dev, work_dev);
            const auto ai = d.adj_shapes.find(F);
            if (ai != d.adj_shapes.end()) {
                adj = ai->second;
                has_adj = true;
            }
        }
        const SpmmShape* adj_sh = has_adj ? &adj : nullptr;
        const size_t plane = (size_t)n * (size_t)F;
        const size_t plane_b = round_up(plane * sizeof(float));
        char* w = static_cast<char*>(work_dev);
        float* gbuf[2] = {reinterpret_cast<float*>(w), reinterpret_cast<float*>(w + plane_b)};
        // recompute slots R[1..every-1]
        auto rslot = [&](int64_t r) {
            return reinterpret_cast<float*>(w + plane_b * (size_t)(1 + r));
        };
        char* tail = w + plane_b * (size_t)(1 + every);
        float* dPT = nullptr;
        if (want_dp) {
            dPT = reinterpret_cast<float*>(tail);
            tail += round_up((size_t)transition->nnz * sizeof(float));
        }
        if (own_values) {
            // the caller's values in transposed order for the adjoint steps
            float* tv = reinterpret_cast<float*>(tail);
            launch_gather_f32(t_src, values_dev, tv, transition->nnz, s);
            av.val = tv;
            with_values(fv, sh, values_dev, F, gbuf[0], rslot(1));
        }
        if (!spmm_shape_valid(sh, F, gbuf[0], rslot(1)))
            sh = spmm_default_shape(F, gbuf[0], rslot(1));
        const int64_t M = (t + every - 1) / every;
        const float* G = dY_dev;
        bool first = true;
        for (int64_t m = M - 1; m >= 0; --m) {
            const int64_t a = m * every, b = std::min<int64_t>(a + every, t);
            const float* base = (a == 0) ? X_dev : ckpt_dev + (size_t)(m - 1) * plane;
            // iterates X_{a+1} .. X_{b-1} of this segment
            const float* prev = base;
            for (int64_t r = 1; r < b - a; ++r) {
                launch_spmm_f32(fv, prev, rslot(r), F, sh, s);
                prev = rslot(r);
            }
            for (int64_t st = b; st > a; --st) {
                const float* Xprev = (st - 1 == a) ? base : rslot(st - 1 - a);
                float* Gn = (st == 1) ? dX_dev : gbuf[st % 2];
                launch_adjoint_edge(av, G, want_dp ? Xprev : nullptr, Gn, dPT, !first, F, s,
                                    adj_sh);
                first = false;
                G = Gn;
            }
        }
        // dP in CSR order: a gather (coalesced stores) through the inverse map
        if (want_dp) launch_gather_f32(t_dst, dPT, dP_dev, transition->nnz, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_sparse_generation(scl_sparse_t matrix, uint64_t* out) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_TRY
        std::lock_guard<std::mutex> lk(matrix->mu);
        *out = matrix->generation;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
