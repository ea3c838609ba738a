// lowp.cpp -- the bf16-iterate diffusion entry point (scl_diffusion_gpu.h).
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_diffusion_gpu.h"

#include <limits>

using namespace scl_amd;

extern "C" {

SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_bf16_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    void* work_dev, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t n_steps, int flags, void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(X_dev && Y_dev, SCL_ERROR_NULL_POINTER, "X or Y is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_CHECK(n_features % 2 == 0, SCL_ERROR_INVALID_ARGUMENT,
                  "bf16 storage needs an even n_features");
    SCL_API_CHECK(n_steps == 1 || work_dev != nullptr, SCL_ERROR_NULL_POINTER,
                  "work buffer is required for n_steps >= 2");
    SCL_API_CHECK((const void*)X_dev != (const void*)Y_dev, SCL_ERROR_INVALID_ARGUMENT,
                  "X and Y must be distinct");
    SCL_API_TRY
        if (transition->rows != transition->cols || transition->rows != n_nodes)
            fail(SCL_ERROR_DIMENSION_MISMATCH,
                 "n_nodes does not match the transition matrix dimensions");
        if (n_features > std::numeric_limits<int32_t>::max())
            fail(SCL_ERROR_INVALID_ARGUMENT, "n_features too large");
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        CsrView v;
        const uint16_t* vb = nullptr;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR& d = device_csr(transition, dev);
            if (flags & SCL_GPU_BF16_VALUES) {
                if (!d.val_bf16) {
                    void* p = nullptr;
                    hip_check(hipMalloc(&p, (size_t)std::max<int64_t>(d.nnz, 1) * 2),
                              "bf16 values");
                    d.val_bf16 = static_cast<uint16_t*>(p);
                    d.bf16_generation = ~uint64_t(0);
                }
                if (d.bf16_generation != d.generation) {
                    launch_to_bf16(d.val, d.val_bf16, d.nnz, s);
                    d.bf16_generation = d.generation;
                }
                vb = d.val_bf16;
            }
            v = CsrView{d.n, d.n_sec, d.nnz, d.row_ptr, d.col, d.val, d.sched};
        }
        const bool sched = (flags & SCL_GPU_BF16_NO_SCHEDULE) == 0;
        const size_t plane = (size_t)n_nodes * (size_t)n_features;
        uint16_t* w[2] = {static_cast<uint16_t*>(work_dev),
                          work_dev ? static_cast<uint16_t*>(work_dev) + plane : nullptr};
        if (n_steps == 1) {
            launch_spmm_lowp(v, vb, LowpIo::F32_TO_F32, X_dev, Y_dev, n_features, sched, s);
        } else {
            launch_spmm_lowp(v, vb, LowpIo::F32_TO_BF16, X_dev, w[1], n_features, sched, s);
            int cur = 1;
            for (int64_t st = 2; st < n_steps; ++st) {
                launch_spmm_lowp(v, vb, LowpIo::BF16_TO_BF16, w[cur], w[cur ^ 1],
                                 n_features, sched, s);
                cur ^= 1;
            }
            launch_spmm_lowp(v, vb, LowpIo::BF16_TO_F32, w[cur], Y_dev, n_features, sched, s);
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
