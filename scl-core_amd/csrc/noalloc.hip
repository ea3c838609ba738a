// noalloc.hip -- Y = P^t X on device buffers without any allocation: the
// variant of scl_gpu_diffuse_matrix_device whose intermediate iterate (the
// caller's work buffer there) lives in a fixed device array of this file.
//
// The array holds kScratchFloats floats on every GPU the library runs on,
// from the first call on that GPU until the process ends.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is handed to kernels: in host code
// the array's own name is not a device address.  Every call on one GPU
// shares that GPU's array, so calls that overlap on one GPU (two streams,
// two threads) race; nothing in it needs to start at zero (each step writes
// every element it later reads).
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "kernels.hpp"
#include "scl_diffusion_gpu.h"

#include <map>
#include <mutex>

using namespace scl_amd;

namespace {

constexpr size_t kScratchFloats = SCL_GPU_NOALLOC_SCRATCH_FLOATS;

static __device__ __align__(256) float scratch[kScratchFloats];

std::mutex g_scratch_mu;
std::map<int, float*> g_scratch;  // device -> address of its copy of scratch

// Address of the current device's copy of `scratch` (looked up on the first
// call on each device).
float* scratch_address(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    const auto it = g_scratch.find(dev);
    if (it != g_scratch.end()) return it->second;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
    if (!p) fail(SCL_ERROR_INTERNAL, "scratch address is null");
    g_scratch[dev] = static_cast<float*>(p);
    return static_cast<float*>(p);
}

}  // namespace

extern "C" {

SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_device_noalloc(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_index_t n_nodes, scl_index_t n_features, scl_index_t n_steps, void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(X_dev, "X is null");
    SCL_API_CHECK_NULL(Y_dev, "Y is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_CHECK((const void*)X_dev != (const void*)Y_dev, SCL_ERROR_INVALID_ARGUMENT,
                  "X and Y must be distinct");
    SCL_API_CHECK(n_steps == 1 || (n_nodes <= (scl_index_t)kScratchFloats / n_features &&
                                   n_nodes * n_features <= (scl_index_t)kScratchFloats),
                  SCL_ERROR_INVALID_ARGUMENT,
                  "n_nodes * n_features exceeds the allocation-free scratch "
                  "(SCL_GPU_NOALLOC_SCRATCH_FLOATS floats) for n_steps >= 2");
    SCL_API_TRY
        if (transition->rows != transition->cols || transition->rows != n_nodes)
            fail(SCL_ERROR_DIMENSION_MISMATCH,
                 "n_nodes does not match the transition matrix dimensions");
        const int dev = current_device();
        float* work = n_steps >= 2 ? scratch_address(dev) : nullptr;
        CsrView v;
        SpmmShape sh;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR* d = device_csr_current(transition, dev);
            if (!d)
                fail(SCL_ERROR_INVALID_ARGUMENT,
                     "operator not on this device or changed since: call "
                     "scl_gpu_sparse_prepare (which allocates) first");
            v = CsrView{d->n, d->n_sec, d->nnz, d->row_ptr, d->col, d->val, d->sched,
                        &d->merge, d->grouped, &d->staged, d->tiled};
            const auto it = d->shapes.find(n_features);
            sh = it != d->shapes.end() ? it->second
                                       : spmm_default_shape(n_features, X_dev, Y_dev);
            const float* bufs[3] = {X_dev, Y_dev, work ? work : Y_dev};
            for (const float* a : bufs)
                for (const float* b : bufs)
                    if (!spmm_shape_valid(sh, n_features, a, b))
                        sh = spmm_default_shape(n_features, X_dev, Y_dev);
        }
        hipStream_t s = static_cast<hipStream_t>(stream);
        // step k (1-based) writes Y when t - k is even, else the scratch, so
        // the last step lands in Y and consecutive steps alternate buffers
        const float* src = X_dev;
        for (int64_t k = 1; k <= n_steps; ++k) {
            float* dst = ((n_steps - k) % 2 == 0) ? Y_dev : work;
            launch_spmm_f32(v, src, dst, n_features, sh, s);
            src = dst;
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
