// lowp.cpp -- the reduced-precision-iterate diffusion entry points
// (scl_diffusion_gpu.h): bf16 or fp16 intermediate iterates, fp32 in and out.
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_diffusion_gpu.h"

#include <cmath>
#include <limits>

using namespace scl_amd;

namespace {

void check_lowp(const scl_sparse_matrix* P, int64_t n, int64_t F) {
    if (P->rows != P->cols || P->rows != n)
        fail(SCL_ERROR_DIMENSION_MISMATCH,
             "n_nodes does not match the transition matrix dimensions");
    if (F > std::numeric_limits<int32_t>::max())
        fail(SCL_ERROR_INVALID_ARGUMENT, "n_features too large");
}

// Bytes of one compensated plane of m values (3 bytes each), 256-aligned.
size_t comp_plane_bytes(size_t m) { return (m * 3 + 255) / 256 * 256; }

CompPlanes comp_planes(void* base, size_t m) {
    CompPlanes c;
    c.hi = static_cast<uint16_t*>(base);
    c.lo = reinterpret_cast<uint8_t*>(static_cast<char*>(base) + m * 2);
    return c;
}

// Y = P^t X with compensated-bf16 storage (lowp.hip): the iterates, and X
// with SCL_GPU_LOWP_ROUND_INPUT, and P's values with SCL_GPU_LOWP_BF16_VALUES;
// fp32 accumulation, fp32 output.
void diffuse_comp(scl_sparse_matrix* P, const float* X, float* Y, void* work, int64_t n,
                  int64_t F, int64_t t, int flags, hipStream_t s) {
    const bool round_input = (flags & SCL_GPU_LOWP_ROUND_INPUT) != 0;
    if ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y)) % 16 != 0)
        fail(SCL_ERROR_INVALID_ARGUMENT, "compensated bf16 needs 16-byte aligned X and Y");
    const int dev = current_device();
    CsrView v;
    const CompPlanes* vals = nullptr;
    CompPlanes vc;
    bool sched = (flags & SCL_GPU_LOWP_NO_SCHEDULE) == 0;
    {
        std::lock_guard<std::mutex> lk(P->mu);
        DeviceCSR& d = device_csr(P, dev);
        const auto ti = d.shapes.find(F);
        if (ti != d.shapes.end() && !ti->second.schedule) sched = false;
        if (flags & SCL_GPU_LOWP_BF16_VALUES) {
            if (!d.val_comp.hi) {
                void* p = nullptr;
                hip_check(hipMalloc(&p, comp_plane_bytes((size_t)std::max<int64_t>(d.nnz, 1))),
                          "compensated values");
                d.val_comp = comp_planes(p, (size_t)std::max<int64_t>(d.nnz, 1));
                d.comp_generation = ~uint64_t(0);
            }
            if (d.comp_generation != d.generation) {
                launch_encode_comp(d.val, d.val_comp, d.nnz, s);
                d.comp_generation = d.generation;
            }
            vc = d.val_comp;
            vals = &vc;
        }
        v = CsrView{d.n, d.n_sec, d.nnz, d.row_ptr, d.col, d.val, d.sched};
    }
    const size_t m = (size_t)n * (size_t)F;
    uint32_t* w[2] = {nullptr, nullptr};
    if (work) {
        w[0] = static_cast<uint32_t*>(work);
        w[1] = reinterpret_cast<uint32_t*>(static_cast<char*>(work) + comp_plane_bytes(m));
    }
    if (round_input) {
        launch_pack_comp(X, w[0], (int64_t)m, s);
        int cur = 0;
        for (int64_t st = 1; st < t; ++st) {
            launch_spmm_comp(v, vals, nullptr, w[cur], nullptr, w[cur ^ 1], F, sched, s);
            cur ^= 1;
        }
        launch_spmm_comp(v, vals, nullptr, w[cur], Y, nullptr, F, sched, s);
    } else if (t == 1) {
        launch_spmm_comp(v, vals, X, nullptr, Y, nullptr, F, sched, s);
    } else {
        launch_spmm_comp(v, vals, X, nullptr, nullptr, w[1], F, sched, s);
        int cur = 1;
        for (int64_t st = 2; st < t; ++st) {
            launch_spmm_comp(v, vals, nullptr, w[cur], nullptr, w[cur ^ 1], F, sched, s);
            cur ^= 1;
        }
        launch_spmm_comp(v, vals, nullptr, w[cur], Y, nullptr, F, sched, s);
    }
}

}  // namespace

extern "C" {

SCL_EXPORT scl_error_t scl_gpu_lowp_workspace(scl_index_t n_nodes, scl_index_t n_features,
                                              int flags, scl_size_t* bytes) {
    SCL_API_CHECK_NULL(bytes, "Output pointer is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    const size_t m = (size_t)n_nodes * (size_t)n_features;
    *bytes = (flags & SCL_GPU_LOWP_COMPENSATED) ? 2 * comp_plane_bytes(m) : m * 4;
    SCL_API_RETURN_OK;
}

SCL_EXPORT scl_error_t scl_gpu_lowp_scale_exp(scl_sparse_t transition,
                                              const scl_real_t* X_dev,
                                              scl_index_t n_nodes,
                                              scl_index_t n_features,
                                              scl_index_t n_steps, int* scale_exp,
                                              void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(X_dev && scale_exp, SCL_ERROR_NULL_POINTER, "X or output is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_TRY
        check_lowp(transition, n_nodes, n_features);
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        // largest row |sum| of P (host copy is current)
        double rho = 0.0;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            const int64_t p = transition->primary();
            for (int64_t i = 0; i < p; ++i) {
                double r = 0.0;
                for (int64_t k = transition->indptr[i]; k < transition->indptr[i + 1]; ++k)
                    r += std::fabs((double)transition->data[k]);
                rho = std::max(rho, r);
            }
        }
        constexpr int kParts = 1024;
        float* part = static_cast<float*>(device_workspace(dev, 3, kParts * sizeof(float)));
        launch_absmax(X_dev, n_nodes * n_features, part, kParts, s);
        float host[kParts];
        hip_check(hipMemcpyAsync(host, part, sizeof(host), hipMemcpyDeviceToHost, s),
                  "absmax download");
        hip_check(hipStreamSynchronize(s), "absmax");
        double mx = 0.0;
        for (float v : host) mx = std::max(mx, (double)v);
        // every intermediate iterate X_s (1 <= s < t) obeys |X_s| <= rho^s max|X|
        double bound = mx;
        for (int64_t st = 1; st < n_steps; ++st) bound = std::max(bound, mx * std::pow(rho, (double)st));
        int k = 0;
        if (bound > 0.0 && std::isfinite(bound)) {
            k = (int)std::floor(std::log2(32768.0 / bound));
            k = std::max(-60, std::min(60, k));
        }
        *scale_exp = k;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_lowp_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    void* work_dev, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t n_steps, int flags, int scale_exp, void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(X_dev && Y_dev, SCL_ERROR_NULL_POINTER, "X or Y is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_CHECK(n_features % 2 == 0, SCL_ERROR_INVALID_ARGUMENT,
                  "16-bit storage needs an even n_features");
    const bool round_input = (flags & SCL_GPU_LOWP_ROUND_INPUT) != 0;
    SCL_API_CHECK((n_steps == 1 && !round_input) || work_dev != nullptr,
                  SCL_ERROR_NULL_POINTER,
                  "work buffer is required for n_steps >= 2 or a rounded input");
    SCL_API_CHECK((const void*)X_dev != (const void*)Y_dev, SCL_ERROR_INVALID_ARGUMENT,
                  "X and Y must be distinct");
    const bool fp16 = (flags & SCL_GPU_LOWP_FP16) != 0;
    const bool comp = (flags & SCL_GPU_LOWP_COMPENSATED) != 0;
    SCL_API_CHECK(scale_exp >= -60 && scale_exp <= 60 && (fp16 || scale_exp == 0),
                  SCL_ERROR_INVALID_ARGUMENT,
                  "scale_exp must be 0 for bf16 and within [-60, 60] for fp16");
    SCL_API_CHECK(!comp || (!fp16 && n_features % 4 == 0), SCL_ERROR_INVALID_ARGUMENT,
                  "compensated bf16 excludes fp16 and needs n_features % 4 == 0");
    if (comp) {
        SCL_API_TRY
            check_lowp(transition, n_nodes, n_features);
            diffuse_comp(transition, X_dev, Y_dev, work_dev, n_nodes, n_features, n_steps,
                         flags, static_cast<hipStream_t>(stream));
            SCL_API_RETURN_OK;
        SCL_API_CATCH
    }
    SCL_API_TRY
        check_lowp(transition, n_nodes, n_features);
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        CsrView v;
        const uint16_t* vb = nullptr;
        bool sched = (flags & SCL_GPU_LOWP_NO_SCHEDULE) == 0;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR& d = device_csr(transition, dev);
            // the visiting order the fp32 SpMM was tuned to for this width
            const auto ti = d.shapes.find(n_features);
            if (ti != d.shapes.end() && !ti->second.schedule) sched = false;
            if (flags & SCL_GPU_LOWP_BF16_VALUES) {
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
        const float up = std::ldexp(1.0f, scale_exp), down = std::ldexp(1.0f, -scale_exp);
        const size_t plane = (size_t)n_nodes * (size_t)n_features;
        uint16_t* w[2] = {static_cast<uint16_t*>(work_dev),
                          work_dev ? static_cast<uint16_t*>(work_dev) + plane : nullptr};
        const int64_t F = n_features;
        if (round_input) {
            // X itself is stored in 16 bits once; every step then gathers
            // 16-bit rows
            launch_narrow_f32(X_dev, w[0], (int64_t)plane, fp16, up, s);
            int cur = 0;
            for (int64_t st = 1; st < n_steps; ++st) {
                launch_spmm_lowp(v, vb, LowpIo::BF16_TO_BF16, w[cur], w[cur ^ 1], F, sched,
                                 fp16, down, up, s);
                cur ^= 1;
            }
            launch_spmm_lowp(v, vb, LowpIo::BF16_TO_F32, w[cur], Y_dev, F, sched, fp16, down,
                             1.0f, s);
        } else if (n_steps == 1) {
            launch_spmm_lowp(v, vb, LowpIo::F32_TO_F32, X_dev, Y_dev, F, sched, fp16, 1.0f,
                             1.0f, s);
        } else {
            launch_spmm_lowp(v, vb, LowpIo::F32_TO_BF16, X_dev, w[1], F, sched, fp16, 1.0f,
                             up, s);
            int cur = 1;
            for (int64_t st = 2; st < n_steps; ++st) {
                launch_spmm_lowp(v, vb, LowpIo::BF16_TO_BF16, w[cur], w[cur ^ 1], F, sched,
                                 fp16, down, up, s);
                cur ^= 1;
            }
            launch_spmm_lowp(v, vb, LowpIo::BF16_TO_F32, w[cur], Y_dev, F, sched, fp16, down,
                             1.0f, s);
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
