// construct.cpp -- entry points that build and inspect the diffusion operator:
// scl_kernel_compute_matrix / scl_kernel_row_sums (kNN distances to kernel
// weights, scl/binding/c_api/kernel.cpp:145-197) and the scl_transition_*
// functions (scl/binding/c_api/transition.cpp:29-150).
//
// Host arrays are copied to the GPU once and results copied back once; the
// dense n x n outputs of symmetrize and matrix_from_velocity are formed on
// the device.  Checks beyond the reference: dimensions that the reference
// would read or write past are SCL_ERROR_DIMENSION_MISMATCH.
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_kernel.h"
#include "scl_transition.h"
#include "session.hpp"

#include <cmath>
#include <limits>
#include <mutex>
#include <utility>

using namespace scl_amd;

namespace {

// The handle's stored values on the device (all of them, stored order).
struct HostValues {
    int64_t nnz;
    DevBuf v;
    HostValues(scl_sparse_matrix* h, hipStream_t s) : nnz(count(h)), v(nnz) {
        std::lock_guard<std::mutex> lk(h->mu);
        if (nnz > 0) upload_host(v.get(), h->data, (size_t)nnz * 4, s, current_device());
    }
    static int64_t count(scl_sparse_matrix* h) {
        std::lock_guard<std::mutex> lk(h->mu);
        return h->indptr[h->primary()] - h->indptr[0];
    }
};

void download(float* dst, const float* src, int64_t m, hipStream_t s) {
    if (m > 0) download_host(dst, src, (size_t)m * 4, s, current_device());
    hip_check(hipStreamSynchronize(s), "download");
}

// square n x n operator, and a dense n x n result that fits
void check_square(const scl_sparse_matrix* T, int64_t n) {
    if (n < 0) fail(SCL_ERROR_INVALID_ARGUMENT, "n must not be negative");
    if (T->rows != T->cols || T->rows != n)
        fail(SCL_ERROR_DIMENSION_MISMATCH, "the matrix must be n x n");
}

void check_dense_square(int64_t n) {
    if (n > 0 && n > std::numeric_limits<int64_t>::max() / n / 4)
        fail(SCL_ERROR_INVALID_ARGUMENT, "n x n output too large");
}

}  // namespace

extern "C" {

// kernel.cpp:145-163
SCL_EXPORT scl_error_t scl_kernel_compute_matrix(scl_sparse_t distances, scl_real_t* kernel_values,
                                                 scl_real_t bandwidth,
                                                 scl_kernel_type_t kernel_type) {
    SCL_API_CHECK_NULL(distances, "Distances matrix handle is null");
    SCL_API_CHECK_NULL(kernel_values, "Kernel values output pointer is null");
    SCL_API_TRY
        const int dev = current_device();
        hipStream_t s = library_stream(dev);
        const HostValues d(distances, s);
        DevBuf out(d.nnz);
        launch_kernel_values(d.v.get(), d.nnz, kernel_weights(bandwidth, (int)kernel_type),
                             out.get(), s);
        download(kernel_values, out.get(), d.nnz, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// kernel.cpp:169-197
SCL_EXPORT scl_error_t scl_kernel_row_sums(scl_sparse_t distances, scl_real_t* sums,
                                           scl_size_t n_points, scl_real_t bandwidth,
                                           scl_kernel_type_t kernel_type) {
    SCL_API_CHECK_NULL(distances, "Distances matrix handle is null");
    SCL_API_CHECK_NULL(sums, "Sums output pointer is null");
    SCL_API_CHECK(n_points > 0, SCL_ERROR_INVALID_ARGUMENT, "Number of points must be positive");
    SCL_API_TRY
        if (distances->rows != (int64_t)n_points)
            fail(SCL_ERROR_DIMENSION_MISMATCH, "Point count mismatch");
        // kernel_row_sums writes one sum per primary slot (kernel.hpp:605)
        const int64_t p = distances->primary();
        if (p > (int64_t)n_points)
            fail(SCL_ERROR_DIMENSION_MISMATCH, "KernelRowSums: output buffer too small");
        const int dev = current_device();
        hipStream_t s = library_stream(dev);
        const HostRows d(distances, p, s);
        DevBuf out(p);
        launch_kernel_row_sums(d.ip.as<int64_t>(), d.v.get(), p,
                               kernel_weights(bandwidth, (int)kernel_type), out.get(), s);
        download(sums, out.get(), p, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// transition.cpp:29-55
SCL_EXPORT scl_error_t scl_transition_matrix_from_velocity(scl_sparse_t velocity_graph,
                                                           scl_index_t n,
                                                           scl_real_t* row_stochastic_out) {
    SCL_API_CHECK_NULL(velocity_graph, "Velocity graph is null");
    SCL_API_CHECK_NULL(row_stochastic_out, "Output is null");
    SCL_API_TRY
        check_square(velocity_graph, n);
        check_dense_square(n);
        const int dev = current_device();
        hipStream_t s = library_stream(dev);
        const HostPattern g(velocity_graph, n, s);
        DevBuf S(n * n);
        launch_fill(S.get(), n * n, 0.0f, s);
        launch_from_velocity(g.ip.as<int64_t>(), g.ix.as<int64_t>(), g.v.get(), n, S.get(), s);
        download(row_stochastic_out, S.get(), n * n, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// transition.cpp:57-87
SCL_EXPORT scl_error_t scl_transition_row_normalize_to_stochastic(scl_sparse_t input,
                                                                  scl_index_t n,
                                                                  scl_real_t* output_values,
                                                                  scl_size_t output_size) {
    SCL_API_CHECK_NULL(input, "Input matrix is null");
    SCL_API_CHECK_NULL(output_values, "Output is null");
    SCL_API_TRY
        if (n < 0) fail(SCL_ERROR_INVALID_ARGUMENT, "n must not be negative");
        if (n > input->primary())
            fail(SCL_ERROR_DIMENSION_MISMATCH, "n exceeds the matrix's rows");
        int64_t need;
        {
            std::lock_guard<std::mutex> lk(input->mu);
            need = input->indptr[n] - input->indptr[0];
        }
        if ((int64_t)output_size < need)
            fail(SCL_ERROR_DIMENSION_MISMATCH, "output_values smaller than the rows' entries");
        const int dev = current_device();
        hipStream_t s = library_stream(dev);
        const HostRows m(input, n, s);
        DevBuf out(need);
        launch_row_normalize(m.ip.as<int64_t>(), m.v.get(), n, out.get(), s);
        download(output_values, out.get(), need, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// transition.cpp:89-115
SCL_EXPORT scl_error_t scl_transition_symmetrize(scl_sparse_t transition_mat, scl_index_t n,
                                                 scl_real_t* symmetric_out) {
    SCL_API_CHECK_NULL(transition_mat, "Transition matrix is null");
    SCL_API_CHECK_NULL(symmetric_out, "Output is null");
    SCL_API_TRY
        check_square(transition_mat, n);
        check_dense_square(n);
        const int dev = current_device();
        hipStream_t s = library_stream(dev);
        const HostPattern t(transition_mat, n, s);
        DevBuf S(n * n);
        launch_fill(S.get(), n * n, 0.0f, s);
        launch_symmetrize(t.ip.as<int64_t>(), t.ix.as<int64_t>(), t.v.get(), n, t.serial,
                          S.get(), s);
        download(symmetric_out, S.get(), n * n, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// transition.cpp:117-150 (stationary_distribution, transition.hpp:571-623)
SCL_EXPORT scl_error_t scl_transition_stationary_distribution(scl_sparse_t transition_mat,
                                                              scl_index_t n,
                                                              scl_real_t* stationary,
                                                              scl_index_t max_iter,
                                                              scl_real_t tolerance) {
    SCL_API_CHECK_NULL(transition_mat, "Transition matrix is null");
    SCL_API_CHECK_NULL(stationary, "Output is null");
    SCL_API_TRY
        check_square(transition_mat, n);
        if (n == 0) SCL_API_RETURN_OK;
        Session S(transition_mat, n);
        ensure_transpose(transition_mat, *S.d, S.s);
        const DeviceCSR& d = *S.d;
        const CsrView tv{d.n_sec, d.n, d.nnz, d.t_row_ptr, d.t_col, d.t_val, d.t_sched, &d.merge};
        DevBuf pi(n), y(n), part(1024), scal(2);
        launch_fill(pi.get(), n, 1.0f / (float)n, S.s);
        SpmmShape sh = spmm_default_shape(1, pi.get(), y.get());
        {
            const auto it = d.t_shapes.find(1);
            if (it != d.t_shapes.end() && spmm_shape_valid(it->second, 1, pi.get(), y.get()))
                sh = it->second;
        }
        float* cur = pi.get();
        float* nxt = y.get();
        float d1 = 1e30f, d2 = 1e30f;  // INF_VALUE, transition.hpp:45
        for (int64_t iter = 0; iter < max_iter; ++iter) {
            launch_spmm_f32(tv, cur, nxt, 1, sh, S.s);
            launch_stationary_update(nxt, cur, n, scal.get(), part.get(), scal.get() + 1, S.s);
            const float diff = S.scalar(scal.get() + 1);
            std::swap(cur, nxt);
            if (diff < tolerance) break;
            if (iter >= 2 && d1 < d2 && diff < d1) {
                const float den = diff - 2.0f * d1 + d2;
                if (std::fabs(den) > 1e-15f && (diff - d1) / den > 0.9f) break;
            }
            d2 = d1;
            d1 = diff;
        }
        S.download(stationary, cur, n);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
