// neighbors.cpp -- exact k-nearest-neighbour search on the GPU:
// scl_neighbors_compute_norms, scl_knn and scl_knn_simple
// (scl/binding/c_api/neighbors.cpp:20-128 over
// scl/kernel/neighbors.hpp:300-476).
//
// The samples are the handle's primary slots.  Their stored values are
// copied to the GPU once, expanded to a zero-padded dense matrix there, and
// every sample is compared with every other by the tiled scan of
// neighbors.hip; the lists come back sorted by (distance, index).
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_neighbors.h"
#include "session.hpp"

#include <cstdint>
#include <limits>

using namespace scl_amd;

namespace {

void download_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes > 0) download_host(dst, src, bytes, s, current_device());
    hip_check(hipStreamSynchronize(s), "download");
}

void knn_impl(scl_sparse_matrix* m, const float* host_norms, int64_t n_samples, int64_t k,
              int64_t* out_idx, float* out_dist) {
    const int64_t N = m->primary(), F = m->secondary();
    // the reference writes N rows of k entries and reads N norms
    if (n_samples < N)
        fail(SCL_ERROR_DIMENSION_MISMATCH, "n_samples is smaller than the matrix's samples");
    if (N >= std::numeric_limits<int32_t>::max() || k > std::numeric_limits<int32_t>::max())
        fail(SCL_ERROR_INVALID_ARGUMENT, "too many samples or neighbours for the device");
    const int64_t n_pad = knn_rows_padded(N), f_pad = knn_features_padded(F);
    if (f_pad > 0 && n_pad > std::numeric_limits<int64_t>::max() / f_pad / 4)
        fail(SCL_ERROR_INVALID_ARGUMENT, "dense sample matrix too large");
    if (N > 0 && k > std::numeric_limits<int64_t>::max() / N / 8)
        fail(SCL_ERROR_INVALID_ARGUMENT, "n_samples * k too large");
    const int dev = current_device();
    hipStream_t s = library_stream(dev);
    const HostPattern P(m, N, s);
    DevBuf X(n_pad * f_pad), norms(N);
    launch_fill(X.get(), n_pad * f_pad, 0.0f, s);
    launch_densify(P.ip.as<int64_t>(), P.ix.as<int64_t>(), P.v.get(), N, n_pad, f_pad, true,
                   P.serial, X.get(), s);
    if (host_norms) {
        if (N > 0) upload_host(norms.get(), host_norms, (size_t)N * 4, s, dev);
    } else {
        launch_sum_squared(P.ip.as<int64_t>(), P.v.get(), N, norms.get(), s);
    }
    DevBuf ld(N * k), lj(N * k, sizeof(int32_t)), od(N * k), oi(N * k, sizeof(int64_t));
    launch_knn(X.get(), norms.get(), N, f_pad, k, ld.get(), lj.as<int32_t>(), od.get(),
               oi.as<int64_t>(), s);
    download_bytes(out_dist, od.get(), (size_t)(N * k) * 4, s);
    download_bytes(out_idx, oi.get(), (size_t)(N * k) * 8, s);
}

}  // namespace

extern "C" {

// neighbors.cpp:20-38
SCL_EXPORT scl_error_t scl_neighbors_compute_norms(scl_sparse_t matrix, scl_real_t* norms_sq) {
    SCL_API_CHECK_NULL(matrix, "Matrix handle is null");
    SCL_API_CHECK_NULL(norms_sq, "Norms output pointer is null");
    SCL_API_TRY
        // the output holds rows() values; compute_norms needs one per
        // primary slot (neighbors.hpp:310)
        const int64_t p = matrix->primary();
        if (p > matrix->rows) fail(SCL_ERROR_DIMENSION_MISMATCH, "Norms size mismatch");
        const int dev = current_device();
        hipStream_t s = library_stream(dev);
        const HostRows P(matrix, p, s);
        DevBuf out(p);
        launch_sum_squared(P.ip.as<int64_t>(), P.v.get(), p, out.get(), s);
        download_bytes(norms_sq, out.get(), (size_t)p * 4, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// neighbors.cpp:44-75
SCL_EXPORT scl_error_t scl_knn(scl_sparse_t matrix, const scl_real_t* norms_sq,
                               scl_size_t n_samples, scl_size_t k, scl_index_t* out_indices,
                               scl_real_t* out_distances) {
    SCL_API_CHECK_NULL(matrix, "Matrix handle is null");
    SCL_API_CHECK_NULL(norms_sq, "Norms input pointer is null");
    SCL_API_CHECK_NULL(out_indices, "Indices output pointer is null");
    SCL_API_CHECK_NULL(out_distances, "Distances output pointer is null");
    SCL_API_CHECK(n_samples > 0 && k > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        knn_impl(matrix, norms_sq, (int64_t)n_samples, (int64_t)k, out_indices, out_distances);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// neighbors.cpp:81-124
SCL_EXPORT scl_error_t scl_knn_simple(scl_sparse_t matrix, scl_size_t n_samples, scl_size_t k,
                                      scl_index_t* out_indices, scl_real_t* out_distances) {
    SCL_API_CHECK_NULL(matrix, "Matrix handle is null");
    SCL_API_CHECK_NULL(out_indices, "Indices output pointer is null");
    SCL_API_CHECK_NULL(out_distances, "Distances output pointer is null");
    SCL_API_CHECK(k > 0, SCL_ERROR_INVALID_ARGUMENT, "k must be positive");
    SCL_API_TRY
        knn_impl(matrix, nullptr, (int64_t)n_samples, (int64_t)k, out_indices, out_distances);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
