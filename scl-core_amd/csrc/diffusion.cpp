// diffusion.cpp -- the scl_diffusion_* drop-in entry points and their
// device-pointer extensions (counterpart of scl/binding/c_api/diffusion.cpp).
//
// The reference runs diffuse_matrix (scl/kernel/diffusion.hpp:638-664) as
// t x { spmm_block into a zeroed temporary; memcpy back }.  Here the t steps
// ping-pong between two device buffers: no zero fill (every output element
// is written once per step) and no copy between steps.
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_diffusion.h"
#include "scl_diffusion_gpu.h"

#include <cstring>
#include <limits>

using namespace scl_amd;

namespace {

constexpr int64_t kInt32Max = std::numeric_limits<int32_t>::max();

CsrView view_of(const DeviceCSR& d) {
    return CsrView{d.n, d.n_sec, d.nnz, d.row_ptr, d.col, d.val, d.sched, &d.merge,
                   d.grouped, &d.staged, d.tiled};
}

CsrView transposed_view_of(const DeviceCSR& d) {
    return CsrView{d.n_sec, d.n, d.nnz, d.t_row_ptr, d.t_col, d.t_val, d.t_sched,
                   &d.merge};
}

// Checks the operator against the dense operand: the reference never checks
// n_nodes against the operator (diffusion.cpp:94-116) and reads out of
// bounds on a mismatch; here it is SCL_ERROR_DIMENSION_MISMATCH.
void check_operator(const scl_sparse_matrix* P, int64_t n_nodes, int64_t F) {
    if (P->rows != P->cols)
        fail(SCL_ERROR_DIMENSION_MISMATCH, "Transition matrix must be square");
    if (P->rows != n_nodes)
        fail(SCL_ERROR_DIMENSION_MISMATCH,
             "n_nodes does not match the transition matrix dimensions");
    if (F > 0 && n_nodes > std::numeric_limits<int64_t>::max() / F / 4)
        fail(SCL_ERROR_INVALID_ARGUMENT, "n_nodes * n_features overflows");
    if (F > kInt32Max)
        fail(SCL_ERROR_INVALID_ARGUMENT, "n_features too large");
}

// Operators smaller than this are not worth a grouped form in the autotuner
// (the form costs a host pass over the pattern).
constexpr int64_t kGroupedMinRows = 4096;

// The measured-best forward shape of wide operators (n >= kGroupedMinRows,
// F >= 128): row groups of 4 (vec 4, 16 union entries per chunk,
// non-temporal stores, at most 4 workgroups per CU).  At n = 200,000, nnz
// 8.9 M, F = 2000 with slab-planar iterates it runs 2.73 ms per step (8
// entries per chunk: 2.76 ms; 16 uncapped: 2.82 ms, its extra resident
// rows crowd the L2).
SpmmShape wide_default_shape() {
    SpmmShape g;
    g.vec = 4;
    g.group = 64;
    g.unroll = 16;
    g.rows_per_group = 4;
    g.wg_per_cu = 4;
    g.nontemporal = true;
    return g;
}

// Whether the iterates between the first and the last of t steps are kept
// slab-planar: when F floats per row leave rows off the 128-byte lines and
// the shape reads planar operands, a full slab's row is then line-aligned
// (one gather touches 8 lines of 128 bytes instead of 9).  The first step
// reads X row-major and the last writes Y row-major; results are the same
// bit for bit.
bool planar_iterates(const CsrView& P, const SpmmShape& shape, int64_t F, int64_t t) {
    return t >= 2 && (F * 4) % 128 != 0 && F > 64 * shape.vec &&
           spmm_layout_supported(shape, P);
}

int step_layout(bool planar, int64_t step, int64_t t) {
    if (!planar) return 0;
    return (step > 1 ? 1 : 0) | (step < t ? 2 : 0);
}

// t steps between two device buffers, starting from buf[start]; returns the
// index of the buffer holding the result.
int pingpong(const CsrView& P, const SpmmShape& shape, float* buf[2], int start,
             int64_t F, int64_t t, hipStream_t s) {
    const bool planar = planar_iterates(P, shape, F, t);
    int cur = start;
    for (int64_t step = 1; step <= t; ++step) {
        launch_spmm_f32(P, buf[cur], buf[cur ^ 1], F, shape, s, step_layout(planar, step, t));
        cur ^= 1;
    }
    return cur;
}

// Host-pointer diffusion: upload X, t steps on the device, download.
void diffuse_host(scl_sparse_matrix* P, float* X, int64_t n, int64_t F,
                  int64_t t) {
    const int dev = current_device();
    std::lock_guard<std::mutex> call_lk(device_mutex(dev));
    std::lock_guard<std::mutex> lk(P->mu);
    // the caller may have edited a wrapped handle's arrays since last time
    if (P->wrapped) P->generation++;
    DeviceCSR& d = device_csr(P, dev);
    hipStream_t s = library_stream(dev);
    const size_t bytes = (size_t)n * (size_t)F * sizeof(float);
    float* buf[2] = {static_cast<float*>(device_workspace(dev, 0, bytes)),
                     static_cast<float*>(device_workspace(dev, 1, bytes))};
    upload_host(buf[0], X, bytes, s, dev);
    const SpmmShape shape = choose_shape(P, d, false, F, buf[0], buf[1]);
    const int fin = pingpong(view_of(d), shape, buf, 0, F, t, s);
    download_host(X, buf[fin], bytes, s, dev);
    hip_check(hipStreamSynchronize(s), "diffusion");
}

void transition_impl(scl_sparse_matrix* A, int symmetric, double* degrees_out,
                     double* residual_out) {
    if (symmetric && A->rows != A->cols)
        fail(SCL_ERROR_DIMENSION_MISMATCH,
             "Symmetric normalisation needs a square matrix");
    // the residual pairs the degree of row i with transposed row i: square only
    if (residual_out && A->rows != A->cols)
        fail(SCL_ERROR_DIMENSION_MISMATCH,
             "The stationarity residual needs a square matrix");
    const int dev = current_device();
    std::lock_guard<std::mutex> call_lk(device_mutex(dev));
    std::lock_guard<std::mutex> lk(A->mu);
    if (A->wrapped) A->generation++;
    DeviceCSR& d = device_csr(A, dev);
    hipStream_t s = library_stream(dev);
    const int32_t n = d.n;
    // degrees (n doubles) + reduction scratch (1025 doubles)
    double* deg = static_cast<double*>(
        device_workspace(dev, 2, sizeof(double) * ((size_t)n + 2048)));
    double* scratch = deg + n;
    launch_row_sums_f64(d.row_ptr, d.val, n, deg, s);
    launch_normalize_rows(d.row_ptr, d.col, d.val, n, deg, symmetric, s);
    refresh_grouped_values(d, s);
    if (A->nnz > 0) download_host(A->data, d.val, (size_t)A->nnz * 4, s, dev);
    if (degrees_out)
        hip_check(hipMemcpyAsync(degrees_out, deg, sizeof(double) * (size_t)n,
                                 hipMemcpyDeviceToHost, s),
                  "download degrees");
    hip_check(hipStreamSynchronize(s), "normalisation");
    // the device copy now holds the new values: advance both generations
    A->generation++;
    d.generation = A->generation;
    if (d.pins > 0 && d.t_row_ptr) {
        launch_gather_f32(d.t_src, d.val, d.t_val, d.nnz, s);
        hip_check(hipStreamSynchronize(s), "refresh transposed values");
        d.t_generation = d.generation;
    } else {
        d.release_transpose();
    }
    if (residual_out) {
        ensure_transpose(A, d, s);
        double* out = scratch + 1100;
        launch_stationarity_residual(d.t_row_ptr, d.t_col, d.t_val, d.n_sec, deg,
                                     scratch, out, s);
        hip_check(hipMemcpyAsync(residual_out, out, sizeof(double),
                                 hipMemcpyDeviceToHost, s),
                  "download residual");
        hip_check(hipStreamSynchronize(s), "residual");
    }
}

}  // namespace

// Launch shape for F features: the one tuned or set for this operator, when
// it fits the buffers; else, for wide forward operators whose rows are
// sorted, the row-grouped shape (its grouped form is built on the first
// such call when `build` is set, and the shape is kept for F); else the
// default.  Caller holds the handle's mutex.
SpmmShape scl_amd::choose_shape(scl_sparse_matrix* h, DeviceCSR& d, bool transposed,
                                int64_t F, const void* X, const void* Y, bool build) {
    auto& m = transposed ? d.t_shapes : d.shapes;
    auto it = m.find(F);
    if (it != m.end() && spmm_shape_valid(it->second, F, X, Y)) return it->second;
    if (!transposed && it == m.end() && d.n >= kGroupedMinRows && F >= 128) {
        const SpmmShape g = wide_default_shape();
        const bool ready = d.grouped[0].R == 4 || (build && d.pins == 0 && ensure_grouped(h, d, 4));
        if (ready && spmm_shape_valid(g, F, nullptr, nullptr)) {
            if (build) m[F] = g;
            if (spmm_shape_valid(g, F, X, Y)) return g;
        } else if (build) {
            m[F] = spmm_default_shape(F, nullptr, nullptr);  // not groupable: do not retry
        }
    }
    return spmm_default_shape(F, X, Y);
}


extern "C" {

// diffusion.cpp:36-62
SCL_EXPORT scl_error_t scl_diffusion_compute_transition_matrix(
    scl_sparse_t adjacency, int symmetric) {
    SCL_API_CHECK_NULL(adjacency, "Adjacency matrix is null");
    SCL_API_TRY
        transition_impl(adjacency, symmetric, nullptr, nullptr);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_compute_transition_matrix_ex(
    scl_sparse_t adjacency, int symmetric, double* degrees_out,
    double* residual_out) {
    SCL_API_CHECK_NULL(adjacency, "Adjacency matrix is null");
    SCL_API_TRY
        transition_impl(adjacency, symmetric, degrees_out, residual_out);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// diffusion.cpp:68-88
SCL_EXPORT scl_error_t scl_diffusion_diffuse_vector(scl_sparse_t transition,
                                                    scl_real_t* x,
                                                    scl_size_t n_nodes,
                                                    scl_index_t n_steps) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(x, "Vector x is null");
    SCL_API_CHECK(n_nodes > 0 && n_steps > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        check_operator(transition, (int64_t)n_nodes, 1);
        diffuse_host(transition, x, (int64_t)n_nodes, 1, n_steps);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// diffusion.cpp:94-116
SCL_EXPORT scl_error_t scl_diffusion_diffuse_matrix(scl_sparse_t transition,
                                                    scl_real_t* X,
                                                    scl_index_t n_nodes,
                                                    scl_index_t n_features,
                                                    scl_index_t n_steps) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(X, "Matrix X is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_TRY
        check_operator(transition, n_nodes, n_features);
        diffuse_host(transition, X, n_nodes, n_features, n_steps);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// diffusion.cpp:318-341
SCL_EXPORT scl_error_t scl_diffusion_magic_impute(scl_sparse_t transition,
                                                  scl_real_t* X,
                                                  scl_index_t n_nodes,
                                                  scl_index_t n_features,
                                                  scl_index_t t) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(X, "Matrix X is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        check_operator(transition, n_nodes, n_features);
        const int64_t steps = (t == 0) ? 3 : t;  // diffusion.cpp:333
        if (steps > 0) diffuse_host(transition, X, n_nodes, n_features, steps);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_device_count(int* out) {
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *out = c;
    SCL_API_RETURN_OK;
}

SCL_EXPORT scl_error_t scl_sparse_locality_order(scl_sparse_t matrix,
                                                 scl_index_t* order_out) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(order_out, "Output pointer is null");
    SCL_API_TRY
        std::lock_guard<std::mutex> lk(matrix->mu);
        const auto order = locality_schedule(matrix->primary(), matrix->indptr,
                                             matrix->indices);
        for (size_t i = 0; i < order.size(); ++i) order_out[i] = order[i];
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_sparse_invalidate(scl_sparse_t matrix) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_TRY
        std::lock_guard<std::mutex> lk(matrix->mu);
        matrix->generation++;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_sparse_prepare(scl_sparse_t matrix) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_TRY
        const int dev = current_device();
        std::lock_guard<std::mutex> lk(matrix->mu);
        device_csr(matrix, dev);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_real_t* work_dev, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t n_steps, void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(X_dev, "X is null");
    SCL_API_CHECK_NULL(Y_dev, "Y is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0 && n_steps > 0,
                  SCL_ERROR_INVALID_ARGUMENT, "Dimensions must be positive");
    SCL_API_CHECK(n_steps == 1 || work_dev != nullptr, SCL_ERROR_NULL_POINTER,
                  "work buffer is required for n_steps >= 2");
    SCL_API_CHECK((const void*)X_dev != (const void*)Y_dev &&
                      (const void*)X_dev != (const void*)work_dev &&
                      (const void*)Y_dev != (const void*)work_dev,
                  SCL_ERROR_INVALID_ARGUMENT, "X, Y and work must be distinct");
    SCL_API_TRY
        check_operator(transition, n_nodes, n_features);
        const int dev = current_device();
        CsrView v;
        SpmmShape sh;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR& d = device_csr(transition, dev);
            v = view_of(d);
            sh = choose_shape(transition, d, false, n_features, X_dev, Y_dev);
            if (work_dev && !spmm_shape_valid(sh, n_features, work_dev, Y_dev))
                sh = spmm_default_shape(n_features, work_dev, Y_dev);
            if (!spmm_shape_valid(sh, n_features, X_dev, Y_dev))
                sh = spmm_default_shape(n_features, X_dev, Y_dev);
            if (work_dev && !spmm_shape_valid(sh, n_features, work_dev, X_dev))
                sh = spmm_default_shape(n_features, work_dev, X_dev);
        }
        hipStream_t s = static_cast<hipStream_t>(stream);
        // step s (1-based) writes Y when t - s is even, else work, so the
        // last step lands in Y and consecutive steps alternate buffers
        const bool planar = planar_iterates(v, sh, n_features, n_steps);
        const float* src = X_dev;
        for (int64_t step = 1; step <= n_steps; ++step) {
            float* dst = ((n_steps - step) % 2 == 0) ? Y_dev : work_dev;
            launch_spmm_f32(v, src, dst, n_features, sh, s, step_layout(planar, step, n_steps));
            src = dst;
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_spmm_device(scl_sparse_t transition,
                                           const scl_real_t* X_dev,
                                           scl_real_t* Y_dev,
                                           scl_index_t n_nodes,
                                           scl_index_t n_features,
                                           int transpose, void* stream) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(X_dev, "X is null");
    SCL_API_CHECK_NULL(Y_dev, "Y is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_CHECK((const void*)X_dev != (const void*)Y_dev,
                  SCL_ERROR_INVALID_ARGUMENT, "X and Y must be distinct");
    SCL_API_TRY
        check_operator(transition, n_nodes, n_features);
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        CsrView v;
        SpmmShape sh;
        {
            std::lock_guard<std::mutex> lk(transition->mu);
            DeviceCSR& d = device_csr(transition, dev);
            if (transpose) {
                ensure_transpose(transition, d, s);
                v = transposed_view_of(d);
            } else {
                v = view_of(d);
            }
            sh = choose_shape(transition, d, transpose != 0, n_features, X_dev, Y_dev);
        }
        launch_spmm_f32(v, X_dev, Y_dev, n_features, sh, s);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_spmm_set_shape(scl_sparse_t transition,
                                              scl_index_t n_features,
                                              int transpose, int vec, int group,
                                              int unroll, int flags) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(n_features > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "n_features must be positive");
    SCL_API_CHECK(transpose >= 0 && transpose <= 2, SCL_ERROR_INVALID_ARGUMENT,
                  "transpose must be 0 (P), 1 (P^T) or 2 (fused adjoint)");
    SpmmShape sh;
    sh.vec = vec;
    sh.group = group;
    sh.unroll = unroll;
    sh.nontemporal = (flags & kShapeNontemporal) != 0;
    sh.schedule = (flags & kShapeNoSchedule) == 0;
    sh.xcd_rows = (flags & kShapeXcdRows) != 0;
    sh.xcd_block = (flags >> kShapeXcdBlockShift) & kShapeXcdBlockMask;
    sh.scalar_rows = (flags & kShapeScalarRows) != 0;
    sh.wg_per_cu = (flags >> kShapeOccupancyShift) & kShapeOccupancyMask;
    sh.merge = (flags >> kShapeMergeShift) & kShapeMergeMask;
    const int gcode = (flags >> kShapeGroupShift) & kShapeGroupMask;
    const bool tiled = (flags & kShapeTiled) != 0;
    SCL_API_CHECK(gcode == 0 || transpose == 0, SCL_ERROR_INVALID_ARGUMENT,
                  "grouped form: code 1 (4 rows), 2 (8 rows) or 3 (LDS-staged groups), "
                  "forward direction only");
    SCL_API_CHECK(!tiled || (transpose == 0 && gcode == 0), SCL_ERROR_INVALID_ARGUMENT,
                  "tiled form: forward direction only, not with a grouped form");
    sh.rows_per_group = tiled ? kTiledRows : gcode == 3 ? kStagedRows : gcode ? 4 * gcode : 0;
    SCL_API_CHECK(sh.merge <= kMergeMax && (sh.merge == 0 || transpose != 2),
                  SCL_ERROR_INVALID_ARGUMENT,
                  "merge-path code must be 0..8, and 0 for the fused adjoint");
    if (transpose == 2) {
        sh.unroll = 8;
        sh.nontemporal = false;
        sh.xcd_rows = false;
        sh.scalar_rows = false;
        SCL_API_CHECK(adjoint_shape_valid(sh, n_features, nullptr, nullptr, nullptr),
                      SCL_ERROR_INVALID_ARGUMENT,
                      "invalid adjoint shape: vec in {1,2,4} dividing n_features, "
                      "group a power of two in 1..64");
    } else {
        SCL_API_CHECK(spmm_shape_valid(sh, n_features, nullptr, nullptr),
                      SCL_ERROR_INVALID_ARGUMENT,
                      "invalid shape: vec in {1,2,4} dividing n_features, group a "
                      "power of two in 1..64, unroll 4, 8 or 16");
    }
    SCL_API_TRY
        const int dev = current_device();
        std::lock_guard<std::mutex> lk(transition->mu);
        DeviceCSR& d = device_csr(transition, dev);
        if (sh.merge > 0)
            ensure_merge_scratch(d, merge_waves(transpose ? d.n_sec : d.n, d.nnz, sh.merge),
                                 n_features);
        if (sh.rows_per_group == kTiledRows) {
            if (!ensure_tiled(transition, d, sh.group, sh.wg_per_cu))
                fail(SCL_ERROR_INVALID_ARGUMENT,
                     "tiled form needs strictly increasing column indices in every row "
                     "and rows that fit one tile");
        } else if (sh.rows_per_group == kStagedRows) {
            if (!ensure_staged(transition, d))
                fail(SCL_ERROR_INVALID_ARGUMENT,
                     "staged form needs strictly increasing column indices in every row");
        } else if (sh.rows_per_group > 0 && !ensure_grouped(transition, d, sh.rows_per_group)) {
            fail(SCL_ERROR_INVALID_ARGUMENT,
                 "grouped form needs strictly increasing column indices in every row "
                 "and fewer than 2^24 columns");
        }
        (transpose == 2 ? d.adj_shapes : transpose ? d.t_shapes : d.shapes)[n_features] = sh;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_spmm_get_shape(scl_sparse_t transition,
                                              scl_index_t n_features,
                                              int transpose, int* vec,
                                              int* group, int* unroll,
                                              int* flags) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK(vec && group && unroll && flags, SCL_ERROR_NULL_POINTER,
                  "Output pointer is null");
    SCL_API_CHECK(n_features > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "n_features must be positive");
    SCL_API_CHECK(transpose >= 0 && transpose <= 2, SCL_ERROR_INVALID_ARGUMENT,
                  "transpose must be 0 (P), 1 (P^T) or 2 (fused adjoint)");
    SCL_API_TRY
        const int dev = current_device();
        std::lock_guard<std::mutex> lk(transition->mu);
        DeviceCSR& d = device_csr(transition, dev);
        // the choice for 16-byte aligned buffers
        SpmmShape sh;
        if (transpose == 2) {
            const auto it = d.adj_shapes.find(n_features);
            if (it != d.adj_shapes.end()) {
                sh = it->second;
            } else {  // the launcher's default: widest vec, one row per group
                sh.vec = n_features % 4 == 0 ? 4 : n_features % 2 == 0 ? 2 : 1;
                sh.group = 64;
            }
            sh.unroll = 8;
            sh.nontemporal = false;
            sh.xcd_rows = false;
            sh.scalar_rows = false;
        } else {
            sh = choose_shape(transition, d, transpose != 0, n_features, nullptr, nullptr,
                              false);
        }
        *vec = sh.vec;
        *group = sh.group;
        *unroll = sh.unroll;
        *flags = (sh.nontemporal ? kShapeNontemporal : 0) |
                 (sh.schedule ? 0 : kShapeNoSchedule) |
                 (sh.xcd_rows ? kShapeXcdRows : 0) |
                 (sh.xcd_block << kShapeXcdBlockShift) |
                 (sh.scalar_rows ? kShapeScalarRows : 0) |
                 (sh.wg_per_cu << kShapeOccupancyShift) |
                 (sh.merge << kShapeMergeShift) |
                 (sh.rows_per_group == kTiledRows ? kShapeTiled
                  : ((sh.rows_per_group == kStagedRows ? 3 : sh.rows_per_group / 4)
                     << kShapeGroupShift));
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_gpu_spmm_autotune(scl_sparse_t transition,
                                             const scl_real_t* X_dev,
                                             scl_real_t* Y_dev,
                                             scl_index_t n_nodes,
                                             scl_index_t n_features,
                                             int transpose, void* stream,
                                             float* best_ms) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(X_dev, "X is null");
    SCL_API_CHECK_NULL(Y_dev, "Y is null");
    SCL_API_CHECK(n_nodes > 0 && n_features > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_CHECK((const void*)X_dev != (const void*)Y_dev,
                  SCL_ERROR_INVALID_ARGUMENT, "X and Y must be distinct");
    SCL_API_CHECK(transpose >= 0 && transpose <= 2, SCL_ERROR_INVALID_ARGUMENT,
                  "transpose must be 0 (P), 1 (P^T) or 2 (fused adjoint)");
    SCL_API_TRY
        check_operator(transition, n_nodes, n_features);
        const int dev = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        std::lock_guard<std::mutex> lk(transition->mu);
        DeviceCSR& d = device_csr(transition, dev);
        if (transpose) ensure_transpose(transition, d, s);
        const CsrView v = transpose ? transposed_view_of(d) : view_of(d);
        // the fused adjoint is timed with its edge term on: X stands in for
        // both G_s and X_{s-1}, and the per-nonzero output goes to scratch
        float* dpt = nullptr;
        if (transpose == 2 && d.nnz > 0)
            dpt = static_cast<float*>(device_workspace(dev, 4, (size_t)d.nnz * sizeof(float)));
        auto cands = transpose == 2
                         ? adjoint_candidate_shapes(n_features, X_dev, Y_dev, nullptr,
                                                    v.sched != nullptr)
                         : spmm_candidate_shapes(n_features, X_dev, Y_dev, v.sched != nullptr);
        if (transpose == 0 && kGroupedMinRows <= d.n) {
            // row groups share gathers of common neighbours: try 4 and 8
            // rows per wave when the pattern allows the grouped form
            // LDS-staged groups of kStagedRows rows (8 waves, rounds of 64
            // union columns: the fastest staged layout)
            if (ensure_staged(transition, d))
                for (int nt = 0; nt < 2; ++nt) {
                    SpmmShape c;
                    c.vec = 4;
                    c.group = 64;
                    c.unroll = 4;
                    c.rows_per_group = kStagedRows;
                    c.nontemporal = nt != 0;
                    if (spmm_shape_valid(c, n_features, X_dev, Y_dev)) cands.push_back(c);
                }
            // LDS tiles: 4 or 8 lanes per row, one workgroup per CU (two
            // measured slower at every width tried)
            for (int lpr : {4, 8})
                for (int budget : {1}) {
                    if (!ensure_tiled(transition, d, lpr, budget)) continue;
                    for (int chunk : {16})
                        for (int nt = 0; nt < 2; ++nt) {
                            SpmmShape c;
                            c.vec = 4;
                            c.group = lpr;
                            c.unroll = chunk;
                            c.wg_per_cu = budget;
                            c.rows_per_group = kTiledRows;
                            c.nontemporal = nt != 0;
                            if (spmm_shape_valid(c, n_features, X_dev, Y_dev) &&
                                (uint64_t)d.n_sec * (uint64_t)n_features <= 0xFFFFFFFFull)
                                cands.push_back(c);
                        }
                }
            for (int R : {4, 8}) {
                if (!ensure_grouped(transition, d, R)) break;
                for (int vec : {4, 2})
                    for (int u : R == 8 ? std::vector<int>{4, 8, 16} : std::vector<int>{8, 16})
                        for (int occ : {0, 4})
                            for (int nt = 0; nt < 2; ++nt) {
                                SpmmShape c;
                                c.vec = vec;
                                c.group = 64;
                                c.unroll = u;
                                c.rows_per_group = R;
                                c.wg_per_cu = occ;
                                c.nontemporal = nt != 0;
                                if (spmm_shape_valid(c, n_features, X_dev, Y_dev))
                                    cands.push_back(c);
                            }
            }
        }
        SCL_API_CHECK(!cands.empty(), SCL_ERROR_INVALID_ARGUMENT,
                      "no launch shape fits these buffers");
        // forward shapes are timed as the inner steps of a diffusion run
        // them: slab-planar in and out where the diffusion keeps its
        // iterates so (Y is scratch here; every address stays in the buffers)
        auto run = [&](const SpmmShape& c) {
            if (transpose == 2)
                launch_adjoint_edge(v, X_dev, dpt ? X_dev : nullptr, Y_dev, dpt, false,
                                    n_features, s, &c);
            else
                launch_spmm_f32(v, X_dev, Y_dev, n_features, c, s,
                                transpose == 0 && planar_iterates(v, c, n_features, 3) ? 3 : 0);
        };
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        float best = -1.0f;
        SpmmShape best_shape = cands.front();
        try {
            for (const SpmmShape& c : cands) {
                run(c);  // warm
                hip_check(hipEventRecord(e0, s), "event");
                for (int r = 0; r < 2; ++r) run(c);
                hip_check(hipEventRecord(e1, s), "event");
                hip_check(hipEventSynchronize(e1), "event");
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
                if (best < 0.0f || ms < best) {
                    best = ms;
                    best_shape = c;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (transpose == 2 ? d.adj_shapes : transpose ? d.t_shapes : d.shapes)[n_features] =
            best_shape;
        if (best_ms) *best_ms = best / 2.0f;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
