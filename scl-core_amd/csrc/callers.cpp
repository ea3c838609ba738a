// callers.cpp -- the other scl-core entry points built on the diffusion SpMM:
// scl_impute_diffusion and scl_impute_magic (scl/binding/c_api/impute.cpp:
// 93-157), scl_pseudotime_diffusion (scl/binding/c_api/pseudotime.cpp:80-118)
// and scl_algebra_spmm (scl/binding/c_api/algebra.cpp:164-195).
//
// Each keeps its dense operands on the GPU for the whole call; host buffers
// are copied in once and out once.  The t-step loops run on the row-split
// SpMM kernels (the stored-order FMA chain of every row).
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_algebra.h"
#include "scl_impute.h"
#include "scl_pseudotime.h"
#include "session.hpp"

#include <cstdint>
#include <limits>
#include <mutex>
#include <utility>
#include <vector>

using namespace scl_amd;

namespace {

constexpr int64_t kPseudotimeIterations = 100;  // pseudotime.hpp:37
constexpr uint64_t kPseudotimeSeed = 42;        // pseudotime.hpp:669
constexpr uint64_t kDegenerateSeedStep = 12345; // pseudotime.hpp:462

void check_dense(int64_t rows, int64_t cols) {
    if (rows > 0 && cols > std::numeric_limits<int64_t>::max() / rows / 4)
        fail(SCL_ERROR_INVALID_ARGUMENT, "dense operand too large");
}

// Operator checks shared by the imputations: square, n_cells rows.
void check_cells_operator(const scl_sparse_matrix* P, int64_t n_cells) {
    if (P->rows != P->cols || P->rows != n_cells)
        fail(SCL_ERROR_DIMENSION_MISMATCH,
             "the transition / affinity matrix must be n_cells x n_cells");
}

// FastRNG (pseudotime.hpp:153-182) for the start of the components.
class FastRng {
public:
    explicit FastRng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s_) {
            z += 0x9e3779b97f4a7c15ULL;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            w = (uint32_t)(z >> 32);
        }
    }
    float uniform() {
        uint32_t t = s_[3];
        const uint32_t x = s_[0];
        s_[3] = s_[2];
        s_[2] = s_[1];
        s_[1] = x;
        t ^= t >> 11;
        t ^= t << 8;
        s_[0] = t ^ x ^ (x << 19);
        return (float)s_[0] * 2.3283064365386963e-10f;
    }

private:
    uint32_t s_[4];
};

}  // namespace

extern "C" {

// impute.cpp:93-119
SCL_EXPORT scl_error_t scl_impute_diffusion(scl_sparse_t matrix,
                                            scl_sparse_t transition_matrix,
                                            scl_index_t n_cells, scl_index_t n_genes,
                                            scl_index_t n_steps, scl_real_t* X_imputed) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(transition_matrix, "Transition matrix is null");
    SCL_API_CHECK_NULL(X_imputed, "Output imputed matrix is null");
    SCL_API_CHECK(n_cells > 0 && n_genes > 0 && n_steps > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        check_dense(n_cells, n_genes);
        check_cells_operator(transition_matrix, n_cells);
        const bool is_csr = matrix->is_csr;
        const int64_t slots = is_csr ? n_cells : n_genes;
        if (matrix->primary() < slots)
            fail(SCL_ERROR_DIMENSION_MISMATCH,
                 "the expression matrix has fewer cells / genes than requested");
        const int dev = current_device();
        const HostPattern X(matrix, slots, library_stream(dev));
        Session S(transition_matrix, n_cells);
        const int64_t total = n_cells * n_genes;
        DevBuf a(total), b(total);
        launch_set_unit(a.get(), total, -1, 0.0f, S.s);
        X.densify(a.get(), n_cells, n_genes, is_csr, S.s);
        float* cur = a.get();
        float* nxt = b.get();
        for (int64_t step = 0; step < n_steps; ++step) {
            S.spmm(cur, nxt, n_genes);
            std::swap(cur, nxt);
        }
        S.download(X_imputed, cur, total);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// impute.cpp:125-157
SCL_EXPORT scl_error_t scl_impute_magic(scl_sparse_t matrix, scl_sparse_t affinity_matrix,
                                        scl_index_t n_cells, scl_index_t n_genes,
                                        scl_index_t t, scl_real_t* X_imputed) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(affinity_matrix, "Affinity matrix is null");
    SCL_API_CHECK_NULL(X_imputed, "Output imputed matrix is null");
    SCL_API_CHECK(n_cells > 0 && n_genes > 0 && t > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        check_dense(n_cells, n_genes);
        check_cells_operator(affinity_matrix, n_cells);
        // only a cells-major (CSR) expression matrix is copied in
        // (impute.hpp:683-697); a CSC one leaves the start at zero
        const bool is_csr = matrix->is_csr;
        if (is_csr && matrix->primary() < n_cells)
            fail(SCL_ERROR_DIMENSION_MISMATCH,
                 "the expression matrix has fewer cells than requested");
        const int dev = current_device();
        std::unique_ptr<HostPattern> X;
        if (is_csr) X = std::make_unique<HostPattern>(matrix, n_cells, library_stream(dev));
        Session S(affinity_matrix, n_cells);
        const int64_t total = n_cells * n_genes;
        DevBuf a(total), b(total), w(S.d->nnz), copy(n_cells, sizeof(int32_t)),
            n_copy(1, sizeof(int));
        launch_set_unit(a.get(), total, -1, 0.0f, S.s);
        if (X) X->densify(a.get(), n_cells, n_genes, true, S.s);
        hip_check(hipMemsetAsync(n_copy.get(), 0, sizeof(int), S.s), "clear count");
        launch_magic_weights(S.d->row_ptr, S.d->val, S.d->n, w.get(), copy.as<int32_t>(),
                             n_copy.as<int>(), S.s);
        int m = 0;
        hip_check(hipMemcpyAsync(&m, n_copy.get(), sizeof(int), hipMemcpyDeviceToHost, S.s),
                  "download count");
        hip_check(hipStreamSynchronize(S.s), "row sums");
        float* cur = a.get();
        float* nxt = b.get();
        for (int64_t step = 0; step < t; ++step) {
            S.spmm_values(w.get(), cur, nxt, n_genes);
            launch_copy_rows(copy.as<int32_t>(), m, cur, nxt, n_genes, S.s);
            std::swap(cur, nxt);
        }
        S.download(X_imputed, cur, total);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// pseudotime.cpp:80-118
SCL_EXPORT scl_error_t scl_pseudotime_diffusion(scl_sparse_t adjacency,
                                                scl_index_t root_cell,
                                                scl_real_t* pseudotime, scl_index_t n,
                                                scl_index_t n_dcs) {
    if (!adjacency || !pseudotime) {
        set_last_error(SCL_ERROR_NULL_POINTER, "Null pointer argument");
        return SCL_ERROR_NULL_POINTER;
    }
    SCL_API_CHECK(root_cell >= 0 && root_cell < n, SCL_ERROR_INVALID_ARGUMENT,
                  "Root cell index out of bounds");
    SCL_API_CHECK(n_dcs >= 0 && n_dcs <= std::numeric_limits<int32_t>::max() / 2,
                  SCL_ERROR_INVALID_ARGUMENT, "n_dcs out of range");
    SCL_API_TRY
        Session S(adjacency, n);
        const int64_t N = S.n, K = n_dcs;
        if (root_cell >= N) fail(SCL_ERROR_INVALID_ARGUMENT, "Root cell index out of bounds");
        check_dense(N, K);
        DevBuf pt(N), part(1024);
        if (K == 0) {
            launch_set_unit(pt.get(), N, -1, 0.0f, S.s);
        } else {
            std::vector<float> start((size_t)(N * K));
            FastRng rng(kPseudotimeSeed);
            for (auto& x : start) x = rng.uniform() - 0.5f;
            const bool chol = K <= kCholQrMaxK;
            DevBuf Q(N * K), Qn(N * K), C(N * K), w(N), gpart(gram_schmidt_part_floats(N, K)),
                sc(8), cw(chol ? cholqr_workspace_doubles(N, K) : 0, sizeof(double)),
                flag(1, sizeof(int));
            // a Cholesky QR (four launches); with more than kCholQrMaxK
            // columns or a degenerate basis, Gram-Schmidt over contiguous
            // columns (C is K x N) across the GPU, which keeps the reference's
            // FastRNG directions for degenerate columns
            auto orthonormalise = [&](float* M, bool use_chol) {
                if (use_chol) {
                    launch_cholqr(M, N, K, cw.as<double>(), flag.as<int>(), S.s);
                    return;
                }
                launch_transpose_rect(M, C.get(), N, K, S.s);
                launch_gram_schmidt(C.get(), w.get(), N, K, gpart.get(), sc.get(), false,
                                    kDegenerateSeedStep, S.s);
                launch_transpose_rect(C.get(), M, K, N, S.s);
            };
            // the Cholesky path iterates in the operator's locality numbering
            // (its renumbered copy: the products gather neighbouring rows);
            // the orthonormalisation does not depend on the row order
            auto components = [&](bool use_chol) {
                const bool ren = use_chol && S.renumbered();
                hip_check(hipMemsetAsync(flag.get(), 0, sizeof(int), S.s), "clear flag");
                S.upload(ren ? Qn.get() : Q.get(), start.data(), N * K);
                if (ren) S.renumber_rows(Qn.get(), Q.get(), K, false);
                orthonormalise(Q.get(), use_chol);
                float* cur = Q.get();
                float* nxt = Qn.get();
                for (int64_t it = 0; it < kPseudotimeIterations; ++it) {
                    if (ren) S.spmm_renumbered(cur, nxt, K);
                    else S.spmm(cur, nxt, K);
                    orthonormalise(nxt, use_chol);
                    std::swap(cur, nxt);
                }
                if (ren) {
                    S.renumber_rows(cur, nxt, K, true);
                    std::swap(cur, nxt);
                }
                return cur;
            };
            float* cur = components(chol);
            if (chol) {
                int degenerate = 0;
                S.download(reinterpret_cast<float*>(&degenerate), flag.get(), 1);
                if (degenerate) cur = components(false);
            }
            launch_pt_distance(cur, N, K, root_cell, pt.get(), part.get(), S.s);
        }
        S.download(pseudotime, pt.get(), N);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

// algebra.cpp:164-195
SCL_EXPORT scl_error_t scl_algebra_spmm(scl_sparse_t A, const scl_real_t* X,
                                        scl_index_t n_cols, scl_real_t* Y, scl_real_t alpha,
                                        scl_real_t beta) {
    SCL_API_CHECK_NULL(A, "Matrix A is null");
    SCL_API_CHECK_NULL(X, "Input matrix X is null");
    SCL_API_CHECK_NULL(Y, "Output matrix Y is null");
    SCL_API_CHECK(n_cols > 0, SCL_ERROR_INVALID_ARGUMENT, "Number of columns must be positive");
    SCL_API_TRY
        Session S(A, 0, false);
        const int64_t m = S.d->n, k = S.d->n_sec, F = n_cols;
        check_dense(std::max<int64_t>(m, k), F);
        DevBuf dX(k * F), dY(m * F);
        if (k > 0) S.upload(dX.get(), X, k * F);
        // beta == 0 never reads Y (a NaN there is cleared)
        if (beta != 0.0f && m > 0) S.upload(dY.get(), Y, m * F);
        launch_spmm_axpby(S.v, dX.get(), dY.get(), F, alpha, beta, S.s);
        if (m > 0) S.download(Y, dY.get(), m * F);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
