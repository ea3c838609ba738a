// walks.cpp -- the vector-side scl_diffusion_* entry points (counterparts of
// scl/binding/c_api/diffusion.cpp:122-401 over scl/kernel/diffusion.hpp:
// 670-1194): diffusion pseudotime (one or many roots), random walk with
// restart, personalized PageRank, diffusion-map embedding, heat kernel
// signature, diffusion distance and the lazy random walk.
//
// Each keeps its vectors on the GPU and runs its SpMV / SpMM steps on the
// row-split kernels (every row the stored-order FMA chain); host buffers are
// copied in once and out once.  Defaults, argument checks and observable
// quirks follow the reference's ABI layer:
//   - 0 selects the default for alpha (0.85), tol (1e-6), max_iter (100),
//     n_iter (50), laziness (0.5) and the heat kernel's steps (10);
//   - out-of-range seeds and roots are skipped (a DPT root out of range is an
//     error);
//   - random walk with restart swaps its output pointer each iteration, so
//     without convergence an odd max_iter leaves the next-to-last iterate in
//     the caller's buffer;
//   - diffusion_distance compares rows of P^t for n <= 1000 and the vectors
//     P^t e_i (columns of P^t) above.
// Checks beyond the reference: the operator must be square and n_nodes must
// cover its rows (SCL_ERROR_DIMENSION_MISMATCH instead of a read past the
// end).
#include "internal.hpp"
#include "kernels.hpp"
#include "scl_diffusion.h"
#include "session.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

using namespace scl_amd;

namespace {

constexpr float kDefaultAlpha = 0.85f;
constexpr float kDefaultTol = 1e-6f;
constexpr int64_t kDefaultMaxIter = 100;
constexpr int64_t kDefaultNIter = 50;
constexpr int64_t kDefaultHksSteps = 10;
constexpr float kDefaultLaziness = 0.5f;
// rows from which the diffusion-map basis is orthogonalised across the GPU
constexpr int64_t kWideGramSchmidtRows = 16384;
// steps of a converging walk between two host reads of its convergence flag;
// the steps after convergence leave the result unchanged on the device
constexpr int64_t kStepsPerCheck = 8;

// compute_dpt (diffusion.hpp:670-768) on device buffers; pt is device.
// With a renumbered operator the walk runs in its locality numbering (the
// maxima it reduces do not depend on the order) and pt is put back at the end.
void dpt_device(Session& S, int64_t root, float* pt, int64_t max_iter, float tol) {
    const int64_t n = S.n;
    const bool ren = S.walk_renumbered();
    DevBuf p(n), pn(n), hit(n), vis(n), part(1024), scal(3);
    float* done = scal.get() + 2;
    launch_set_unit(ren ? pn.get() : p.get(), n, root, 1.0f, S.s);
    if (ren) {
        S.renumber_rows(pn.get(), p.get(), 1, false);
        hip_check(hipMemcpyAsync(vis.get(), p.get(), (size_t)n * 4, hipMemcpyDeviceToDevice,
                                 S.s),
                  "copy");
    } else {
        launch_set_unit(vis.get(), n, root, 1.0f, S.s);
    }
    launch_set_unit(hit.get(), n, -1, 0.0f, S.s);
    launch_fill(done, 1, 0.0f, S.s);
    float* cur = p.get();
    float* nxt = pn.get();
    for (int64_t it = 1; it <= max_iter;) {
        // the step whose max(1 - visited) falls below tol sets the flag; the
        // later steps of the batch leave visited and hit as they are
        for (const int64_t end = std::min(max_iter, it + kStepsPerCheck - 1); it <= end; ++it) {
            if (ren) S.spmv_renumbered(cur, nxt);
            else S.spmv(cur, nxt);
            launch_dpt_update(nxt, vis.get(), hit.get(), n, (float)it, part.get(), scal.get(),
                              tol, done, S.s);
            std::swap(cur, nxt);
        }
        if (S.scalar(done) != 0.0f) break;
    }
    launch_dpt_finish(vis.get(), hit.get(), ren ? nxt : pt, n, part.get(), scal.get() + 1,
                      S.s);
    if (ren) S.renumber_rows(nxt, pt, 1, true);
}

}  // namespace

extern "C" {

SCL_EXPORT scl_error_t scl_diffusion_lazy_random_walk(scl_sparse_t transition, scl_real_t* x,
                                                      scl_size_t n_nodes,
                                                      scl_index_t n_steps,
                                                      scl_real_t laziness) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(x, "Vector x is null");
    SCL_API_CHECK(n_nodes > 0 && n_steps > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_CHECK(laziness >= 0 && laziness <= 1, SCL_ERROR_INVALID_ARGUMENT,
                  "Laziness must be in [0, 1]");
    SCL_API_TRY
        Session S(transition, (int64_t)n_nodes);
        const int64_t n = S.n;
        if (n > 0) {
            const float lazy = laziness == 0 ? kDefaultLaziness : laziness;
            const float move = 1.0f - lazy;
            // in the locality numbering when the operator has a renumbered
            // copy (the blend is element-wise)
            const bool ren = S.walk_renumbered();
            DevBuf xd(n), t(n);
            S.upload(ren ? t.get() : xd.get(), x, n);
            if (ren) S.renumber_rows(t.get(), xd.get(), 1, false);
            for (int64_t st = 0; st < n_steps; ++st) {
                if (ren) S.spmv_renumbered(xd.get(), t.get());
                else S.spmv(xd.get(), t.get());
                launch_blend(xd.get(), t.get(), n, lazy, move, S.s);
            }
            if (ren) S.renumber_rows(xd.get(), t.get(), 1, true);
            S.download(x, ren ? t.get() : xd.get(), n);
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_diffusion_heat_kernel_signature(scl_sparse_t transition,
                                                           scl_real_t* signature,
                                                           scl_size_t n_nodes, scl_real_t t,
                                                           scl_index_t n_steps) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(signature, "Output signature array is null");
    SCL_API_CHECK(n_nodes > 0, SCL_ERROR_INVALID_ARGUMENT, "Number of nodes must be positive");
    SCL_API_CHECK(t > 0, SCL_ERROR_INVALID_ARGUMENT, "Time parameter t must be positive");
    SCL_API_TRY
        Session S(transition, (int64_t)n_nodes);
        const int64_t n = S.n;
        if (n > 0) {
            const int64_t steps = n_steps == 0 ? kDefaultHksSteps : n_steps;
            const float dt = t / (float)steps, keep = 1.0f - dt;
            // in the locality numbering when the operator has a renumbered
            // copy (the start is all ones, the blend element-wise)
            const bool ren = S.walk_renumbered();
            DevBuf sig(n), tmp(n);
            std::vector<float> ones((size_t)n, 1.0f);
            S.upload(sig.get(), ones.data(), n);
            for (int64_t st = 0; st < steps; ++st) {
                if (ren) S.spmv_renumbered(sig.get(), tmp.get());
                else S.spmv(sig.get(), tmp.get());
                launch_blend(sig.get(), tmp.get(), n, keep, dt, S.s);
            }
            if (ren) S.renumber_rows(sig.get(), tmp.get(), 1, true);
            S.download(signature, ren ? tmp.get() : sig.get(), n);
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_diffusion_random_walk_with_restart(
    scl_sparse_t transition, const scl_index_t* seed_nodes, scl_size_t n_seeds,
    scl_real_t* scores, scl_size_t n_nodes, scl_real_t alpha, scl_index_t max_iter,
    scl_real_t tol) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(seed_nodes, "Seed nodes array is null");
    SCL_API_CHECK_NULL(scores, "Output scores array is null");
    SCL_API_CHECK(n_seeds > 0 && n_nodes > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_CHECK(alpha >= 0 && alpha <= 1, SCL_ERROR_INVALID_ARGUMENT,
                  "Alpha must be in [0, 1]");
    SCL_API_TRY
        Session S(transition, (int64_t)n_nodes);
        const int64_t n = S.n;
        if (n > 0) {
            const float a = alpha == 0 ? kDefaultAlpha : alpha;
            const int64_t iters = max_iter == 0 ? kDefaultMaxIter : max_iter;
            const float tl = tol == 0 ? kDefaultTol : tol;
            std::vector<float> restart((size_t)n, 0.0f);
            int64_t cnt = 0;
            for (size_t q = 0; q < n_seeds; ++q) {
                const int64_t idx = seed_nodes[q];
                if (idx >= 0 && idx < n) {
                    restart[(size_t)idx] = 1.0f;
                    ++cnt;
                }
            }
            if (cnt == 0) {
                std::memset(scores, 0, (size_t)n * sizeof(float));
                SCL_API_RETURN_OK;
            }
            const float inv = 1.0f / (float)cnt;
            for (float& r : restart) r *= inv;
            // in the locality numbering when the operator has a renumbered
            // copy: the update is element-wise; its convergence total sums
            // in that order
            const bool ren = S.walk_renumbered();
            DevBuf out(n), other(n), rs(n), part(1024), total(2);
            float* done = total.get() + 1;
            S.upload(ren ? other.get() : rs.get(), restart.data(), n);
            if (ren) S.renumber_rows(other.get(), rs.get(), 1, false);
            hip_check(hipMemcpyAsync(out.get(), rs.get(), (size_t)n * 4,
                                     hipMemcpyDeviceToDevice, S.s),
                      "copy");
            launch_fill(done, 1, 0.0f, S.s);
            const float beta = 1.0f - a;
            // out is the caller's buffer in the reference; cur/nxt swap as its
            // pointer does
            float* cur = out.get();
            float* nxt = other.get();
            for (int64_t it = 0; it < iters;) {
                for (const int64_t end = std::min(iters, it + kStepsPerCheck); it < end; ++it) {
                    if (ren) S.spmv_renumbered(cur, nxt);
                    else S.spmv(cur, nxt);
                    launch_rwr_update(nxt, cur, rs.get(), n, a, beta, part.get(), total.get(),
                                      tl, done, S.s);
                    std::swap(cur, nxt);
                }
                if (S.scalar(done) != 0.0f) {
                    // the reference copies the converged vector over the
                    // other buffer and stops: cur holds it (the steps after
                    // convergence copied it from buffer to buffer)
                    hip_check(hipMemcpyAsync(nxt, cur, (size_t)n * 4, hipMemcpyDeviceToDevice,
                                             S.s),
                              "copy");
                    break;
                }
            }
            if (ren) S.renumber_rows(out.get(), rs.get(), 1, true);
            S.download(scores, ren ? rs.get() : out.get(), n);
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_diffusion_personalized_pagerank(scl_sparse_t transition,
                                                           scl_index_t seed_node,
                                                           scl_real_t* scores,
                                                           scl_size_t n_nodes,
                                                           scl_real_t alpha,
                                                           scl_index_t max_iter,
                                                           scl_real_t tol) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(scores, "Output scores array is null");
    SCL_API_CHECK(n_nodes > 0, SCL_ERROR_INVALID_ARGUMENT, "Number of nodes must be positive");
    SCL_API_CHECK(alpha >= 0 && alpha <= 1, SCL_ERROR_INVALID_ARGUMENT,
                  "Alpha must be in [0, 1]");
    // personalized_pagerank (diffusion.hpp:1122-1135): one seed
    const scl_index_t seeds[1] = {seed_node};
    return scl_diffusion_random_walk_with_restart(transition, seeds, 1, scores, n_nodes, alpha,
                                                  max_iter, tol);
}

SCL_EXPORT scl_error_t scl_diffusion_compute_dpt(scl_sparse_t transition,
                                                 scl_index_t root_cell,
                                                 scl_real_t* pseudotime, scl_size_t n_nodes,
                                                 scl_index_t max_iter, scl_real_t tol) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(pseudotime, "Output pseudotime array is null");
    SCL_API_CHECK(n_nodes > 0, SCL_ERROR_INVALID_ARGUMENT, "Number of nodes must be positive");
    SCL_API_TRY
        Session S(transition, (int64_t)n_nodes);
        const int64_t n = S.n;
        if (root_cell < 0 || root_cell >= n)
            fail(SCL_ERROR_INVALID_ARGUMENT, "DPT: root_cell out of range");
        const int64_t iters = max_iter == 0 ? kDefaultMaxIter : max_iter;
        const float tl = tol == 0 ? kDefaultTol : tol;
        DevBuf pt(n);
        dpt_device(S, root_cell, pt.get(), iters, tl);
        S.download(pseudotime, pt.get(), n);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_diffusion_compute_dpt_multi_root(
    scl_sparse_t transition, const scl_index_t* root_cells, scl_size_t n_roots,
    scl_real_t* pseudotime, scl_size_t n_nodes, scl_index_t max_iter) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(root_cells, "Root cells array is null");
    SCL_API_CHECK_NULL(pseudotime, "Output pseudotime array is null");
    SCL_API_CHECK(n_roots > 0 && n_nodes > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        Session S(transition, (int64_t)n_nodes);
        const int64_t n = S.n;
        const int64_t iters = max_iter == 0 ? kDefaultMaxIter : max_iter;
        if (n > 0) {
            DevBuf sum(n), one(n);
            launch_set_unit(sum.get(), n, -1, 0.0f, S.s);
            int64_t valid = 0;
            for (size_t r = 0; r < n_roots; ++r) {
                const int64_t root = root_cells[r];
                if (root < 0 || root >= n) continue;
                ++valid;
                // the reference calls compute_dpt with its default tolerance
                dpt_device(S, root, one.get(), iters, kDefaultTol);
                launch_add(sum.get(), one.get(), n, S.s);
            }
            if (valid == 0) {
                for (int64_t i = 0; i < n; ++i) pseudotime[i] = 1.0f;
            } else {
                launch_scale(sum.get(), n, 1.0f / (float)valid, S.s);
                S.download(pseudotime, sum.get(), n);
            }
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_diffusion_diffusion_map_embedding(scl_sparse_t transition,
                                                             scl_real_t* embedding,
                                                             scl_index_t n_nodes,
                                                             scl_index_t n_components,
                                                             scl_index_t n_iter) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(embedding, "Output embedding array is null");
    SCL_API_CHECK(n_nodes > 0 && n_components > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        Session S(transition, (int64_t)n_nodes);
        const int64_t n = S.n, K = n_components;
        if (n > 0) {
            if (K > std::numeric_limits<int32_t>::max() / 2)
                fail(SCL_ERROR_INVALID_ARGUMENT, "n_components too large");
            const int64_t iters = n_iter == 0 ? kDefaultNIter : n_iter;
            DevBuf Q(n * K), Qn(n * K), w(n);
            // one workgroup orthogonalises small bases; from kWideGramSchmidtRows
            // rows a Cholesky QR (four launches) replaces it, or, with more
            // than kCholQrMaxK columns or a degenerate basis, the columns are
            // transposed to contiguous vectors and every dot product of the
            // Gram-Schmidt spreads over the GPU
            const bool wide = n >= kWideGramSchmidtRows;
            const bool chol = wide && K <= kCholQrMaxK;
            DevBuf C(wide ? n * K : 0), part(wide ? gram_schmidt_part_floats(n, K) : 0),
                sc(wide ? 8 : 0), cw(chol ? cholqr_workspace_doubles(n, K) : 0, sizeof(double)),
                flag(1, sizeof(int));
            auto orthogonalise = [&](float* M, bool use_chol) {
                if (!wide) {
                    launch_mgs(M, w.get(), n, K, S.s);
                } else if (use_chol) {
                    launch_cholqr(M, n, K, cw.as<double>(), flag.as<int>(), S.s);
                } else {
                    launch_transpose_rect(M, C.get(), n, K, S.s);
                    launch_gram_schmidt(C.get(), w.get(), n, K, part.get(), sc.get(), true, 0,
                                        S.s);
                    launch_transpose_rect(C.get(), M, K, n, S.s);
                }
            };
            // the Cholesky path iterates in the operator's locality numbering
            // (its renumbered copy: the products gather neighbouring rows);
            // the orthonormalisation does not depend on the row order
            auto embed = [&](bool use_chol) {
                const bool ren = use_chol && S.renumbered();
                hip_check(hipMemsetAsync(flag.get(), 0, sizeof(int), S.s), "clear flag");
                launch_dmap_init(ren ? Qn.get() : Q.get(), n, K, S.s);
                if (ren) S.renumber_rows(Qn.get(), Q.get(), K, false);
                orthogonalise(Q.get(), use_chol);
                float* cur = Q.get();
                float* nxt = Qn.get();
                for (int64_t it = 0; it < iters; ++it) {
                    if (ren) S.spmm_renumbered(cur, nxt, K);
                    else S.spmm(cur, nxt, K);
                    orthogonalise(nxt, use_chol);
                    std::swap(cur, nxt);
                }
                if (ren) {
                    S.renumber_rows(cur, nxt, K, true);
                    std::swap(cur, nxt);
                }
                return cur;
            };
            float* out = embed(chol);
            if (chol) {
                int degenerate = 0;
                S.download(reinterpret_cast<float*>(&degenerate), flag.get(), 1);
                if (degenerate) out = embed(false);
            }
            S.download(embedding, out, n * K);
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_diffusion_diffusion_distance(scl_sparse_t transition,
                                                        scl_real_t* distances,
                                                        scl_size_t n_nodes,
                                                        scl_index_t n_steps) {
    SCL_API_CHECK_NULL(transition, "Transition matrix is null");
    SCL_API_CHECK_NULL(distances, "Output distances array is null");
    SCL_API_CHECK(n_nodes > 0 && n_steps > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Dimensions must be positive");
    SCL_API_TRY
        Session S(transition, (int64_t)n_nodes);
        const int64_t n = S.n;
        if (n > 0) {
            if (n > std::numeric_limits<int32_t>::max() / 2)
                fail(SCL_ERROR_INVALID_ARGUMENT, "too many nodes for a dense distance matrix");
            // T = P^t from the identity, one SpMM step per diffusion step
            // (column i of T is the vector P^t e_i)
            DevBuf A(n * n), B(n * n);
            launch_identity(A.get(), n, S.s);
            float* cur = A.get();
            float* nxt = B.get();
            for (int64_t st = 0; st < n_steps; ++st) {
                S.spmm(cur, nxt, n);
                std::swap(cur, nxt);
            }
            // n <= 1000: distances between rows of T; above, between its
            // columns (diffusion.hpp:1034-1116)
            const float* rows = cur;
            if (n > 1000) {
                launch_transpose_square(cur, nxt, n, S.s);
                rows = nxt;
                std::swap(cur, nxt);  // the transpose now lives in cur
            }
            launch_pairwise_rows(rows, nxt, n, S.s);
            S.download(distances, nxt, n * n);
        }
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // extern "C"
