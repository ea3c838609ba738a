/*
 * scl_diffusion.h -- the drop-in diffusion entry points.
 *
 * Replaces scl/binding/c_api/diffusion.h:15-35 and 99-105 (implemented in
 * scl/binding/c_api/diffusion.cpp:36-116 and 318-341).  Same signatures,
 * same argument checks and error codes, same in-place host-buffer semantics;
 * the work runs on the current HIP device and the call returns when the
 * result is back in the caller's buffer.
 *
 * Checks the reference lacks, added here (all forced from the host, none
 * touches the GPU):
 *   - n_nodes must equal the operator's rows and cols -> SCL_ERROR_DIMENSION_MISMATCH
 *   - n_nodes * n_features must fit in memory         -> SCL_ERROR_INVALID_ARGUMENT
 *   - nnz and n must fit int32 on the device          -> SCL_ERROR_INVALID_ARGUMENT
 *   - device allocation failure                       -> SCL_ERROR_OUT_OF_MEMORY
 *   - no usable GPU                                   -> SCL_ERROR_FEATURE_UNAVAILABLE
 *
 * A CSC handle is applied through its primary (column) slots, exactly as the
 * reference's visit() dispatch does (core/internal.hpp:133-151), so
 * diffusion through a CSC handle of P computes (P^T)^t X.
 */
#ifndef SCL_DIFFUSION_H
#define SCL_DIFFUSION_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* diffusion.h:15-18 / diffusion.cpp:36-62.  Normalises the handle's values in
 * place.  symmetric == 0: row-stochastic D^-1 A; symmetric != 0:
 * D^-1/2 A D^-1/2.  Row sums, reciprocals and square roots are FP64 on the
 * GPU and each value is rounded to float once; rows whose sum is <= 1e-15
 * become zero (diffusion.hpp:41, 565-600). */
SCL_EXPORT scl_error_t scl_diffusion_compute_transition_matrix(
    scl_sparse_t adjacency, int symmetric);

/* diffusion.h:21-26 / diffusion.cpp:68-88: x <- P^t x (x on the host). */
SCL_EXPORT scl_error_t scl_diffusion_diffuse_vector(scl_sparse_t transition,
                                                    scl_real_t* x,
                                                    scl_size_t n_nodes,
                                                    scl_index_t n_steps);

/* diffusion.h:29-35 / diffusion.cpp:94-116: X <- P^t X, X an n_nodes x
 * n_features row-major host buffer, overwritten in place. */
SCL_EXPORT scl_error_t scl_diffusion_diffuse_matrix(scl_sparse_t transition,
                                                    scl_real_t* X,
                                                    scl_index_t n_nodes,
                                                    scl_index_t n_features,
                                                    scl_index_t n_steps);

/* diffusion.h:99-105 / diffusion.cpp:318-341: alias of diffuse_matrix with
 * t == 0 meaning 3 (diffusion.cpp:333); t < 0 leaves X unchanged, as the
 * reference's unchecked loop does. */
SCL_EXPORT scl_error_t scl_diffusion_magic_impute(scl_sparse_t transition,
                                                  scl_real_t* X,
                                                  scl_index_t n_nodes,
                                                  scl_index_t n_features,
                                                  scl_index_t t);

/* ---- vector-side consumers (diffusion.h:38-122, diffusion.cpp:122-401) ---
 * Vectors live on the GPU for the whole call (host buffers are copied in and
 * out once); every SpMV is the stored-order FMA chain.  0 selects the
 * reference's default for alpha (0.85), tol (1e-6), max_iter (100), n_iter
 * (50), laziness (0.5) and the heat kernel's n_steps (10).  The operator
 * must be square and n_nodes must cover its rows (SCL_ERROR_DIMENSION_MISMATCH;
 * the reference checks only the latter). */

/* diffusion.h:38-45 / diffusion.hpp:670-768: diffusion pseudotime from
 * root_cell (hitting times of the walk from the root, scaled to [0, 1];
 * nodes visited with probability <= 0.5 get 1).  A root out of range is
 * SCL_ERROR_INVALID_ARGUMENT. */
SCL_EXPORT scl_error_t scl_diffusion_compute_dpt(scl_sparse_t transition,
                                                 scl_index_t root_cell,
                                                 scl_real_t* pseudotime,
                                                 scl_size_t n_nodes,
                                                 scl_index_t max_iter, scl_real_t tol);

/* diffusion.h:48-55 / diffusion.hpp:774-827: mean of the single-root
 * pseudotimes over the in-range roots (each with tolerance 1e-6); all ones
 * when no root is in range. */
SCL_EXPORT scl_error_t scl_diffusion_compute_dpt_multi_root(
    scl_sparse_t transition, const scl_index_t* root_cells, scl_size_t n_roots,
    scl_real_t* pseudotime, scl_size_t n_nodes, scl_index_t max_iter);

/* diffusion.h:58-67 / diffusion.hpp:833-895: s <- alpha P s + (1 - alpha) r,
 * r uniform over the in-range seeds, until sum |s_new - s| < tol.  Without
 * convergence an odd max_iter leaves the next-to-last iterate in scores (the
 * reference swaps its output pointer each iteration); all zero without an
 * in-range seed. */
SCL_EXPORT scl_error_t scl_diffusion_random_walk_with_restart(
    scl_sparse_t transition, const scl_index_t* seed_nodes, scl_size_t n_seeds,
    scl_real_t* scores, scl_size_t n_nodes, scl_real_t alpha, scl_index_t max_iter,
    scl_real_t tol);

/* diffusion.h:70-78 / diffusion.hpp:1122-1135: the same with one seed. */
SCL_EXPORT scl_error_t scl_diffusion_personalized_pagerank(
    scl_sparse_t transition, scl_index_t seed_node, scl_real_t* scores,
    scl_size_t n_nodes, scl_real_t alpha, scl_index_t max_iter, scl_real_t tol);

/* diffusion.h:81-87 / diffusion.hpp:901-955: n_iter rounds of
 * Q <- GramSchmidt(P Q) from a fixed cosine start (n_nodes x n_components,
 * row-major); the SpMM runs on the row-split kernels with F = n_components,
 * Gram-Schmidt in one workgroup. */
SCL_EXPORT scl_error_t scl_diffusion_diffusion_map_embedding(scl_sparse_t transition,
                                                             scl_real_t* embedding,
                                                             scl_index_t n_nodes,
                                                             scl_index_t n_components,
                                                             scl_index_t n_iter);

/* diffusion.h:90-96 / diffusion.hpp:961-1003: sig = 1, then n_steps rounds of
 * sig <- (1 - dt) sig + dt P sig with dt = t / n_steps. */
SCL_EXPORT scl_error_t scl_diffusion_heat_kernel_signature(scl_sparse_t transition,
                                                           scl_real_t* signature,
                                                           scl_size_t n_nodes, scl_real_t t,
                                                           scl_index_t n_steps);

/* diffusion.h:108-113 / diffusion.hpp:1024-1116: the n_nodes x n_nodes
 * matrix of Euclidean distances between the rows of P^t (n <= 1000) or
 * between the vectors P^t e_i (n > 1000), as the reference's two branches
 * compute.  P^t is built densely on the GPU (2 n^2 floats of device memory). */
SCL_EXPORT scl_error_t scl_diffusion_diffusion_distance(scl_sparse_t transition,
                                                        scl_real_t* distances,
                                                        scl_size_t n_nodes,
                                                        scl_index_t n_steps);

/* diffusion.h:116-122 / diffusion.hpp:1157-1194: n_steps rounds of
 * x <- laziness x + (1 - laziness) P x. */
SCL_EXPORT scl_error_t scl_diffusion_lazy_random_walk(scl_sparse_t transition,
                                                      scl_real_t* x, scl_size_t n_nodes,
                                                      scl_index_t n_steps,
                                                      scl_real_t laziness);

#ifdef __cplusplus
}
#endif
#endif
