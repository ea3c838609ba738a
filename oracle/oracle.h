/*
 * oracle.h -- CPU restatement of scl-core's graph-diffusion path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (libscl_f32_i64.so) never links,
 * loads or calls it.
 *
 * Two families of functions:
 *   orc_*_f32   restate the reference's FP32 algorithm step by step (same
 *               loop structure, same per-row accumulation order, FMA where the
 *               reference's Highway MulAdd / contracted "+= a*b" produce one).
 *   orc_*_f64   an FP64 oracle of the same mathematics (double accumulation),
 *               used to state errors; it is not a restatement of any code.
 *
 * Index arrays are int64 as at the reference ABI (scl_index_t = int64_t).
 * Reference citations are relative to the scl-core source tree.
 *
 * Parity status of each function is recorded next to it; every function is
 * pinned by tests in tests/test_oracle*.py (closed forms, reference known
 * answers, published generator definitions, invariants, brute force).  Two
 * orders taken from Google Highway, which the reference fetches and which is
 * not in its tree, are parity unpinned: the vector Exp of the Gaussian
 * kernel batch (construct.c) and the SumOfLanes tree of sum_squared
 * (neighbors.c); both affect rounding only.
 */
#ifndef SCL_ORACLE_H
#define SCL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One step Y = P X, following scl/kernel/diffusion.hpp:227-278 (spmm_block):
 * zero Y, then for each row i and each stored nonzero k in stored order,
 * Y[i,:] = fma(v_k, X[j_k,:], Y[i,:]) (axpy_simd, diffusion.hpp:100-134).
 * OpenMP static over rows when n >= 256 (diffusion.hpp:240); results do not
 * depend on the thread count. */
void orc_spmm_step_f32(int64_t n, int64_t F, const int64_t* indptr,
                       const int64_t* indices, const float* vals,
                       const float* X, float* Y);

/* X <- P^t X in place, following diffusion.hpp:638-664 (diffuse_matrix):
 * value-initialised n*F temporary, t x { spmm_block; memcpy back }.
 * Returns 0, or -1 when the temporary cannot be allocated. */
int orc_diffuse_matrix_f32(int64_t n, int64_t F, int64_t t,
                           const int64_t* indptr, const int64_t* indices,
                           const float* vals, float* X);

/* Column subset of orc_diffuse_matrix_f32: X_sub is n x n_sub (row-major) and
 * holds columns of the full X; since columns never mix, the result equals the
 * same columns of the full computation bit for bit. */
int orc_diffuse_columns_f32(int64_t n, int64_t n_sub, int64_t t,
                            const int64_t* indptr, const int64_t* indices,
                            const float* vals, float* X_sub);

/* x <- P^t x, following diffusion.hpp:609-632 (diffuse_vector) and
 * spmv_parallel 169-221: sum = fma(v_k, x[j_k], sum) in stored order. */
int orc_diffuse_vector_f32(int64_t n, int64_t t, const int64_t* indptr,
                           const int64_t* indices, const float* vals,
                           float* x);

/* In-place normalisation of values, following diffusion.hpp:492-603
 * (compute_transition_matrix) with compute_row_sums_parallel 452-484:
 * FP32 sequential row sums; symmetric == 0: v *= (s > 1e-15f ? 1/s : 0);
 * symmetric != 0: v *= d_i^-1/2 * d_j^-1/2. */
void orc_transition_f32(int64_t n, const int64_t* indptr,
                        const int64_t* indices, float* vals, int symmetric);

/* FP64 normalisation (what the device computes): d_i = sum_k (double)a_ik in
 * stored order; symmetric == 0: v = (float)(a / d_i), 0 when d_i <= 1e-15;
 * symmetric != 0: v = (float)(a * (1/sqrt(d_i)) * (1/sqrt(d_j))).
 * Writes d to deg (length n) when deg != NULL. */
void orc_transition_f64(int64_t n, const int64_t* indptr,
                        const int64_t* indices, float* vals, int symmetric,
                        double* deg);

/* Reduced-precision restatement: Y = P^t X with the t-1 intermediate
 * iterates rounded to bf16 (round to nearest even) and fp32 FMA accumulation
 * in stored order; the input X and the output Y are fp32; P's values are
 * rounded to bf16 first when bf16_values != 0.  This is the arithmetic the
 * bf16 device path performs (the reference has no reduced-precision path;
 * parity of this function is pinned by tests against the FP32 restatement
 * and closed forms with bf16-exact data). */
int orc_diffuse_bf16_f32(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                         const int64_t* indices, const float* vals,
                         int bf16_values, const float* X, float* Y);
/* fp32 -> nearest-even bf16, returned as fp32. */
float orc_bf16_round(float f);

/* The same with fp16 iterates when fp16 == 1: each intermediate iterate is
 * stored as round_fp16(y * 2^scale_exp) and read back times 2^-scale_exp
 * (round to nearest even, overflow to inf, subnormals kept); compensated
 * bf16 when fp16 == 2 (orc_comp_round: values too when bf16_values); bf16
 * iterates otherwise (scale_exp ignored).  round_input != 0 stores X the
 * same way before the first step. */
int orc_diffuse_lowp_f32(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                         const int64_t* indices, const float* vals,
                         int bf16_values, int fp16, int scale_exp, int round_input,
                         const float* X, float* Y);
/* fp32 -> nearest-even binary16, returned as fp32. */
float orc_fp16_round(float f);
/* fp32 -> compensated bf16 -> fp32: rounded once to 24 bits (15 mantissa
 * bits, nearest even; nan kept quiet) and read back exactly.  The device's
 * compensated storage (a bf16 plane of the upper 16 bits plus an 8-bit
 * plane of the next 8; no reference counterpart; pinned by its error bound,
 * rounding direction, idempotence, symmetry and exact cases in
 * tests/test_oracle.py). */
float orc_comp_round(float f);

/* FP64 oracle: Y = P^t X with double values/accumulators, X and Y double. */
int orc_diffuse_f64(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                    const int64_t* indices, const double* vals,
                    const double* X, double* Y);

/* FP64 oracle of the transposed operator: Y = (P^T)^t X. */
int orc_diffuse_transpose_f64(int64_t n, int64_t F, int64_t t,
                              const int64_t* indptr, const int64_t* indices,
                              const double* vals, const double* X, double* Y);

/* FP64 reference gradient of L = 0.5 ||P^t X - T||^2:
 *   dX = (P^T)^t (Y - T),   dP_k (k = nonzero (i,j)) = sum_{s=1..t} <G_s[i,:], X_{s-1}[j,:]>
 * with G_t = Y - T and G_{s-1} = P^T G_s.  Writes Y, dX (n x F) and dP (nnz).
 * Returns L. */
double orc_grad_f64(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                    const int64_t* indices, const double* vals,
                    const double* X, const double* T, double* Y, double* dX,
                    double* dP);

/* FP64 loss only (for finite differences). */
double orc_loss_f64(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                    const int64_t* indices, const double* vals,
                    const double* X, const double* T);

/* FP64 stationarity residual r = || P^T pi - pi ||_1 with pi = deg / sum(deg),
 * P given in FP32 values; counterpart of scl/kernel/transition.hpp:398-416
 * (sparse_matvec_transpose) used by stationary_distribution 571-627. */
double orc_stationarity_residual_f64(int64_t n, const int64_t* indptr,
                                     const int64_t* indices, const float* vals,
                                     const double* deg);

/* Relative error in the reference tests' convention
 * (tests/c_api/include/precision.hpp:102-120): max|a-b| / max(max|a|,max|b|),
 * or max|a-b| when both are below 1e-15. a is float, b is double. */
double orc_rel_err_f32_f64(size_t len, const float* a, const double* b);
double orc_rel_err_f32_f32(size_t len, const float* a, const float* b);
/* Median over elements of |a-b| / max|b|. */
double orc_median_rel_err_f32_f64(size_t len, const float* a, const double* b);

/* Thread count the restatement runs with (omp_get_max_threads). */
int orc_num_threads(void);
/* Sets the OpenMP thread count of later calls (n > 0). */
void orc_set_num_threads(int n);

/* ---- vector-side consumers (walks.c): diffusion.hpp:670-1194 ---------- */
int orc_lazy_walk_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                      float* x, int64_t n_steps, float laziness);
int orc_heat_kernel_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                        float* sig, float t, int64_t n_steps);
int64_t orc_rwr_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                    const int64_t* seeds, int64_t n_seeds, float* scores, float alpha,
                    int64_t max_iter, float tol);
int orc_dpt_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                int64_t root, float* pt, int64_t max_iter, float tol);
int orc_dpt_multi_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                      const int64_t* roots, int64_t n_roots, float* pt, int64_t max_iter);
void orc_dmap_init_f32(int64_t n, int64_t K, float* Q);
int orc_dmap_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                 float* emb, int64_t K, int64_t n_iter);
/* The same in FP64 from the same fp32 start vectors (an error anchor). */
int orc_dmap_f64(int64_t n, const int64_t* ip, const int64_t* ix, const float* v, double* emb,
                 int64_t K, int64_t n_iter);
int orc_mgs_f32(int64_t n, int64_t K, float* Q);
int orc_diffusion_distance_f32(int64_t n, const int64_t* ip, const int64_t* ix,
                               const float* v, float* D, int64_t n_steps);

/* ---- other callers of the SpMM (callers.c) ---- */
/* Dense copy of a sparse cells x genes matrix (impute.hpp:612-640, 683-697). */
void orc_densify_f32(int64_t n_cells, int64_t n_genes, const int64_t* ip,
                     const int64_t* ix, const float* v, int is_csr, float* out);
/* diffusion imputation steps (impute.hpp:599-658) on the dense X (n x G). */
int orc_impute_diffusion_f32(int64_t n, int64_t G, int64_t n_steps, const int64_t* ip,
                             const int64_t* ix, const float* v, float* X);
/* MAGIC steps with row-sum normalisation (impute.hpp:669-741, 355-389). */
int orc_magic_f32(int64_t n, int64_t G, int64_t t, const int64_t* ip, const int64_t* ix,
                  const float* v, float* X);
/* FastRNG of pseudotime.hpp:153-182: seeding (the four 64-bit z and the
 * state), raw steps from a state, and uniform() draws from a seed. */
void orc_fastrng_seed(uint64_t seed, uint64_t* z4, uint32_t* state4);
void orc_fastrng_step(uint32_t* state4, int64_t m, uint32_t* out);
void orc_fastrng_uniform(uint64_t seed, int64_t m, float* out);
/* orthonormalize_columns (pseudotime.hpp:418-474) of row-major n x K Q. */
int orc_pt_orthonormalise_f32(int64_t n, int64_t K, float* Q);
/* diffusion_pseudotime (pseudotime.hpp:646-712). */
int orc_pseudotime_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                       int64_t root, int64_t K, int64_t n_iter, float* pt);
/* The same in FP64 from the same fp32 start (an error anchor). */
int orc_pseudotime_f64(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                       int64_t root, int64_t K, int64_t n_iter, double* pt);
/* algebra::spmm, Y = alpha A X + beta Y (algebra.hpp:546-609); A has m rows. */
void orc_algebra_spmm_f32(int64_t m, int64_t F, const int64_t* ip, const int64_t* ix,
                          const float* v, const float* X, float* Y, float alpha,
                          float beta);

/* ---- construction of the operator (construct.c) ---- */
/* compute_kernel_matrix (kernel.hpp:530-590): kernel of each stored distance;
 * type is the C ABI's scl_kernel_type_t value. */
void orc_kernel_matrix_f32(int64_t nnz, const float* d, float bandwidth, int type, float* out);
/* kernel_row_sums (kernel.hpp:596-645). */
void orc_kernel_row_sums_f32(int64_t n, const int64_t* ip, const float* d, float bandwidth,
                             int type, float* sums);
/* vector_sum of transition.hpp:165-179 (8 strided partials, then the tail). */
float orc_vector_sum8_f32(const float* v, int64_t n);
/* row_normalize_to_stochastic (transition.hpp:493-525). */
void orc_row_normalize_f32(int64_t n, const int64_t* ip, const float* v, float* out);
/* symmetrize_transition (transition.hpp:530-565), dense n x n output. */
void orc_symmetrize_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                        float* S);
/* transition_matrix_from_velocity (transition.hpp:449-487), dense n x n. */
void orc_from_velocity_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                           float* S);
/* stationary_distribution (transition.hpp:571-623); returns iterations run. */
int64_t orc_stationary_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                           int64_t max_iter, float tol, float* pi);

/* ---- k-nearest neighbours (neighbors.c) ---- */
/* vectorize::sum_squared, 8-lane float model (vectorize.hpp:782-822). */
float orc_sum_squared_f32(const float* v, int64_t n);
/* compute_norms (neighbors.hpp:303-317). */
void orc_norms_f32(int64_t n, const int64_t* ip, const float* v, float* norms);
/* knn (neighbors.hpp:319-476): out_idx / out_dist are n x k. */
int orc_knn_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                const float* norms, int64_t k, int64_t* out_idx, float* out_dist);

#ifdef __cplusplus
}
#endif
#endif
