/*
 * oracle.c -- CPU restatement + FP64 oracle of scl-core's diffusion path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): never linked into the product.
 * Built by oracle/Makefile with -ffp-contract=off, so every fused multiply-add
 * below is an explicit fmaf()/fma() and nothing else is fused.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* scl/kernel/diffusion.hpp:40 (config::PARALLEL_THRESHOLD) */
#define ORC_PARALLEL_THRESHOLD 256
/* scl/kernel/diffusion.hpp:41 (config::MIN_PROB), as Real = float */
#define ORC_MIN_PROB_F32 1e-15f
#define ORC_MIN_PROB_F64 1e-15

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of later parallel regions (a timing knob of the CPU baseline) */
void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* axpy_simd (diffusion.hpp:100-134): y[c] += a * x[c], one FMA per element.
 * Highway's MulAdd is a fused multiply-add on FMA targets; the scalar tail
 * `y[k] += alpha * x[k]` (131-133) is contracted to an FMA by GCC/Clang at
 * -O3 with their default contraction; both give fmaf(a, x, y). */
static inline void orc_axpy_f32(float a, const float* restrict x,
                                float* restrict y, int64_t F) {
    for (int64_t c = 0; c < F; ++c) y[c] = fmaf(a, x[c], y[c]);
}

/* spmm_block (diffusion.hpp:227-278). */
static void orc_spmm_strided_f32(int64_t n, int64_t F, const int64_t* indptr,
                                 const int64_t* indices, const float* vals,
                                 const float* X, float* Y) {
    /* diffusion.hpp:238 -- zero the whole output first */
    memset(Y, 0, (size_t)n * (size_t)F * sizeof(float));
    /* diffusion.hpp:240-257 (parallel) / 258-277 (serial): same order */
#pragma omp parallel for schedule(static) if (n >= ORC_PARALLEL_THRESHOLD)
    for (int64_t i = 0; i < n; ++i) {
        float* Yi = Y + i * F;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) {
            const int64_t j = indices[k];
            orc_axpy_f32(vals[k], X + j * F, Yi, F);
        }
    }
}

void orc_spmm_step_f32(int64_t n, int64_t F, const int64_t* indptr,
                       const int64_t* indices, const float* vals,
                       const float* X, float* Y) {
    orc_spmm_strided_f32(n, F, indptr, indices, vals, X, Y);
}

/* diffuse_matrix (diffusion.hpp:638-664). */
int orc_diffuse_matrix_f32(int64_t n, int64_t F, int64_t t,
                           const int64_t* indptr, const int64_t* indices,
                           const float* vals, float* X) {
    if (n == 0 || F == 0 || t == 0) return 0; /* diffusion.hpp:652 */
    const size_t total = (size_t)n * (size_t)F;
    /* diffusion.hpp:655: aligned_alloc value-initialises (memory.hpp:71) */
    float* Xn = (float*)calloc(total, sizeof(float));
    if (!Xn) return -1;
    for (int64_t s = 0; s < t; ++s) {               /* 658-661 */
        orc_spmm_strided_f32(n, F, indptr, indices, vals, X, Xn);
        memcpy(X, Xn, total * sizeof(float));
    }
    free(Xn);
    return 0;
}

int orc_diffuse_columns_f32(int64_t n, int64_t n_sub, int64_t t,
                            const int64_t* indptr, const int64_t* indices,
                            const float* vals, float* X_sub) {
    return orc_diffuse_matrix_f32(n, n_sub, t, indptr, indices, vals, X_sub);
}

/* diffuse_vector (diffusion.hpp:609-632) over spmv_parallel (169-221).  The
 * 4-way unrolled body adds into one accumulator in stored order, so the
 * result is the plain sequential chain sum = fma(v, x[j], sum). */
int orc_diffuse_vector_f32(int64_t n, int64_t t, const int64_t* indptr,
                           const int64_t* indices, const float* vals,
                           float* x) {
    if (n == 0 || t == 0) return 0;
    float* xn = (float*)calloc((size_t)n, sizeof(float));
    if (!xn) return -1;
    for (int64_t s = 0; s < t; ++s) {
#pragma omp parallel for schedule(static) if (n >= ORC_PARALLEL_THRESHOLD)
        for (int64_t i = 0; i < n; ++i) {
            float sum = 0.0f;
            for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k)
                sum = fmaf(vals[k], x[indices[k]], sum);
            xn[i] = sum;
        }
        memcpy(x, xn, (size_t)n * sizeof(float));
    }
    free(xn);
    return 0;
}

/* compute_transition_matrix (diffusion.hpp:492-603). */
void orc_transition_f32(int64_t n, const int64_t* indptr,
                        const int64_t* indices, float* vals, int symmetric) {
    float* s = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    /* compute_row_sums_parallel (452-484): sequential float sum per row */
    for (int64_t i = 0; i < n; ++i) {
        float acc = 0.0f;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) acc += vals[k];
        s[i] = acc;
    }
    if (symmetric) {
        /* 508-564: d_inv_sqrt = s > MIN_PROB ? 1/sqrt(s) : 0; v *= di * dj */
        float* dis = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
        for (int64_t i = 0; i < n; ++i)
            dis[i] = (s[i] > ORC_MIN_PROB_F32) ? 1.0f / sqrtf(s[i]) : 0.0f;
        for (int64_t i = 0; i < n; ++i) {
            const float di = dis[i];
            for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k)
                vals[k] *= di * dis[indices[k]];
        }
        free(dis);
    } else {
        /* 565-600: inv_sum = s > MIN_PROB ? 1/s : 0; v *= inv_sum */
        for (int64_t i = 0; i < n; ++i) {
            const float inv = (s[i] > ORC_MIN_PROB_F32) ? 1.0f / s[i] : 0.0f;
            for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) vals[k] *= inv;
        }
    }
    free(s);
}

void orc_transition_f64(int64_t n, const int64_t* indptr,
                        const int64_t* indices, float* vals, int symmetric,
                        double* deg) {
    double* d = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) acc += (double)vals[k];
        d[i] = acc;
        if (deg) deg[i] = acc;
    }
    if (symmetric) {
        double* dis = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        for (int64_t i = 0; i < n; ++i)
            dis[i] = (d[i] > ORC_MIN_PROB_F64) ? 1.0 / sqrt(d[i]) : 0.0;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k)
                vals[k] = (float)(((double)vals[k] * dis[i]) * dis[indices[k]]);
        free(dis);
    } else {
        for (int64_t i = 0; i < n; ++i) {
            const double di = d[i];
            for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k)
                vals[k] = (di > ORC_MIN_PROB_F64) ? (float)((double)vals[k] / di) : 0.0f;
        }
    }
    free(d);
}

/* ------------------------------------------------------- bf16 iterates -- */

/* fp32 -> bf16 -> fp32, round to nearest even (NaN kept quiet) */
static inline float orc_round_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) {
        u = (u | 0x00400000u) & 0xffff0000u;
    } else {
        u = (u + (((u >> 16) & 1u) + 0x7fffu)) & 0xffff0000u;
    }
    float r;
    memcpy(&r, &u, 4);
    return r;
}

float orc_bf16_round(float f) { return orc_round_bf16(f); }

/* fp32 -> binary16 -> fp32, round to nearest even; overflow -> inf,
 * subnormals kept (quantum 2^-24). */
static inline float orc_round_fp16(float f) {
    if (f != f) return f;
    const float a = fabsf(f);
    if (a >= 65520.0f) return copysignf(INFINITY, f);
    if (a < 6.103515625e-05f) {
        /* a * 2^24 < 2^10 is exact; rintf rounds half to even */
        const float r = rintf(a * 16777216.0f) / 16777216.0f;
        return copysignf(r, f);
    }
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u + 0x0fffu + ((u >> 13) & 1u)) & 0xffffe000u;
    float r;
    memcpy(&r, &u, 4);
    if (fabsf(r) > 65504.0f) return copysignf(INFINITY, f);
    return r;
}

float orc_fp16_round(float f) { return orc_round_fp16(f); }

/* compensated bf16: q rounded once to 24 bits (round to nearest even at
 * bit 8 of the fp32 word; nan kept quiet), stored as its upper 16 bits (a
 * bf16) and the next 8, and read back as that 24-bit word shifted left 8 */
static inline float orc_round_comp(float q) {
    uint32_t u;
    memcpy(&u, &q, 4);
    const uint32_t w = ((u & 0x7fffffffu) > 0x7f800000u) ? ((u >> 8) | 0x4000u)
                                                         : ((u + 0x7fu + ((u >> 8) & 1u)) >> 8);
    const uint32_t r = w << 8;
    float f;
    memcpy(&f, &r, 4);
    return f;
}

float orc_comp_round(float f) { return orc_round_comp(f); }

int orc_diffuse_lowp_f32(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                         const int64_t* indices, const float* vals,
                         int bf16_values, int fp16, int scale_exp, int round_input,
                         const float* X, float* Y) {
    const size_t total = (size_t)n * (size_t)F;
    const int64_t nnz = indptr[n];
    float* v = (float*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(float));
    float* a = (float*)malloc(total * sizeof(float));
    float* b = (float*)malloc(total * sizeof(float));
    if (!v || !a || !b) { free(v); free(a); free(b); return -1; }
    const float up = ldexpf(1.0f, scale_exp), down = ldexpf(1.0f, -scale_exp);
    /* fp16: 0 bf16, 1 fp16 (scaled by 2^scale_exp), 2 compensated bf16 */
    for (int64_t k = 0; k < nnz; ++k)
        v[k] = !bf16_values ? vals[k]
               : fp16 == 2  ? orc_round_comp(vals[k])
                            : orc_round_bf16(vals[k]);
    memcpy(a, X, total * sizeof(float));
    if (round_input) {
        /* X stored as round16(X * 2^k) and read back as that * 2^-k */
        if (fp16 == 1)
            for (size_t e = 0; e < total; ++e) a[e] = orc_round_fp16(a[e] * up) * down;
        else if (fp16 == 2)
            for (size_t e = 0; e < total; ++e) a[e] = orc_round_comp(a[e]);
        else
            for (size_t e = 0; e < total; ++e) a[e] = orc_round_bf16(a[e]);
    }
    for (int64_t s = 1; s <= t; ++s) {
        orc_spmm_strided_f32(n, F, indptr, indices, v, a, b);
        if (s < t) {
            /* stored as round16(y * 2^k), read back as that * 2^-k */
            if (fp16 == 1)
                for (size_t e = 0; e < total; ++e) b[e] = orc_round_fp16(b[e] * up) * down;
            else if (fp16 == 2)
                for (size_t e = 0; e < total; ++e) b[e] = orc_round_comp(b[e]);
            else
                for (size_t e = 0; e < total; ++e) b[e] = orc_round_bf16(b[e]);
        }
        float* tmp = a; a = b; b = tmp;
    }
    memcpy(Y, a, total * sizeof(float));
    free(v); free(a); free(b);
    return 0;
}

int orc_diffuse_bf16_f32(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                         const int64_t* indices, const float* vals,
                         int bf16_values, const float* X, float* Y) {
    return orc_diffuse_lowp_f32(n, F, t, indptr, indices, vals, bf16_values, 0, 0, 0, X, Y);
}

/* ---------------------------------------------------------------- FP64 -- */

static void orc_spmm_f64(int64_t n, int64_t F, const int64_t* indptr,
                         const int64_t* indices, const double* vals,
                         const double* X, double* Y) {
#pragma omp parallel for schedule(static) if (n >= ORC_PARALLEL_THRESHOLD)
    for (int64_t i = 0; i < n; ++i) {
        double* Yi = Y + i * F;
        for (int64_t c = 0; c < F; ++c) Yi[c] = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) {
            const double v = vals[k];
            const double* Xj = X + indices[k] * F;
            for (int64_t c = 0; c < F; ++c) Yi[c] += v * Xj[c];
        }
    }
}

/* Y = P^T X via a row-parallel pass over the transposed pattern, built by a
 * counting sort (the FP64 restatement of transition.hpp:398-416's scatter,
 * reordered so no two threads write one output row). */
typedef struct { int64_t* ptr; int64_t* idx; double* val; } orc_csr_t;

static int orc_transpose(int64_t n, const int64_t* indptr,
                         const int64_t* indices, const double* vals,
                         orc_csr_t* out) {
    const int64_t nnz = indptr[n];
    out->ptr = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    out->idx = (int64_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
    out->val = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    if (!out->ptr || !out->idx || !out->val) return -1;
    for (int64_t k = 0; k < nnz; ++k) out->ptr[indices[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) out->ptr[i + 1] += out->ptr[i];
    int64_t* cur = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!cur) return -1;
    memcpy(cur, out->ptr, (size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) {
            const int64_t p = cur[indices[k]]++;
            out->idx[p] = i;
            out->val[p] = vals[k];
        }
    free(cur);
    return 0;
}

static void orc_free_csr(orc_csr_t* m) {
    free(m->ptr); free(m->idx); free(m->val);
}

int orc_diffuse_f64(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                    const int64_t* indices, const double* vals,
                    const double* X, double* Y) {
    const size_t total = (size_t)n * (size_t)F;
    double* A = (double*)malloc(total * sizeof(double));
    if (!A) return -1;
    memcpy(Y, X, total * sizeof(double));
    for (int64_t s = 0; s < t; ++s) {
        orc_spmm_f64(n, F, indptr, indices, vals, Y, A);
        memcpy(Y, A, total * sizeof(double));
    }
    free(A);
    return 0;
}

int orc_diffuse_transpose_f64(int64_t n, int64_t F, int64_t t,
                              const int64_t* indptr, const int64_t* indices,
                              const double* vals, const double* X, double* Y) {
    orc_csr_t T;
    if (orc_transpose(n, indptr, indices, vals, &T)) return -1;
    const int rc = orc_diffuse_f64(n, F, t, T.ptr, T.idx, T.val, X, Y);
    orc_free_csr(&T);
    return rc;
}

double orc_loss_f64(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                    const int64_t* indices, const double* vals,
                    const double* X, const double* T) {
    const size_t total = (size_t)n * (size_t)F;
    double* Y = (double*)malloc(total * sizeof(double));
    if (!Y) return NAN;
    orc_diffuse_f64(n, F, t, indptr, indices, vals, X, Y);
    double L = 0.0;
    for (size_t e = 0; e < total; ++e) {
        const double r = Y[e] - T[e];
        L += 0.5 * r * r;
    }
    free(Y);
    return L;
}

double orc_grad_f64(int64_t n, int64_t F, int64_t t, const int64_t* indptr,
                    const int64_t* indices, const double* vals,
                    const double* X, const double* T, double* Y, double* dX,
                    double* dP) {
    const size_t total = (size_t)n * (size_t)F;
    const int64_t nnz = indptr[n];
    /* keep every iterate X_0..X_t */
    double* it = (double*)malloc(total * (size_t)(t + 1) * sizeof(double));
    double* G = (double*)malloc(total * sizeof(double));
    double* Gn = (double*)malloc(total * sizeof(double));
    orc_csr_t PT;
    if (!it || !G || !Gn || orc_transpose(n, indptr, indices, vals, &PT)) {
        free(it); free(G); free(Gn);
        return NAN;
    }
    memcpy(it, X, total * sizeof(double));
    for (int64_t s = 1; s <= t; ++s)
        orc_spmm_f64(n, F, indptr, indices, vals, it + (size_t)(s - 1) * total,
                     it + (size_t)s * total);
    memcpy(Y, it + (size_t)t * total, total * sizeof(double));
    double L = 0.0;
    for (size_t e = 0; e < total; ++e) {
        G[e] = Y[e] - T[e];
        L += 0.5 * G[e] * G[e];
    }
    for (int64_t k = 0; k < nnz; ++k) dP[k] = 0.0;
    for (int64_t s = t; s >= 1; --s) {
        /* dP_ij += <G_s[i,:], X_{s-1}[j,:]> */
        const double* Xp = it + (size_t)(s - 1) * total;
#pragma omp parallel for schedule(static) if (n >= ORC_PARALLEL_THRESHOLD)
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) {
                const double* gi = G + i * F;
                const double* xj = Xp + indices[k] * F;
                double acc = 0.0;
                for (int64_t c = 0; c < F; ++c) acc += gi[c] * xj[c];
                dP[k] += acc;
            }
        /* G_{s-1} = P^T G_s */
        orc_spmm_f64(n, F, PT.ptr, PT.idx, PT.val, G, Gn);
        memcpy(G, Gn, total * sizeof(double));
    }
    memcpy(dX, G, total * sizeof(double));
    orc_free_csr(&PT);
    free(it); free(G); free(Gn);
    return L;
}

double orc_stationarity_residual_f64(int64_t n, const int64_t* indptr,
                                     const int64_t* indices, const float* vals,
                                     const double* deg) {
    double D = 0.0;
    for (int64_t i = 0; i < n; ++i) D += deg[i];
    double* y = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!y) return NAN;
    /* y = P^T pi, serial scatter as transition.hpp:398-416 */
    for (int64_t i = 0; i < n; ++i) {
        const double pi = deg[i] / D;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k)
            y[indices[k]] += (double)vals[k] * pi;
    }
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += fabs(y[i] - deg[i] / D);
    free(y);
    return r;
}

/* ---------------------------------------------------------------- errors -- */

double orc_rel_err_f32_f64(size_t len, const float* a, const double* b) {
    double mv = 0.0, md = 0.0;
    for (size_t e = 0; e < len; ++e) {
        const double x = (double)a[e], y = b[e];
        const double m = fabs(x) > fabs(y) ? fabs(x) : fabs(y);
        if (m > mv) mv = m;
        const double d = fabs(x - y);
        if (d > md || d != d) md = d;
    }
    return mv < 1e-15 ? md : md / mv;
}

double orc_rel_err_f32_f32(size_t len, const float* a, const float* b) {
    double mv = 0.0, md = 0.0;
    for (size_t e = 0; e < len; ++e) {
        const double x = (double)a[e], y = (double)b[e];
        const double m = fabs(x) > fabs(y) ? fabs(x) : fabs(y);
        if (m > mv) mv = m;
        const double d = fabs(x - y);
        if (d > md || d != d) md = d;
    }
    return mv < 1e-15 ? md : md / mv;
}

static int orc_cmp_double(const void* p, const void* q) {
    const double a = *(const double*)p, b = *(const double*)q;
    return (a > b) - (a < b);
}

double orc_median_rel_err_f32_f64(size_t len, const float* a, const double* b) {
    if (len == 0) return 0.0;
    double mb = 0.0;
    for (size_t e = 0; e < len; ++e)
        if (fabs(b[e]) > mb) mb = fabs(b[e]);
    double* r = (double*)malloc(len * sizeof(double));
    if (!r) return NAN;
    for (size_t e = 0; e < len; ++e)
        r[e] = fabs((double)a[e] - b[e]) / (mb > 0 ? mb : 1.0);
    qsort(r, len, sizeof(double), orc_cmp_double);
    const double med = r[len / 2];
    free(r);
    return med;
}
