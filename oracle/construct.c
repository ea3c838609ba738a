/*
 * construct.c -- CPU restatement of the steps that build the diffusion
 * operator in scl-core: kernel weights of kNN distances
 * (scl/kernel/kernel.hpp:530-645), row normalisation, symmetrisation and the
 * velocity transition (scl/kernel/transition.hpp:449-565) and the stationary
 * distribution (transition.hpp:571-623).
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: loaded by tests/ as the checker.
 *
 * transition.hpp's vector helpers (97-293) are the AVX/AVX2 forms that any
 * current x86 build with -march=native selects: vector_sum keeps 8 strided
 * partial sums, adds them left to right, then adds the tail; max_abs_diff
 * is a maximum (order-free).  The build has -ffp-contract=off: every fused
 * multiply-add below is an explicit fmaf, as the reference's compiler
 * contracts "y += a * x" and "1 + a * b".
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define T_EPS 1e-15f

/* kernel.hpp:25-52 kernel kinds (the C ABI's scl_kernel_type_t values) */
enum { K_GAUSS = 0, K_EPAN = 1, K_COS = 2, K_LIN = 3, K_POLY = 4, K_LAPL = 5, K_CAUCHY = 6,
       K_SIGM = 7, K_UNIF = 8, K_TRI = 9 };

typedef struct {
    float inv_h, inv_h2, inv_2h2, h2;
    int type;
} kparams;

/* KernelParams (kernel.hpp:108-122) */
static kparams kp_make(float h, int type) {
    kparams p;
    const float hh = h > 1e-10f ? h : 1e-10f;
    p.inv_h = 1.0f / hh;
    p.inv_h2 = p.inv_h * p.inv_h;
    p.inv_2h2 = 0.5f * p.inv_h2;
    p.h2 = hh * hh;
    p.type = type;
    return p;
}

/* apply_kernel (kernel.hpp:125-144) with the kernels of 58-104; kinds
 * without a formula (linear, polynomial, sigmoid) fall back to Gaussian. */
static float kp_apply(float dsq, const kparams* p) {
    switch (p->type) {
        case K_EPAN: {
            const float u2 = dsq * p->inv_h2;
            return u2 > 1.0f ? 0.0f : 0.75f * (1.0f - u2);
        }
        case K_COS: {
            const float u = sqrtf(dsq) * p->inv_h;
            const float pi = 3.14159265358979323846f;
            return u > 1.0f ? 0.0f : (pi / 4.0f) * cosf(pi * u / 2.0f);
        }
        case K_LAPL:
            return expf(-sqrtf(dsq) * p->inv_h);
        case K_CAUCHY:
            return 1.0f / fmaf(dsq, p->inv_h2, 1.0f);
        case K_UNIF:
            return dsq < p->h2 ? 0.5f : 0.0f;
        case K_TRI: {
            const float u = sqrtf(dsq) * p->inv_h;
            return u >= 1.0f ? 0.0f : 1.0f - u;
        }
        default:
            return expf(-dsq * p->inv_2h2);
    }
}

/* compute_kernel_matrix (kernel.hpp:530-590): kernel value of every stored
 * distance d, in stored order, of apply_kernel(d * d).  The reference's
 * Gaussian batch evaluates most entries with Highway's vector Exp, which
 * may differ from expf in the last bit: parity unpinned in those bits. */
void orc_kernel_matrix_f32(int64_t nnz, const float* d, float bandwidth, int type, float* out) {
    const kparams p = kp_make(bandwidth, type);
    for (int64_t k = 0; k < nnz; ++k) out[k] = kp_apply(d[k] * d[k], &p);
}

/* kernel_row_sums (kernel.hpp:596-645): sum of the row's kernel values in
 * stored order. */
void orc_kernel_row_sums_f32(int64_t n, const int64_t* ip, const float* d, float bandwidth,
                             int type, float* sums) {
    const kparams p = kp_make(bandwidth, type);
    for (int64_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s += kp_apply(d[k] * d[k], &p);
        sums[i] = s;
    }
}

/* vector_sum, AVX form (transition.hpp:165-179) */
float orc_vector_sum8_f32(const float* v, int64_t n) {
    float lane[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int l = 0; l < 8; ++l) lane[l] += v[i + l];
    float r = lane[0] + lane[1] + lane[2] + lane[3] + lane[4] + lane[5] + lane[6] + lane[7];
    for (; i < n; ++i) r += v[i];
    return r;
}

/* row_normalize_to_stochastic (transition.hpp:493-525): per row, the sum in
 * stored order (plain adds); out = v * (sum > 1e-15 ? 1 / sum : 0). */
void orc_row_normalize_f32(int64_t n, const int64_t* ip, const float* v, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s += v[k];
        const float inv = s > T_EPS ? 1.0f / s : 0.0f;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) out[k] = v[k] * inv;
    }
}

/* symmetrize_transition (transition.hpp:530-565): dense n x n
 * S = (T + T^T) / 2 accumulated entry by entry in stored order (S[i][j] and
 * S[j][i] each += v * 0.5), then every row with vector_sum > 1e-15 scaled by
 * 1 / sum. */
void orc_symmetrize_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                        float* S) {
    memset(S, 0, (size_t)n * (size_t)n * sizeof(float));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
            const int64_t j = ix[k];
            const float h = v[k] * 0.5f;
            S[i * n + j] += h;
            S[j * n + i] += h;
        }
    for (int64_t i = 0; i < n; ++i) {
        float* row = S + i * n;
        const float s = orc_vector_sum8_f32(row, n);
        if (s > T_EPS) {
            const float inv = 1.0f / s;
            for (int64_t j = 0; j < n; ++j) row[j] *= inv;
        }
    }
}

/* transition_matrix_from_velocity (transition.hpp:449-487): dense n x n;
 * row i: sum of max(v, 0) in stored order; when above 1e-15 each positive
 * entry is stored as v / sum (later entries overwrite), else S[i][i] = 1. */
void orc_from_velocity_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                           float* S) {
    memset(S, 0, (size_t)n * (size_t)n * sizeof(float));
    for (int64_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s += v[k] > 0.0f ? v[k] : 0.0f;
        float* row = S + i * n;
        if (s > T_EPS) {
            const float inv = 1.0f / s;
            for (int64_t k = ip[i]; k < ip[i + 1]; ++k)
                if (v[k] > 0.0f) row[ix[k]] = v[k] * inv;
        } else {
            row[i] = 1.0f;
        }
    }
}

/* stationary_distribution (transition.hpp:571-623): pi = 1/n; per iteration
 * y = T^T pi by the row-ordered scatter (y[j] = fma(v, pi_i, y[j]),
 * transition.hpp:399-416), y /= vector_sum(y) when that exceeds 1e-15,
 * diff = max |y - pi|, pi = y; stop when diff < tol, or when the Aitken test
 * (from the third iteration, with two shrinking differences) finds
 * (diff - d1) / (diff - 2 d1 + d2) > 0.9 with |denominator| > 1e-15.
 * Returns the number of iterations run. */
int64_t orc_stationary_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                           int64_t max_iter, float tol, float* pi) {
    if (n <= 0) return 0;
    const float init = 1.0f / (float)n;
    for (int64_t i = 0; i < n; ++i) pi[i] = init;
    float* y = (float*)malloc((size_t)n * sizeof(float));
    if (!y) return -1;
    float d1 = 1e30f, d2 = 1e30f;
    int64_t it = 0;
    for (; it < max_iter; ++it) {
        memset(y, 0, (size_t)n * sizeof(float));
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = ip[i]; k < ip[i + 1]; ++k) y[ix[k]] = fmaf(v[k], pi[i], y[ix[k]]);
        const float s = orc_vector_sum8_f32(y, n);
        if (s > T_EPS) {
            const float inv = 1.0f / s;
            for (int64_t i = 0; i < n; ++i) y[i] *= inv;
        }
        float diff = 0.0f;
        for (int64_t i = 0; i < n; ++i) {
            const float d = fabsf(y[i] - pi[i]);
            diff = d > diff ? d : diff;
        }
        memcpy(pi, y, (size_t)n * sizeof(float));
        if (diff < tol) {
            ++it;
            break;
        }
        if (it >= 2 && d1 < d2 && diff < d1) {
            const float den = diff - 2.0f * d1 + d2;
            if (fabsf(den) > T_EPS && (diff - d1) / den > 0.9f) {
                ++it;
                break;
            }
        }
        d2 = d1;
        d1 = diff;
    }
    free(y);
    return it;
}
