/*
 * callers.c -- CPU restatement of the other callers of the diffusion SpMM in
 * scl-core: diffusion imputation and MAGIC (scl/kernel/impute.hpp), diffusion
 * pseudotime from diffusion components (scl/kernel/pseudotime.hpp) and the
 * generic Y = alpha A X + beta Y (scl/kernel/algebra.hpp).
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: loaded by tests/ as the checker.
 *
 * Every "s += a*b" / "y += a*x" of the reference is the fused form its
 * compiler contracts it to (an explicit fmaf); the build has
 * -ffp-contract=off, so nothing else is fused.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Dense n_cells x n_genes copy of a sparse expression matrix, as the two
 * imputations start (impute.hpp:612-640 and 683-697): zero, then every stored
 * entry in range, in stored order (a repeated index keeps the last value).
 * CSR (is_csr != 0): primary slot c is cell c; CSC: primary slot g is gene g.
 * Negative indices are skipped as well (handles never hold them). */
void orc_densify_f32(int64_t n_cells, int64_t n_genes, const int64_t* ip,
                     const int64_t* ix, const float* v, int is_csr, float* out) {
    memset(out, 0, (size_t)n_cells * (size_t)n_genes * sizeof(float));
    if (is_csr) {
        for (int64_t c = 0; c < n_cells; ++c)
            for (int64_t k = ip[c]; k < ip[c + 1]; ++k)
                if (ix[k] >= 0 && ix[k] < n_genes) out[c * n_genes + ix[k]] = v[k];
    } else {
        for (int64_t g = 0; g < n_genes; ++g)
            for (int64_t k = ip[g]; k < ip[g + 1]; ++k)
                if (ix[k] >= 0 && ix[k] < n_cells) out[ix[k] * n_genes + g] = v[k];
    }
}

/* diffusion_impute_sparse_transition (impute.hpp:599-658) after the densify:
 * n_steps of spmm_diffusion (impute.hpp:321-347: zero the row, then
 * axpy_simd of each stored neighbour in order, i.e. the FMA chain of
 * orc_spmm_step_f32), double-buffered.  X is n x G, updated in place. */
int orc_impute_diffusion_f32(int64_t n, int64_t G, int64_t n_steps, const int64_t* ip,
                             const int64_t* ix, const float* v, float* X) {
    if (n_steps <= 0 || n == 0 || G == 0) return 0;
    float* tmp = (float*)malloc((size_t)n * (size_t)G * sizeof(float));
    if (!tmp) return -1;
    float *in = X, *out = tmp;
    for (int64_t s = 0; s < n_steps; ++s) {
        orc_spmm_step_f32(n, G, ip, ix, v, in, out);
        float* w = in;
        in = out;
        out = w;
    }
    if (in != X) memcpy(X, in, (size_t)n * (size_t)G * sizeof(float));
    free(tmp);
    return 0;
}

/* magic_impute (impute.hpp:669-741) after the densify: FP32 row sums of the
 * affinity in stored order (plain adds), then t steps of spmm_normalized
 * (impute.hpp:355-389): a row whose sum is below 1e-12 copies its input row;
 * otherwise the row is zeroed and each neighbour added with weight
 * a_ij * (1 / sum_i), one FMA per element. */
int orc_magic_f32(int64_t n, int64_t G, int64_t t, const int64_t* ip, const int64_t* ix,
                  const float* v, float* X) {
    if (t <= 0 || n == 0 || G == 0) return 0;
    float* tmp = (float*)malloc((size_t)n * (size_t)G * sizeof(float));
    float* rs = (float*)malloc((size_t)n * sizeof(float));
    if (!tmp || !rs) {
        free(tmp);
        free(rs);
        return -1;
    }
    for (int64_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s += v[k];
        rs[i] = s;
    }
    float *in = X, *out = tmp;
    for (int64_t s = 0; s < t; ++s) {
#pragma omp parallel for schedule(static) if (n >= 256)
        for (int64_t i = 0; i < n; ++i) {
            float* y = out + i * G;
            if (rs[i] < 1e-12f) {
                memcpy(y, in + i * G, (size_t)G * sizeof(float));
                continue;
            }
            for (int64_t f = 0; f < G; ++f) y[f] = 0.0f;
            const float inv = 1.0f / rs[i];
            for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
                const float w = v[k] * inv;
                const float* x = in + ix[k] * G;
                for (int64_t f = 0; f < G; ++f) y[f] = fmaf(w, x[f], y[f]);
            }
        }
        float* w = in;
        in = out;
        out = w;
    }
    if (in != X) memcpy(X, in, (size_t)n * (size_t)G * sizeof(float));
    free(tmp);
    free(rs);
    return 0;
}

/* ---------------------------------------------------------------- FastRNG --
 * pseudotime.hpp:153-182.  Seeding: four rounds of z += 0x9e3779b97f4a7c15;
 * z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9, keeping the high 32 bits of each
 * z (the first half of a SplitMix64 output).  Step: the bit-mirrored form of
 * Marsaglia's xorshift128 (shifts 11 right, 8 left, 19 left); the newest
 * word is the output.  uniform() = next() * 2^-32 in float. */
typedef struct {
    uint32_t s[4];
} orc_rng;

static void rng_seed(orc_rng* r, uint64_t seed, uint64_t* z_out) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
        z += 0x9e3779b97f4a7c15ULL;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        if (z_out) z_out[i] = z;
        r->s[i] = (uint32_t)(z >> 32);
    }
}

static uint32_t rng_next(orc_rng* r) {
    uint32_t t = r->s[3];
    const uint32_t x = r->s[0];
    r->s[3] = r->s[2];
    r->s[2] = r->s[1];
    r->s[1] = x;
    t ^= t >> 11;
    t ^= t << 8;
    r->s[0] = t ^ x ^ (x << 19);
    return r->s[0];
}

static float rng_uniform(orc_rng* r) { return (float)rng_next(r) * 2.3283064365386963e-10f; }

void orc_fastrng_seed(uint64_t seed, uint64_t* z4, uint32_t* state4) {
    orc_rng r;
    rng_seed(&r, seed, z4);
    memcpy(state4, r.s, sizeof r.s);
}

void orc_fastrng_step(uint32_t* state4, int64_t m, uint32_t* out) {
    orc_rng r;
    memcpy(r.s, state4, sizeof r.s);
    for (int64_t i = 0; i < m; ++i) out[i] = rng_next(&r);
    memcpy(state4, r.s, sizeof r.s);
}

void orc_fastrng_uniform(uint64_t seed, int64_t m, float* out) {
    orc_rng r;
    rng_seed(&r, seed, NULL);
    for (int64_t i = 0; i < m; ++i) out[i] = rng_uniform(&r);
}

/* orthonormalize_columns (pseudotime.hpp:418-474): one Gram-Schmidt pass
 * over the K columns of row-major n x K Q (dot and norm of each earlier
 * column in one loop, projection when its norm exceeds 1e-15), then the
 * column is normalised, or, when its norm is at most 1e-15, replaced by the
 * normalised vector uniform() - 0.5 from FastRNG(c * 12345).  w: n floats. */
static void pt_orthonormalise(float* Q, int64_t n, int64_t K, float* w) {
    for (int64_t c = 0; c < K; ++c) {
        for (int64_t i = 0; i < n; ++i) w[i] = Q[i * K + c];
        for (int64_t p = 0; p < c; ++p) {
            float dot = 0.0f, nrm = 0.0f;
            for (int64_t i = 0; i < n; ++i) {
                const float q = Q[i * K + p];
                dot = fmaf(w[i], q, dot);
                nrm = fmaf(q, q, nrm);
            }
            if (nrm > 1e-15f) {
                const float coeff = dot / nrm;
                for (int64_t i = 0; i < n; ++i) w[i] = fmaf(-coeff, Q[i * K + p], w[i]);
            }
        }
        float nrm = 0.0f;
        for (int64_t i = 0; i < n; ++i) nrm = fmaf(w[i], w[i], nrm);
        if (nrm > 1e-15f) {
            const float inv = 1.0f / sqrtf(nrm);
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = w[i] * inv;
        } else {
            orc_rng r;
            rng_seed(&r, (uint64_t)c * 12345u, NULL);
            nrm = 0.0f;
            for (int64_t i = 0; i < n; ++i) {
                w[i] = rng_uniform(&r) - 0.5f;
                nrm = fmaf(w[i], w[i], nrm);
            }
            const float inv = 1.0f / sqrtf(nrm);
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = w[i] * inv;
        }
    }
}

int orc_pt_orthonormalise_f32(int64_t n, int64_t K, float* Q) {
    float* w = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    if (!w) return -1;
    pt_orthonormalise(Q, n, K, w);
    free(w);
    return 0;
}

/* diffusion_pseudotime (pseudotime.hpp:646-712): Q (n x K) filled with
 * FastRNG(42) uniform() - 0.5 in row-major order and orthonormalised;
 * n_iter times Q <- orthonormalise(P Q) (spmm_block, pseudotime.hpp:382-411:
 * the stored-order FMA chain); then pt_i = sqrt(sum_c (Q[i,c] - Q[root,c])^2)
 * and, when max pt > 0, pt *= 1 / max pt. */
int orc_pseudotime_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                       int64_t root, int64_t K, int64_t n_iter, float* pt) {
    if (n == 0) return 0;
    const size_t sz = (size_t)n * (size_t)(K > 0 ? K : 1);
    float* Q = (float*)malloc(sz * sizeof(float));
    float* Qn = (float*)malloc(sz * sizeof(float));
    float* w = (float*)malloc((size_t)n * sizeof(float));
    if (!Q || !Qn || !w) {
        free(Q);
        free(Qn);
        free(w);
        return -1;
    }
    orc_rng r;
    rng_seed(&r, 42, NULL);
    for (int64_t i = 0; i < n * K; ++i) Q[i] = rng_uniform(&r) - 0.5f;
    pt_orthonormalise(Q, n, K, w);
    for (int64_t it = 0; it < n_iter; ++it) {
        orc_spmm_step_f32(n, K, ip, ix, v, Q, Qn);
        pt_orthonormalise(Qn, n, K, w);
        float* s = Q;
        Q = Qn;
        Qn = s;
    }
    const float* rq = Q + root * K;
    for (int64_t i = 0; i < n; ++i) {
        float d2 = 0.0f;
        for (int64_t c = 0; c < K; ++c) {
            const float d = Q[i * K + c] - rq[c];
            d2 = fmaf(d, d, d2);
        }
        pt[i] = sqrtf(d2);
    }
    float mx = pt[0];
    for (int64_t i = 1; i < n; ++i) mx = pt[i] > mx ? pt[i] : mx;
    if (mx > 0.0f) {
        const float inv = 1.0f / mx;
        for (int64_t i = 0; i < n; ++i) pt[i] *= inv;
    }
    free(Q);
    free(Qn);
    free(w);
    return 0;
}

/* ------------------------------------------------------------ FP64 anchor --
 * orc_pseudotime_f32 in FP64 (double orthonormalisation and SpMM) from the
 * same fp32 FastRNG(42) start: the exact-arithmetic counterpart that FP32
 * implementations are measured against. */
static void pt_orthonormalise_f64(double* Q, int64_t n, int64_t K, double* w) {
    for (int64_t c = 0; c < K; ++c) {
        for (int64_t i = 0; i < n; ++i) w[i] = Q[i * K + c];
        for (int64_t p = 0; p < c; ++p) {
            double dot = 0.0, nrm = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                const double q = Q[i * K + p];
                dot += w[i] * q;
                nrm += q * q;
            }
            if (nrm > 1e-15) {
                const double coeff = dot / nrm;
                for (int64_t i = 0; i < n; ++i) w[i] -= coeff * Q[i * K + p];
            }
        }
        double nrm = 0.0;
        for (int64_t i = 0; i < n; ++i) nrm += w[i] * w[i];
        if (nrm > 1e-15) {
            const double inv = 1.0 / sqrt(nrm);
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = w[i] * inv;
        } else {
            orc_rng r;
            rng_seed(&r, (uint64_t)c * 12345u, NULL);
            nrm = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                w[i] = (double)(rng_uniform(&r) - 0.5f);
                nrm += w[i] * w[i];
            }
            const double inv = 1.0 / sqrt(nrm);
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = w[i] * inv;
        }
    }
}

int orc_pseudotime_f64(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                       int64_t root, int64_t K, int64_t n_iter, double* pt) {
    if (n == 0) return 0;
    const size_t sz = (size_t)n * (size_t)(K > 0 ? K : 1);
    double* Q = (double*)malloc(sz * sizeof(double));
    double* Qn = (double*)malloc(sz * sizeof(double));
    double* w = (double*)malloc((size_t)n * sizeof(double));
    if (!Q || !Qn || !w) {
        free(Q);
        free(Qn);
        free(w);
        return -1;
    }
    orc_rng r;
    rng_seed(&r, 42, NULL);
    for (int64_t i = 0; i < n * K; ++i) Q[i] = (double)(rng_uniform(&r) - 0.5f);
    pt_orthonormalise_f64(Q, n, K, w);
    for (int64_t it = 0; it < n_iter; ++it) {
#pragma omp parallel for schedule(static) if (n >= 256)
        for (int64_t i = 0; i < n; ++i) {
            double* y = Qn + i * K;
            for (int64_t c = 0; c < K; ++c) y[c] = 0.0;
            for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
                const double a = (double)v[k];
                for (int64_t c = 0; c < K; ++c) y[c] += a * Q[ix[k] * K + c];
            }
        }
        pt_orthonormalise_f64(Qn, n, K, w);
        double* s = Q;
        Q = Qn;
        Qn = s;
    }
    const double* rq = Q + root * K;
    for (int64_t i = 0; i < n; ++i) {
        double d2 = 0.0;
        for (int64_t c = 0; c < K; ++c) {
            const double d = Q[i * K + c] - rq[c];
            d2 += d * d;
        }
        pt[i] = sqrt(d2);
    }
    double mx = pt[0];
    for (int64_t i = 1; i < n; ++i) mx = pt[i] > mx ? pt[i] : mx;
    if (mx > 0.0)
        for (int64_t i = 0; i < n; ++i) pt[i] /= mx;
    free(Q);
    free(Qn);
    free(w);
    return 0;
}

/* algebra::spmm (algebra.hpp:546-609): Y (primary x F) = beta Y first
 * (scale_output 59-: beta 0 zeroes, beta 1 keeps, else multiplies); nothing
 * more when alpha == 0; else for each row and each stored entry in order,
 * w = a_ij * alpha and, unless w == 0 (axpy returns early, 110-111),
 * Y[i,:] = fma(w, X[j,:], Y[i,:]).  Column blocks of 64 do not change any
 * element's chain. */
void orc_algebra_spmm_f32(int64_t m, int64_t F, const int64_t* ip, const int64_t* ix,
                          const float* v, const float* X, float* Y, float alpha,
                          float beta) {
    for (int64_t i = 0; i < m * F; ++i) {
        if (beta == 0.0f) Y[i] = 0.0f;
        else if (beta != 1.0f) Y[i] = Y[i] * beta;
    }
    if (alpha == 0.0f) return;
#pragma omp parallel for schedule(static) if (m >= 128)
    for (int64_t i = 0; i < m; ++i) {
        float* y = Y + i * F;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
            const float w = v[k] * alpha;
            if (w == 0.0f) continue;
            const float* x = X + ix[k] * F;
            for (int64_t f = 0; f < F; ++f) y[f] = fmaf(w, x[f], y[f]);
        }
    }
}
