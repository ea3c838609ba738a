/*
 * walks.c -- CPU restatement of the vector-side consumers of the diffusion
 * operator in scl/kernel/diffusion.hpp: lazy random walk, heat kernel
 * signature, random walk with restart / personalized PageRank, diffusion
 * pseudotime (single and multi-root), diffusion-map embedding and diffusion
 * distance.
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: loaded by tests/ as the checker.
 *
 * Every "a*b + c*d" or "s += a*b" of the reference is written as the fused
 * form the reference's compiler may contract it to (an explicit fmaf); the
 * build has -ffp-contract=off, so nothing else is fused.  The SpMV is the
 * stored-order FMA chain of spmv_parallel (diffusion.hpp:169-221).  Sums over
 * all nodes (the convergence tests, the Gram-Schmidt dot products) run in
 * index order here; the device sums in another order, so the two agree to
 * rounding, not bit for bit.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define W_MIN_PROB 1e-15f

static void spmv(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                 const float* x, float* y) {
#pragma omp parallel for schedule(static) if (n >= 256)
    for (int64_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) s = fmaf(v[k], x[ix[k]], s);
        y[i] = s;
    }
}

/* lazy_random_walk (diffusion.hpp:1157-1194): x <- laziness*x + (1-laziness)*P x. */
int orc_lazy_walk_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                      float* x, int64_t n_steps, float laziness) {
    if (n == 0 || n_steps == 0) return 0;
    float* t = (float*)malloc((size_t)n * sizeof(float));
    if (!t) return -1;
    const float move = 1.0f - laziness;
    for (int64_t s = 0; s < n_steps; ++s) {
        spmv(n, ip, ix, v, x, t);
        for (int64_t i = 0; i < n; ++i) x[i] = fmaf(laziness, x[i], move * t[i]);
    }
    free(t);
    return 0;
}

/* heat_kernel_signature (diffusion.hpp:961-1003): sig = 1, then n_steps of
 * sig <- (1-dt) sig + dt P sig with dt = t / n_steps. */
int orc_heat_kernel_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                        float* sig, float t, int64_t n_steps) {
    if (n == 0) return 0;
    for (int64_t i = 0; i < n; ++i) sig[i] = 1.0f;
    float* tmp = (float*)malloc((size_t)n * sizeof(float));
    if (!tmp) return -1;
    const float dt = t / (float)n_steps, keep = 1.0f - dt;
    for (int64_t s = 0; s < n_steps; ++s) {
        spmv(n, ip, ix, v, sig, tmp);
        for (int64_t i = 0; i < n; ++i) sig[i] = fmaf(keep, sig[i], dt * tmp[i]);
    }
    free(tmp);
    return 0;
}

/* random_walk_with_restart (diffusion.hpp:833-895), with spmv_fused_linear
 * (285-331) and check_convergence_simd (337-364).  The swap of the output
 * pointer is kept: without convergence, the caller's buffer ends with the
 * last iterate when max_iter is even and with the one before when it is
 * odd.  Returns the iterations run (-1 on allocation failure). */
int64_t orc_rwr_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                    const int64_t* seeds, int64_t n_seeds, float* scores, float alpha,
                    int64_t max_iter, float tol) {
    if (n == 0) return 0;
    float* restart = (float*)calloc((size_t)n, sizeof(float));
    float* other = (float*)malloc((size_t)n * sizeof(float));
    if (!restart || !other) { free(restart); free(other); return -1; }
    int64_t cnt = 0;
    for (int64_t s = 0; s < n_seeds; ++s)
        if (seeds[s] >= 0 && seeds[s] < n) { restart[seeds[s]] = 1.0f; ++cnt; }
    if (cnt == 0) {
        memset(scores, 0, (size_t)n * sizeof(float));
        free(restart); free(other);
        return 0;
    }
    const float inv = 1.0f / (float)cnt;
    for (int64_t i = 0; i < n; ++i) restart[i] *= inv;
    memcpy(scores, restart, (size_t)n * sizeof(float));
    const float beta = 1.0f - alpha;
    float* cur = scores;
    float* nxt = other;
    int64_t it = 0;
    for (; it < max_iter; ++it) {
        spmv(n, ip, ix, v, cur, nxt);
        for (int64_t i = 0; i < n; ++i) nxt[i] = fmaf(alpha, nxt[i], beta * restart[i]);
        float total = 0.0f;
        for (int64_t i = 0; i < n; ++i) total += fabsf(nxt[i] - cur[i]);
        if (total < tol) {
            memcpy(cur, nxt, (size_t)n * sizeof(float));
            ++it;
            break;
        }
        float* tmp = cur; cur = nxt; nxt = tmp;
    }
    free(restart); free(other);
    return it;
}

/* compute_dpt (diffusion.hpp:670-768). */
int orc_dpt_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                int64_t root, float* pt, int64_t max_iter, float tol) {
    if (n == 0) return 0;
    float* p = (float*)calloc((size_t)n, sizeof(float));
    float* pn = (float*)calloc((size_t)n, sizeof(float));
    float* hit = (float*)calloc((size_t)n, sizeof(float));
    float* vis = (float*)calloc((size_t)n, sizeof(float));
    if (!p || !pn || !hit || !vis) { free(p); free(pn); free(hit); free(vis); return -1; }
    p[root] = 1.0f;
    vis[root] = 1.0f;
    for (int64_t it = 1; it <= max_iter; ++it) {
        spmv(n, ip, ix, v, p, pn);
        for (int64_t i = 0; i < n; ++i) {
            if (vis[i] < 0.99f && pn[i] > W_MIN_PROB) {
                const float rem = 1.0f - vis[i];
                const float np = pn[i] * rem;
                hit[i] = fmaf((float)it, np, hit[i]);
                vis[i] += np;
                vis[i] = vis[i] < 1.0f ? vis[i] : 1.0f;
            }
        }
        float mx = 0.0f;
        for (int64_t i = 0; i < n; ++i) {
            const float u = 1.0f - vis[i];
            mx = mx > u ? mx : u;
        }
        if (mx < tol) break;
        float* tmp = p; p = pn; pn = tmp;
    }
    float max_time = 0.0f;
    for (int64_t i = 0; i < n; ++i)
        if (vis[i] > 0.5f) {
            hit[i] /= vis[i];
            max_time = max_time > hit[i] ? max_time : hit[i];
        }
    const float inv_max = max_time > 0.0f ? 1.0f / max_time : 1.0f;
    for (int64_t i = 0; i < n; ++i) pt[i] = vis[i] > 0.5f ? hit[i] * inv_max : 1.0f;
    free(p); free(pn); free(hit); free(vis);
    return 0;
}

/* compute_dpt_multi_root (diffusion.hpp:774-827): the mean of the single-root
 * pseudotimes of the in-range roots (default tolerance), all ones when none. */
int orc_dpt_multi_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                      const int64_t* roots, int64_t n_roots, float* pt, int64_t max_iter) {
    if (n == 0 || n_roots == 0) return 0;
    float* one = (float*)malloc((size_t)n * sizeof(float));
    if (!one) return -1;
    int64_t valid = 0;
    memset(pt, 0, (size_t)n * sizeof(float));
    for (int64_t r = 0; r < n_roots; ++r) {
        if (roots[r] < 0 || roots[r] >= n) continue;
        ++valid;
        if (orc_dpt_f32(n, ip, ix, v, roots[r], one, max_iter, 1e-6f)) { free(one); return -1; }
        for (int64_t i = 0; i < n; ++i) pt[i] += one[i];
    }
    if (valid == 0) {
        for (int64_t i = 0; i < n; ++i) pt[i] = 1.0f;
    } else {
        const float inv = 1.0f / (float)valid;
        for (int64_t i = 0; i < n; ++i) pt[i] *= inv;
    }
    free(one);
    return 0;
}

/* orthogonalize_vectors (diffusion.hpp:370-446): modified Gram-Schmidt over
 * the K columns of the row-major n x K matrix Q, one reorthogonalisation
 * pass when some |<w, q_p>| / |w| reaches 0.7, degenerate columns replaced
 * by the unit vector e_(v mod n). */
static void mgs(float* Q, int64_t n, int64_t K, float* w) {
    for (int64_t c = 0; c < K; ++c) {
        for (int64_t i = 0; i < n; ++i) w[i] = Q[i * K + c];
        for (int reorth = 0; reorth <= 2; ++reorth) {
            for (int64_t p = 0; p < c; ++p) {
                float dot = 0.0f, nrm = 0.0f;
                for (int64_t i = 0; i < n; ++i) {
                    const float q = Q[i * K + p];
                    dot = fmaf(w[i], q, dot);
                    nrm = fmaf(q, q, nrm);
                }
                if (nrm > W_MIN_PROB) {
                    const float coeff = dot / nrm;
                    for (int64_t i = 0; i < n; ++i) w[i] = fmaf(-coeff, Q[i * K + p], w[i]);
                }
            }
            float nv = 0.0f;
            for (int64_t i = 0; i < n; ++i) nv = fmaf(w[i], w[i], nv);
            if (reorth == 0 && c > 0) {
                float max_cos = 0.0f;
                for (int64_t p = 0; p < c; ++p) {
                    float dot = 0.0f;
                    for (int64_t i = 0; i < n; ++i) dot = fmaf(w[i], Q[i * K + p], dot);
                    const float cs = nv > W_MIN_PROB ? fabsf(dot) / sqrtf(nv) : 0.0f;
                    max_cos = max_cos > cs ? max_cos : cs;
                }
                if (max_cos < 0.7f) break;
            } else {
                break;
            }
        }
        float nrm = 0.0f;
        for (int64_t i = 0; i < n; ++i) nrm = fmaf(w[i], w[i], nrm);
        if (nrm > W_MIN_PROB) {
            const float inv = 1.0f / sqrtf(nrm);
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = w[i] * inv;
        } else {
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = (i == c % n) ? 1.0f : 0.0f;
        }
    }
}

/* Initial Q of diffusion_map_embedding (diffusion.hpp:922-929). */
void orc_dmap_init_f32(int64_t n, int64_t K, float* Q) {
    for (int64_t i = 0; i < n; ++i)
        for (int64_t c = 0; c < K; ++c) {
            const float t = (float)i / (float)n;
            const float ph = (float)(c + 1) * 3.14159265358979323846f;
            Q[i * K + c] = cosf(ph * t) + 0.1f * sinf(ph * t * 2.7f);
        }
}

/* diffusion_map_embedding (diffusion.hpp:901-955): n_iter rounds of
 * Q <- MGS(P Q) after an initial MGS of the fixed start. */
int orc_dmap_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                 float* emb, int64_t K, int64_t n_iter) {
    if (n == 0 || K == 0) return 0;
    float* Q = (float*)malloc((size_t)(n * K) * sizeof(float));
    float* Qn = (float*)malloc((size_t)(n * K) * sizeof(float));
    float* w = (float*)malloc((size_t)n * sizeof(float));
    if (!Q || !Qn || !w) { free(Q); free(Qn); free(w); return -1; }
    orc_dmap_init_f32(n, K, Q);
    mgs(Q, n, K, w);
    for (int64_t it = 0; it < n_iter; ++it) {
        orc_spmm_step_f32(n, K, ip, ix, v, Q, Qn);
        mgs(Qn, n, K, w);
        float* tmp = Q; Q = Qn; Qn = tmp;
    }
    memcpy(emb, Q, (size_t)(n * K) * sizeof(float));
    free(Q); free(Qn); free(w);
    return 0;
}

/* Gram-Schmidt alone on a given n x K matrix (for pinning tests). */
int orc_mgs_f32(int64_t n, int64_t K, float* Q) {
    float* w = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    if (!w) return -1;
    mgs(Q, n, K, w);
    free(w);
    return 0;
}

/* diffusion_distance (diffusion.hpp:1024-1116).  n <= 1000: rows of P^t
 * (t SpMM steps from the identity); n > 1000: the vectors P^t e_i, i.e. the
 * rows of (P^T)^t.  D[i][j] = sqrt(sum_k (a_ik - a_jk)^2), summed in k order. */
int orc_diffusion_distance_f32(int64_t n, const int64_t* ip, const int64_t* ix,
                               const float* v, float* D, int64_t n_steps) {
    if (n == 0) return 0;
    const size_t nn = (size_t)n * (size_t)n;
    float* A = (float*)calloc(nn, sizeof(float));
    float* B = (float*)malloc(nn * sizeof(float));
    if (!A || !B) { free(A); free(B); return -1; }
    if (n <= 1000) {
        for (int64_t i = 0; i < n; ++i) A[i * n + i] = 1.0f;
        for (int64_t s = 0; s < n_steps; ++s) {
            orc_spmm_step_f32(n, n, ip, ix, v, A, B);
            float* tmp = A; A = B; B = tmp;
        }
    } else {
        float* e = (float*)malloc((size_t)n * sizeof(float));
        float* tmp = (float*)malloc((size_t)n * sizeof(float));
        if (!e || !tmp) { free(e); free(tmp); free(A); free(B); return -1; }
        for (int64_t i = 0; i < n; ++i) {
            memset(e, 0, (size_t)n * sizeof(float));
            e[i] = 1.0f;
            for (int64_t s = 0; s < n_steps; ++s) {
                spmv(n, ip, ix, v, e, tmp);
                memcpy(e, tmp, (size_t)n * sizeof(float));
            }
            memcpy(A + (size_t)i * n, e, (size_t)n * sizeof(float));
        }
        free(e); free(tmp);
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
        D[i * n + i] = 0.0f;
        for (int64_t j = i + 1; j < n; ++j) {
            float s = 0.0f;
            for (int64_t k = 0; k < n; ++k) {
                const float d = A[i * n + k] - A[j * n + k];
                s = fmaf(d, d, s);
            }
            const float dist = sqrtf(s);
            D[i * n + j] = dist;
            D[j * n + i] = dist;
        }
    }
    free(A); free(B);
    return 0;
}

/* ------------------------------------------------------------ FP64 anchors --
 * The diffusion map of orc_dmap_f32 in FP64 (double accumulators, double
 * SpMM), from the same fp32 start vectors: the exact-arithmetic counterpart
 * that FP32 implementations are measured against. */
static void spmm_step_f64(int64_t n, int64_t F, const int64_t* ip, const int64_t* ix,
                          const float* v, const double* X, double* Y) {
#pragma omp parallel for schedule(static) if (n >= 256)
    for (int64_t i = 0; i < n; ++i) {
        double* y = Y + i * F;
        for (int64_t c = 0; c < F; ++c) y[c] = 0.0;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
            const double a = (double)v[k];
            const double* x = X + ix[k] * F;
            for (int64_t c = 0; c < F; ++c) y[c] += a * x[c];
        }
    }
}

static void mgs_f64(double* Q, int64_t n, int64_t K, double* w) {
    for (int64_t c = 0; c < K; ++c) {
        for (int64_t i = 0; i < n; ++i) w[i] = Q[i * K + c];
        for (int reorth = 0; reorth <= 2; ++reorth) {
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
            double nv = 0.0;
            for (int64_t i = 0; i < n; ++i) nv += w[i] * w[i];
            if (reorth == 0 && c > 0) {
                double max_cos = 0.0;
                for (int64_t p = 0; p < c; ++p) {
                    double dot = 0.0;
                    for (int64_t i = 0; i < n; ++i) dot += w[i] * Q[i * K + p];
                    const double cs = nv > 1e-15 ? fabs(dot) / sqrt(nv) : 0.0;
                    max_cos = max_cos > cs ? max_cos : cs;
                }
                if (max_cos < 0.7) break;
            } else {
                break;
            }
        }
        double nrm = 0.0;
        for (int64_t i = 0; i < n; ++i) nrm += w[i] * w[i];
        if (nrm > 1e-15) {
            const double inv = 1.0 / sqrt(nrm);
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = w[i] * inv;
        } else {
            for (int64_t i = 0; i < n; ++i) Q[i * K + c] = (i == c % n) ? 1.0 : 0.0;
        }
    }
}

int orc_dmap_f64(int64_t n, const int64_t* ip, const int64_t* ix, const float* v, double* emb,
                 int64_t K, int64_t n_iter) {
    if (n == 0 || K == 0) return 0;
    float* q32 = (float*)malloc((size_t)(n * K) * sizeof(float));
    double* Q = (double*)malloc((size_t)(n * K) * sizeof(double));
    double* Qn = (double*)malloc((size_t)(n * K) * sizeof(double));
    double* w = (double*)malloc((size_t)n * sizeof(double));
    if (!q32 || !Q || !Qn || !w) { free(q32); free(Q); free(Qn); free(w); return -1; }
    orc_dmap_init_f32(n, K, q32);
    for (int64_t e = 0; e < n * K; ++e) Q[e] = (double)q32[e];
    mgs_f64(Q, n, K, w);
    for (int64_t it = 0; it < n_iter; ++it) {
        spmm_step_f64(n, K, ip, ix, v, Q, Qn);
        mgs_f64(Qn, n, K, w);
        double* tmp = Q; Q = Qn; Qn = tmp;
    }
    memcpy(emb, Q, (size_t)(n * K) * sizeof(double));
    free(q32); free(Q); free(Qn); free(w);
    return 0;
}
