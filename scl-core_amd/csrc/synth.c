/*
 * synth.c -- seeded synthetic inputs for the diffusion workloads.
 *
 * Everything here is deterministic for a given seed and independent of the
 * thread count: every random draw comes from a counter-based stream
 * (SplitMix64 over (seed, stream, counter)), never from a shared generator,
 * and never from std::*_distribution (whose output is implementation-defined).
 *
 * Graphs:
 *   synth_uniform_knn  k distinct neighbours per row drawn uniformly from
 *                      {0..n-1} \ {i} (Floyd sampling), indices sorted.
 *   synth_knn_graph    points from a Gaussian mixture in R^dim, exact kNN
 *                      (brute force), then union symmetrisation with the
 *                      larger weight kept: W_ij = max(w_ij, w_ji).
 * Edge weights follow w = exp(-d^2 / sigma^2), sigma = 1, d = |N(0,1)|.
 *
 * Dense inputs:
 *   synth_fill_normal         X_e ~ N(0,1)
 *   synth_fill_log1p_poisson  X_ig = log1p(Poisson(lambda_g)),
 *                             lambda_g ~ LogNormal(-1, 1.5)
 * Both accept an optional column list so a subset of the columns of a large
 * matrix can be regenerated bit-identically on its own.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SYN_EXPORT __attribute__((visibility("default")))

static inline uint64_t syn_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s; } syn_rng;

static inline syn_rng syn_stream(uint64_t seed, uint64_t stream) {
    syn_rng r;
    r.s = syn_mix(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL) ^
          syn_mix(stream + 0xD1B54A32D192ED03ULL);
    return r;
}

static inline uint64_t syn_next(syn_rng* r) {
    r->s += 0x9E3779B97F4A7C15ULL;
    return syn_mix(r->s);
}

/* uniform in (0, 1] -- never 0, so log() is safe */
static inline double syn_unif(syn_rng* r) {
    return ((double)(syn_next(r) >> 11) + 1.0) * 0x1.0p-53;
}

static inline double syn_normal(syn_rng* r) {
    const double u1 = syn_unif(r), u2 = syn_unif(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

/* uniform integer in [0, m) by rejection (no modulo bias) */
static inline uint64_t syn_below(syn_rng* r, uint64_t m) {
    const uint64_t lim = UINT64_MAX - (UINT64_MAX % m);
    uint64_t x;
    do { x = syn_next(r); } while (x >= lim);
    return x % m;
}

static inline float syn_edge_weight(syn_rng* r) {
    const double d = fabs(syn_normal(r));
    return (float)exp(-d * d);
}

static int syn_cmp_i64(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------ uniform kNN */

SYN_EXPORT int synth_uniform_knn(int64_t n, int64_t k, uint64_t seed,
                                 int64_t* indptr, int64_t* indices, float* w) {
    if (n < 2 || k < 1 || k > n - 1) return -1;
    for (int64_t i = 0; i <= n; ++i) indptr[i] = i * k;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        syn_rng r = syn_stream(seed, (uint64_t)i);
        int64_t* row = indices + i * k;
        /* Floyd's algorithm: k distinct values from [0, n-2] */
        const int64_t N = n - 1;
        int64_t cnt = 0;
        for (int64_t j = N - k; j < N; ++j) {
            const int64_t t = (int64_t)syn_below(&r, (uint64_t)(j + 1));
            int dup = 0;
            for (int64_t q = 0; q < cnt; ++q)
                if (row[q] == t) { dup = 1; break; }
            row[cnt++] = dup ? j : t;
        }
        for (int64_t q = 0; q < k; ++q)
            if (row[q] >= i) row[q] += 1; /* skip the self loop */
        qsort(row, (size_t)k, sizeof(int64_t), syn_cmp_i64);
        for (int64_t q = 0; q < k; ++q) w[i * k + q] = syn_edge_weight(&r);
    }
    return 0;
}

/* ------------------------------------------------------ geometric kNN graph */

#define SYN_QB 16    /* queries per block */
#define SYN_CB 512   /* candidates per block */

typedef struct { float d; int64_t j; } syn_cand;

static inline int syn_better(float d, int64_t j, const syn_cand* c) {
    return d < c->d || (d == c->d && j < c->j);
}

/* insert (d, j) into best[0..k) kept sorted ascending; caller checked it beats best[k-1] */
static inline void syn_insert(syn_cand* best, int64_t k, float d, int64_t j) {
    int64_t p = k - 1;
    while (p > 0 && syn_better(d, j, &best[p - 1])) {
        best[p] = best[p - 1];
        --p;
    }
    best[p].d = d;
    best[p].j = j;
}

/* exact kNN by brute force, squared Euclidean distance in float, ties broken
 * by the smaller index; soa is dim x n */
static void syn_knn(int64_t n, int64_t dim, int64_t k, const float* soa,
                    int64_t* nbr) {
    const int64_t nqb = (n + SYN_QB - 1) / SYN_QB;
#pragma omp parallel
    {
        syn_cand* best = (syn_cand*)malloc(sizeof(syn_cand) * (size_t)(SYN_QB * k));
        float* d2 = (float*)malloc(sizeof(float) * SYN_QB * SYN_CB);
        float q[64 * SYN_QB];
#pragma omp for schedule(dynamic, 4)
        for (int64_t qb = 0; qb < nqb; ++qb) {
            const int64_t q0 = qb * SYN_QB;
            const int64_t nq = (q0 + SYN_QB <= n) ? SYN_QB : n - q0;
            for (int64_t a = 0; a < nq; ++a) {
                for (int64_t d = 0; d < dim; ++d) q[d * SYN_QB + a] = soa[d * n + q0 + a];
                for (int64_t m = 0; m < k; ++m) {
                    best[a * k + m].d = INFINITY;
                    best[a * k + m].j = INT64_MAX;
                }
            }
            for (int64_t c0 = 0; c0 < n; c0 += SYN_CB) {
                const int64_t nc = (c0 + SYN_CB <= n) ? SYN_CB : n - c0;
                for (int64_t a = 0; a < nq; ++a) {
                    float* da = d2 + a * SYN_CB;
                    for (int64_t b = 0; b < nc; ++b) da[b] = 0.0f;
                    for (int64_t d = 0; d < dim; ++d) {
                        const float qv = q[d * SYN_QB + a];
                        const float* col = soa + d * n + c0;
                        for (int64_t b = 0; b < nc; ++b) {
                            const float t = col[b] - qv;
                            da[b] += t * t;
                        }
                    }
                    const float thr = best[a * k + k - 1].d;
                    const int64_t self = q0 + a;
                    for (int64_t b = 0; b < nc; ++b) {
                        if (da[b] <= thr) {
                            const int64_t j = c0 + b;
                            if (j == self) continue;
                            if (syn_better(da[b], j, &best[a * k + k - 1]))
                                syn_insert(best + a * k, k, da[b], j);
                        }
                    }
                }
            }
            for (int64_t a = 0; a < nq; ++a)
                for (int64_t m = 0; m < k; ++m) nbr[(q0 + a) * k + m] = best[a * k + m].j;
        }
        free(best);
        free(d2);
    }
}

typedef struct { int64_t j; float w; } syn_edge;

static int syn_cmp_edge(const void* a, const void* b) {
    const syn_edge* x = (const syn_edge*)a;
    const syn_edge* y = (const syn_edge*)b;
    return (x->j > y->j) - (x->j < y->j);
}

/* Gaussian-mixture points (cluster centres ~ N(0, spread^2) per coordinate,
 * points = centre + N(0,1)), exact kNN with k neighbours, C1 weight law per
 * directed edge, union symmetrisation keeping the larger weight.
 * indices / w must hold 2*n*k entries; returns nnz, or -1 on bad input. */
SYN_EXPORT int64_t synth_knn_graph(int64_t n, int64_t dim, int64_t k,
                                   int64_t n_clusters, double spread,
                                   uint64_t seed, int64_t* indptr,
                                   int64_t* indices, float* w) {
    if (n < 2 || k < 1 || k > n - 1 || dim < 1 || dim > 64 || n_clusters < 1)
        return -1;
    float* centres = (float*)malloc(sizeof(float) * (size_t)(n_clusters * dim));
    float* soa = (float*)malloc(sizeof(float) * (size_t)(n * dim));
    int64_t* nbr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n * k));
    int64_t* deg = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    if (!centres || !soa || !nbr || !deg) {
        free(centres); free(soa); free(nbr); free(deg);
        return -1;
    }
    {
        syn_rng r = syn_stream(seed, UINT64_C(0xC0FFEE));
        for (int64_t c = 0; c < n_clusters * dim; ++c)
            centres[c] = (float)(spread * syn_normal(&r));
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        syn_rng r = syn_stream(seed ^ UINT64_C(0x5EED), (uint64_t)i);
        const int64_t c = (int64_t)syn_below(&r, (uint64_t)n_clusters);
        for (int64_t d = 0; d < dim; ++d)
            soa[d * n + i] = centres[c * dim + d] + (float)syn_normal(&r);
    }
    syn_knn(n, dim, k, soa, nbr);
    free(soa);
    free(centres);

    /* directed weights, then both directions into per-row edge lists */
    float* dw = (float*)malloc(sizeof(float) * (size_t)(n * k));
    if (!dw) { free(nbr); free(deg); return -1; }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        syn_rng r = syn_stream(seed ^ UINT64_C(0xAB5EED), (uint64_t)i);
        for (int64_t m = 0; m < k; ++m) dw[i * k + m] = syn_edge_weight(&r);
    }
    for (int64_t e = 0; e < n * k; ++e) {
        deg[e / k + 1]++;
        deg[nbr[e] + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) deg[i + 1] += deg[i];
    syn_edge* lists = (syn_edge*)malloc(sizeof(syn_edge) * (size_t)(2 * n * k));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    if (!lists || !fill) {
        free(dw); free(nbr); free(deg); free(lists); free(fill);
        return -1;
    }
    memcpy(fill, deg, sizeof(int64_t) * (size_t)n);
    for (int64_t e = 0; e < n * k; ++e) {
        const int64_t i = e / k, j = nbr[e];
        syn_edge a = {j, dw[e]};
        syn_edge b = {i, dw[e]};
        lists[fill[i]++] = a;
        lists[fill[j]++] = b;
    }
    free(dw);
    free(nbr);
    free(fill);
    /* sort each row by column, merge duplicates with max, compact */
    int64_t* rowlen = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        syn_edge* L = lists + deg[i];
        const int64_t m = deg[i + 1] - deg[i];
        qsort(L, (size_t)m, sizeof(syn_edge), syn_cmp_edge);
        int64_t o = 0;
        for (int64_t q = 0; q < m; ++q) {
            if (o > 0 && L[o - 1].j == L[q].j) {
                if (L[q].w > L[o - 1].w) L[o - 1].w = L[q].w;
            } else {
                L[o++] = L[q];
            }
        }
        rowlen[i] = o;
    }
    indptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) indptr[i + 1] = indptr[i] + rowlen[i];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const syn_edge* L = lists + deg[i];
        for (int64_t q = 0; q < rowlen[i]; ++q) {
            indices[indptr[i] + q] = L[q].j;
            w[indptr[i] + q] = L[q].w;
        }
    }
    const int64_t nnz = indptr[n];
    free(lists);
    free(rowlen);
    free(deg);
    return nnz;
}

/* Row normalisation used to prepare P from W for the workloads: FP64 row
 * sums in stored order, p = (float)(w / d), rows with d <= 1e-15 set to 0. */
SYN_EXPORT void synth_row_normalize(int64_t n, const int64_t* indptr, float* w) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q) d += (double)w[q];
        for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q)
            w[q] = d > 1e-15 ? (float)((double)w[q] / d) : 0.0f;
    }
}

/* ----------------------------------------------------------- dense inputs */

/* X[i, cols[c]] for c < n_sub (cols == NULL: all F columns, n_sub = F);
 * out is n x n_sub row-major. */
SYN_EXPORT void synth_fill_normal(int64_t n, int64_t F, uint64_t seed,
                                  const int64_t* cols, int64_t n_sub,
                                  float* out) {
    if (!cols) n_sub = F;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t c = 0; c < n_sub; ++c) {
            const int64_t g = cols ? cols[c] : c;
            syn_rng r = syn_stream(seed, (uint64_t)(i * F + g));
            out[i * n_sub + c] = (float)syn_normal(&r);
        }
    }
}

static inline int64_t syn_poisson(syn_rng* r, double lam) {
    /* inversion by sequential search */
    const double u = syn_unif(r);
    double p = exp(-lam), s = p;
    int64_t x = 0;
    while (u > s && x < 100000) {
        ++x;
        p *= lam / (double)x;
        s += p;
        if (p == 0.0 && (double)x > lam) break;
    }
    return x;
}

SYN_EXPORT void synth_fill_log1p_poisson(int64_t n, int64_t F, uint64_t seed,
                                         const int64_t* cols, int64_t n_sub,
                                         float* out) {
    if (!cols) n_sub = F;
    double* lam = (double*)malloc(sizeof(double) * (size_t)(n_sub > 0 ? n_sub : 1));
    if (!lam) return;
    for (int64_t c = 0; c < n_sub; ++c) {
        const int64_t g = cols ? cols[c] : c;
        syn_rng r = syn_stream(seed ^ UINT64_C(0x1A4BDA), (uint64_t)g);
        lam[c] = exp(-1.0 + 1.5 * syn_normal(&r));
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t c = 0; c < n_sub; ++c) {
            const int64_t g = cols ? cols[c] : c;
            syn_rng r = syn_stream(seed, (uint64_t)(i * F + g));
            out[i * n_sub + c] = (float)log1p((double)syn_poisson(&r, lam[c]));
        }
    }
    free(lam);
}
