/*
 * neighbors.c -- CPU restatement of scl-core's exact k-nearest-neighbour
 * search (scl/kernel/neighbors.hpp:300-476, C ABI
 * scl/binding/c_api/neighbors.cpp:20-128).
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: loaded by tests/ as the checker.
 *
 * Norms follow vectorize::sum_squared (scl/core/vectorize.hpp:782-822) for
 * 8-lane float vectors (an AVX2 build): four accumulators of 8 lanes over
 * blocks of 32 values, folded as (a0 + a1) + (a2 + a3), then single vectors,
 * then the 8 lanes summed as ((l0 + l4) + (l2 + l6)) + ((l1 + l5) + (l3 + l7))
 * (the 256-bit halves added, then the 128-bit shuffle tree), then the tail.
 * The lane tree is the one Highway's x86 SumOfLanes uses; Highway is not in
 * the reference tree (fetched unpinned), so that order is parity unpinned.
 *
 * The sparse dot products (dot_linear / dot_binary / dot_gallop,
 * neighbors.hpp:42-178) all add the products of matching indices in
 * increasing index order, "sum += a * b" contracted to an FMA.  The search
 * visits j = 0..N-1, skips i, prunes with the Cauchy-Schwarz bound
 * (sqrt n_i - sqrt n_j)^2 >= current k-th, keeps a max-heap of k entries
 * (replace the top only when strictly closer), then heap-sorts ascending.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

float orc_sum_squared_f32(const float* v, int64_t n) {
    if (n == 0) return 0.0f;
    float acc[4][8];
    memset(acc, 0, sizeof acc);
    int64_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int r = 0; r < 4; ++r)
            for (int l = 0; l < 8; ++l) {
                const float x = v[i + 8 * r + l];
                acc[r][l] = fmaf(x, x, acc[r][l]);
            }
    float a[8];
    for (int l = 0; l < 8; ++l) a[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);
    for (; i + 8 <= n; i += 8)
        for (int l = 0; l < 8; ++l) a[l] = fmaf(v[i + l], v[i + l], a[l]);
    float r = ((a[0] + a[4]) + (a[2] + a[6])) + ((a[1] + a[5]) + (a[3] + a[7]));
    for (; i < n; ++i) r = fmaf(v[i], v[i], r);
    return r;
}

void orc_norms_f32(int64_t n, const int64_t* ip, const float* v, float* norms) {
    for (int64_t i = 0; i < n; ++i) norms[i] = orc_sum_squared_f32(v + ip[i], ip[i + 1] - ip[i]);
}

/* matches of two sorted index lists, products added in index order */
static float sparse_dot(const int64_t* i1, const float* v1, int64_t n1, const int64_t* i2,
                        const float* v2, int64_t n2) {
    float s = 0.0f;
    int64_t a = 0, b = 0;
    while (a < n1 && b < n2) {
        if (i1[a] == i2[b]) {
            s = fmaf(v1[a], v2[b], s);
            ++a;
            ++b;
        } else if (i1[a] < i2[b]) {
            ++a;
        } else {
            ++b;
        }
    }
    return s;
}

typedef struct {
    float d;
    int64_t j;
} kent;

static void swap_e(kent* a, kent* b) {
    kent t = *a;
    *a = *b;
    *b = t;
}

int orc_knn_f32(int64_t n, const int64_t* ip, const int64_t* ix, const float* v,
                const float* norms, int64_t k, int64_t* out_idx, float* out_dist) {
    int fail = 0;
#pragma omp parallel
    {
        kent* h = (kent*)malloc((size_t)(k > 0 ? k : 1) * sizeof(kent));
        if (!h) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < n; ++i) {
            if (!h) continue;
            int64_t cnt = 0;
            float maxd = 3.40282347e38f;
            const float ni = norms[i], si = sqrtf(ni);
            for (int64_t j = 0; j < n; ++j) {
                if (j == i) continue;
                const float nj = norms[j];
                const float df = si - sqrtf(nj);
                if (df * df >= maxd) continue;
                const float dot = sparse_dot(ix + ip[i], v + ip[i], ip[i + 1] - ip[i],
                                             ix + ip[j], v + ip[j], ip[j + 1] - ip[j]);
                float d2 = fmaf(-2.0f, dot, ni + nj);
                if (d2 < 0.0f) d2 = 0.0f;
                if (cnt < k) {
                    h[cnt].d = d2;
                    h[cnt].j = j;
                    int64_t p = cnt;
                    while (p > 0) { /* sift up (max-heap) */
                        const int64_t par = (p - 1) / 2;
                        if (h[p].d <= h[par].d) break;
                        swap_e(&h[p], &h[par]);
                        p = par;
                    }
                    ++cnt;
                    if (cnt == k) maxd = h[0].d;
                } else if (d2 < h[0].d) {
                    h[0].d = d2;
                    h[0].j = j;
                    int64_t p = 0;
                    for (;;) { /* sift down (max-heap) */
                        int64_t big = p;
                        const int64_t l = 2 * p + 1, r = 2 * p + 2;
                        if (l < cnt && h[l].d > h[big].d) big = l;
                        if (r < cnt && h[r].d > h[big].d) big = r;
                        if (big == p) break;
                        swap_e(&h[p], &h[big]);
                        p = big;
                    }
                    maxd = h[0].d;
                }
            }
            /* heap sort: min-heap build, extract to the back, reverse */
            for (int64_t m = cnt / 2; m > 0 && cnt >= 2; --m) {
                int64_t p = m - 1;
                for (;;) {
                    int64_t s = p;
                    const int64_t l = 2 * p + 1, r = 2 * p + 2;
                    if (l < cnt && h[l].d < h[s].d) s = l;
                    if (r < cnt && h[r].d < h[s].d) s = r;
                    if (s == p) break;
                    swap_e(&h[p], &h[s]);
                    p = s;
                }
            }
            for (int64_t m = cnt; m > 1; --m) {
                swap_e(&h[0], &h[m - 1]);
                int64_t p = 0;
                for (;;) {
                    int64_t s = p;
                    const int64_t l = 2 * p + 1, r = 2 * p + 2;
                    if (l < m - 1 && h[l].d < h[s].d) s = l;
                    if (r < m - 1 && h[r].d < h[s].d) s = r;
                    if (s == p) break;
                    swap_e(&h[p], &h[s]);
                    p = s;
                }
            }
            for (int64_t m = 0; m < cnt / 2; ++m) swap_e(&h[m], &h[cnt - 1 - m]);
            for (int64_t m = 0; m < cnt; ++m) {
                out_dist[i * k + m] = sqrtf(h[m].d);
                out_idx[i * k + m] = h[m].j;
            }
            for (int64_t m = cnt; m < k; ++m) {
                out_dist[i * k + m] = INFINITY;
                out_idx[i * k + m] = -1;
            }
        }
        free(h);
    }
    return fail ? -1 : 0;
}
