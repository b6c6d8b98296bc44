/* FP64 brute-force k-nearest-neighbour oracle (test infrastructure only).
 *
 * Distance: squared L2, the metric of scarf's hnswlib index (space 'l2',
 * ann.py:14-28), evaluated directly as sum_t (q_t - r_t)^2 in double precision,
 * terms added in feature order t = 0..D-1, with no fused multiply-add (built with
 * -ffp-contract=off).  Ranking: ascending distance, ties broken by the lower
 * reference index.  The HIP re-rank kernel evaluates distances in this same order,
 * so the two agree bit for bit.
 */
#include <stdint.h>
#include <stdlib.h>
#include <math.h>

static inline double sqdist(const float *q, const float *r, int D) {
    double acc = 0.0;
    for (int t = 0; t < D; ++t) {
        double df = (double)q[t] - (double)r[t];
        double sq = df * df;
        acc = acc + sq;
    }
    return acc;
}

/* (d, j) lexicographic: 1 if (d1, j1) < (d2, j2). */
static inline int better(double d1, int64_t j1, double d2, int64_t j2) {
    return d1 < d2 || (d1 == d2 && j1 < j2);
}

/* Q: nq x ldq floats, R: nr x ldr floats.  Outputs idx[nq*k] (-1 padded when nr < k)
 * and dist[nq*k] ascending. */
void oracle_knn_f64(const float *Q, int64_t nq, int64_t ldq, const float *R, int64_t nr,
                    int64_t ldr, int D, int k, int64_t *idx, double *dist) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nq; ++i) {
        int64_t *bi = idx + i * k;
        double *bd = dist + i * k;
        int cnt = 0;
        for (int c = 0; c < k; ++c) { bi[c] = -1; bd[c] = INFINITY; }
        const float *q = Q + i * ldq;
        for (int64_t j = 0; j < nr; ++j) {
            double d = sqdist(q, R + j * ldr, D);
            if (cnt == k && !better(d, j, bd[k - 1], bi[k - 1])) continue;
            int pos = cnt < k ? cnt : k - 1;
            if (cnt < k) ++cnt;
            while (pos > 0 && better(d, j, bd[pos - 1], bi[pos - 1])) {
                bd[pos] = bd[pos - 1]; bi[pos] = bi[pos - 1]; --pos;
            }
            bd[pos] = d; bi[pos] = j;
        }
    }
}

/* Exact distances for given (query, reference) pairs, same evaluation order. */
void oracle_pair_dist_f64(const float *Q, int64_t ldq, const float *R, int64_t ldr, int D,
                          const int64_t *qi, const int64_t *rj, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) out[p] = sqdist(Q + qi[p] * ldq, R + rj[p] * ldr, D);
}
