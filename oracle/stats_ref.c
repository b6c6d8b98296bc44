/* FP64 normalisation statistics and z Gram of CSR counts (test infrastructure only).
 *
 * The same arithmetic as the numpy oracle (oracle/norm.py), written in C with OpenMP so
 * that it finishes in seconds on million-cell matrices:
 *   libsize_i = sum_j c_ij over all features, 0 -> 1        (assay.py:818-823)
 *   x_ij = log1p(sf * c_ij / libsize_i)                      (assay.py:54-64)
 *   mu_j = sum_i x_ij / n, sigma_j = sqrt(sum_i (x_ij - mu_j)^2 / n), both cleaned as
 *   clean_array does (NaN/inf -> 0, then 0 -> 0 / 1)         (graph_datastore.py:767-796,
 *                                                             utils.py:143-154)
 *   z_ij = clip((x_ij - mu_j) / sigma_j, -clip, clip)        (ann.py:191-192; clip <= 0:
 *                                                             none)
 * and the uncentred Gram Z^T Z with the column sums of z, the matrix IncrementalPCA's
 * first partial_fit decomposes (sklearn _incremental_pca.py:257-385).
 *
 * Zeros of the count matrix all have z = z0_j = clip(-mu_j / sigma_j).  Writing each row
 * as z_i = z0 + delta_i, where delta_i is non-zero only at the row's stored entries,
 *   Z^T Z = n z0 z0^T + z0 s^T + s z0^T + sum_i delta_i delta_i^T,  s = sum_i delta_i,
 * which costs sum_i nnz_i^2 / 2 instead of n d^2 (one triangle, mirrored at the end).  Row sums run in row order per thread;
 * threads' partial sums are added in thread order.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline double clean(double v, double fill) {
  if (isnan(v) || isinf(v)) v = 0.0;
  return v == 0.0 ? fill : v;
}

static inline double xval(uint32_t c, double ls, double sf, int log_on) {
  const double v = (sf * (double)c) / ls;
  return log_on ? log1p(v) : v;
}

static inline double zval(double x, double mu, double sg, double clip) {
  double z = (x - mu) / sg;
  if (clip > 0) z = z < -clip ? -clip : (z > clip ? clip : z);
  return z;
}

void oracle_csr_stats(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                      int64_t n, int64_t d, double sf, int log_on, double *libsize,
                      double *mu, double *sigma) {
  int nt = omp_get_max_threads();
  double *acc = calloc((size_t)nt * d, sizeof(double));
  int64_t *cnt = calloc((size_t)nt * d, sizeof(int64_t));
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    double *a = acc + (size_t)t * d;
    int64_t *c = cnt + (size_t)t * d;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) s += (double)vals[e];
      libsize[i] = s == 0.0 ? 1.0 : s;
    }
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        a[cols[e]] += xval(vals[e], libsize[i], sf, log_on);
        c[cols[e]] += 1;
      }
  }
  for (int64_t j = 0; j < d; ++j) {
    double s = 0.0;
    int64_t k = 0;
    for (int t = 0; t < nt; ++t) {
      s += acc[(size_t)t * d + j];
      k += cnt[(size_t)t * d + j];
    }
    mu[j] = s / (double)n;
    sigma[j] = (double)(n - k) * mu[j] * mu[j]; /* the zeros' share of the second pass */
  }
  memset(acc, 0, (size_t)nt * d * sizeof(double));
#pragma omp parallel
  {
    double *a = acc + (size_t)omp_get_thread_num() * d;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        const double dv = xval(vals[e], libsize[i], sf, log_on) - mu[cols[e]];
        a[cols[e]] += dv * dv;
      }
  }
  for (int64_t j = 0; j < d; ++j) {
    double s = sigma[j];
    for (int t = 0; t < nt; ++t) s += acc[(size_t)t * d + j];
    sigma[j] = clean(sqrt(s / (double)n), 1.0);
    mu[j] = clean(mu[j], 0.0);
  }
  free(acc);
  free(cnt);
}

/* G (d x d, row-major, both triangles) = Z^T Z and colsum (d) = sum_i z_i over the rows
 * [0, n); libsize per row as above. */
void oracle_csr_zgram(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                      int64_t n, int64_t d, const double *libsize, double sf, int log_on,
                      const double *mu, const double *sigma, double clip, double *G,
                      double *colsum) {
  double *z0 = malloc((size_t)d * sizeof(double));
  for (int64_t j = 0; j < d; ++j) z0[j] = zval(0.0, mu[j], sigma[j], clip);
  int nt = omp_get_max_threads();
  double *gt = calloc((size_t)nt * d * d, sizeof(double));
  double *st = calloc((size_t)nt * d, sizeof(double));
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    double *g = gt + (size_t)t * d * d;
    double *s = st + (size_t)t * d;
    int64_t cap = 0;
    int32_t *jj = NULL;
    double *dl = NULL;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      const int64_t a = rowptr[i], m = rowptr[i + 1] - a;
      if (m > cap) {
        cap = m;
        jj = realloc(jj, (size_t)cap * sizeof(int32_t));
        dl = realloc(dl, (size_t)cap * sizeof(double));
      }
      for (int64_t e = 0; e < m; ++e) {
        const int32_t j = cols[a + e];
        jj[e] = j;
        dl[e] = zval(xval(vals[a + e], libsize[i], sf, log_on), mu[j], sigma[j], clip) - z0[j];
        s[j] += dl[e];
      }
      /* upper triangle: entry (min, max) of each pair; the lower one is filled at the end */
      for (int64_t p = 0; p < m; ++p) {
        const int32_t jp = jj[p];
        const double dp = dl[p];
        for (int64_t q = p; q < m; ++q) {
          const int32_t jq = jj[q];
          if (jp <= jq)
            g[(size_t)jp * d + jq] += dp * dl[q];
          else
            g[(size_t)jq * d + jp] += dp * dl[q];
        }
      }
    }
    free(jj);
    free(dl);
  }
  double *sv = calloc((size_t)d, sizeof(double));
  for (int t = 0; t < nt; ++t)
    for (int64_t j = 0; j < d; ++j) sv[j] += st[(size_t)t * d + j];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < d; ++r)
    for (int64_t c = r; c < d; ++c) {
      double v = 0.0;
      for (int t = 0; t < nt; ++t) v += gt[(size_t)t * d * d + (size_t)r * d + c];
      v += (double)n * z0[r] * z0[c] + z0[r] * sv[c] + sv[r] * z0[c];
      G[(size_t)r * d + c] = v;
      G[(size_t)c * d + r] = v;
    }
  for (int64_t j = 0; j < d; ++j) colsum[j] = (double)n * z0[j] + sv[j];
  free(sv);
  free(gt);
  free(st);
  free(z0);
}
