// MiniBatchKMeans seed partitions.  scarf's AnnStream._fit_kmeans (ann.py:328-346) fits
// scikit-learn 1.7.2's MiniBatchKMeans(n_clusters, random_state, batch_size, n_init=3)
// with one partial_fit per block of cells, then labels every cell with predict.  The
// random draws stay on the host, in scikit-learn's order; every distance, argmin and
// center update runs here, in FP64, on rows given as FP32 (promoted exactly) or FP64.
//
//   k-means++ seeding      sklearn/cluster/_kmeans.py:174-272 (_kmeans_plusplus)
//   nearest-center labels  sklearn/cluster/_k_means_lloyd.pyx:172-215 (_update_chunk_dense)
//   minibatch update       sklearn/cluster/_k_means_minibatch.pyx:59-108 (update_center_dense)
//
// Dot products are sequential fused sums over features from 0; a row's squared norm is its
// dot with itself, so the distance of a point to itself is exactly zero here, as in
// scikit-learn's candidate scoring.
#include "sh_common.h"

namespace sh {

template <typename T>
__device__ __forceinline__ double ld_f64(const T *p) {
  return (double)*p;
}

// ---- squared row norms (centers before each assignment) ----------------------------------
__global__ void k_km_row_norms(const double *__restrict__ C, int32_t k, int32_t d,
                               double *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  const double *r = C + (int64_t)c * d;
  double s = 0.0;
  for (int f = 0; f < d; ++f) s = fma(r[f], r[f], s);
  out[c] = s;
}

// ---- nearest center ------------------------------------------------------------------------
// A workgroup labels 64 rows.  Centers stream through LDS 64 at a time, features 32 at a
// time; each thread holds a 4 x 4 block of dot products (rows tr*4.., centers tc*4..).
// score = ||c||^2 - 2 x.c (the row norm does not change the argmin); the first center of
// the smallest score wins, as scikit-learn's strict '<' scan over increasing j.
constexpr int kAsRows = 64, kAsCenters = 64, kAsFeat = 32, kAsPad = 2;

template <typename T>
__global__ __launch_bounds__(256) void k_km_assign(const T *__restrict__ X, int64_t n,
                                                   int32_t d, int64_t ldx,
                                                   const double *__restrict__ C, int32_t k,
                                                   const double *__restrict__ cn,
                                                   int32_t *__restrict__ labels) {
  __shared__ __align__(16) double sX[kAsFeat][kAsRows + kAsPad];
  __shared__ __align__(16) double sC[kAsFeat][kAsCenters + kAsPad];
  __shared__ double sBv[16][kAsRows];
  __shared__ int32_t sBi[16][kAsRows];
  const int tid = threadIdx.x, tr = tid & 15, tc = tid >> 4;
  const int64_t r0 = (int64_t)blockIdx.x * kAsRows;
  double bv[4];
  int32_t bi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    bv[i] = __builtin_inf();
    bi[i] = INT32_MAX;
  }
  for (int c0 = 0; c0 < k; c0 += kAsCenters) {
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int f0 = 0; f0 < d; f0 += kAsFeat) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < (kAsRows * kAsFeat) / 256; ++q) {
        const int e = tid + 256 * q;
        const int p = e / kAsFeat, f = e % kAsFeat;
        const int ff = f0 + f;
        const int64_t row = r0 + p;
        sX[f][p] = (row < n && ff < d) ? ld_f64(X + row * ldx + ff) : 0.0;
        const int cc = c0 + p;
        sC[f][p] = (cc < k && ff < d) ? C[(int64_t)cc * d + ff] : 0.0;
      }
      __syncthreads();
      const int fe = min(kAsFeat, d - f0);
      for (int f = 0; f < fe; ++f) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = sX[f][tr * 4 + i];
          b[i] = sC[f][tc * 4 + i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int cc = c0 + tc * 4 + j;
      if (cc < k) {
        const double nc = cn[cc];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double v = fma(-2.0, acc[i][j], nc);
          if (v < bv[i]) {
            bv[i] = v;
            bi[i] = cc;
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    sBv[tc][tr * 4 + i] = bv[i];
    sBi[tc][tr * 4 + i] = bi[i];
  }
  __syncthreads();
  if (tid < kAsRows) {
    double v = sBv[0][tid];
    int32_t id = sBi[0][tid];
    for (int t = 1; t < 16; ++t) {
      const double v2 = sBv[t][tid];
      const int32_t i2 = sBi[t][tid];
      if (v2 < v || (v2 == v && i2 < id)) {
        v = v2;
        id = i2;
      }
    }
    const int64_t row = r0 + tid;
    if (row < n) labels[row] = id == INT32_MAX ? 0 : id;
  }
}

// ---- minibatch center update -------------------------------------------------------------
// One workgroup per center, one thread per feature.  Members are visited in increasing
// row order (a wave ballot over 64 labels at a time), so each (center, feature) is the
// same chain of roundings as scikit-learn's loop: c*w, then + x per member, then
// * (1 / (w + count)).  Contraction into fused multiply-adds is off for that reason.
template <typename T>
__global__ void k_km_update(const T *__restrict__ X, int64_t n, int32_t d, int64_t ldx,
                            const int32_t *__restrict__ labels, double *__restrict__ C,
                            double *__restrict__ ws, int32_t *__restrict__ counts) {
#pragma clang fp contract(off)
  const int c = blockIdx.x;
  const int f = threadIdx.x;
  const int lane = f & 63;
  const bool act = f < d;
  const double w0 = ws[c];
  double acc = act ? C[(int64_t)c * d + f] * w0 : 0.0;
  int cnt = 0;
  for (int64_t base = 0; base < n; base += 64) {
    const int64_t i = base + lane;
    const bool m = i < n && labels[i] == c;
    uint64_t mask = __ballot(m);
    cnt += __popcll(mask);
    if (act) {
      while (mask) {
        const int b = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        acc = acc + ld_f64(X + (base + b) * ldx + f);
      }
    }
  }
  __syncthreads();  // every wave has read ws[c] before it changes
  if (cnt > 0) {
    const double wn = w0 + (double)cnt;
    const double alpha = 1.0 / wn;
    if (act) C[(int64_t)c * d + f] = acc * alpha;
    if (f == 0) ws[c] = wn;
  }
  if (f == 0) counts[c] = cnt;
}

// ---- row gather: centers[dst[r]] = X[src[r]] -------------------------------------------------
template <typename T>
__global__ void k_km_gather(const T *__restrict__ X, int32_t d, int64_t ldx,
                            const int64_t *__restrict__ src, const int32_t *__restrict__ dst,
                            int64_t m, double *__restrict__ C) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * d) return;
  const int64_t r = t / d;
  const int f = (int)(t % d);
  const int64_t drow = dst ? (int64_t)dst[r] : r;
  C[drow * d + f] = ld_f64(X + src[r] * ldx + f);
}

// ---- k-means++ seeding -------------------------------------------------------------------
// One workgroup runs all n_clusters greedy rounds (they are sequential by construction).
// Thread j owns the contiguous points [j*S, (j+1)*S) and keeps their closest-center
// distances and per-candidate distances in the workspace; only candidate rows, prefix
// sums and potentials are shared through LDS.  Round c (c >= 1) turns the host's uniform
// draws u_t into thresholds u_t * potential and takes, for each, the first point whose
// running sum of closest distances reaches it (np.searchsorted over the cumulative sum;
// past the end -> the last point).  Every candidate is scored by the potential
// sum_i min(closest_i, dist(cand, i)); the first smallest wins.
constexpr int kPpThreads = 512, kPpWaves = kPpThreads / 64;
constexpr int kPpMaxTrials = 16, kPpMaxD = 256;

template <typename T>
__global__ __launch_bounds__(512) void k_km_pp(const T *__restrict__ X, int64_t ldx,
                                                const int64_t *__restrict__ rows, int64_t n,
                                                int32_t d, int32_t k, int32_t L,
                                                int64_t first, const double *__restrict__ U,
                                                int64_t *__restrict__ out,
                                                double *__restrict__ xn,
                                                double *__restrict__ closest,
                                                double *__restrict__ dmin) {
  __shared__ double sCand[kPpMaxTrials * kPpMaxD];
  __shared__ double sCn[kPpMaxTrials];
  __shared__ unsigned long long sCid[kPpMaxTrials];
  __shared__ double sIncl[kPpThreads];
  __shared__ double sTot[kPpWaves];
  __shared__ double sPart[kPpWaves][kPpMaxTrials];
  __shared__ double sPot;
  __shared__ int sBest;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t S = (n + kPpThreads - 1) / kPpThreads;
  const int64_t i0 = min(n, (int64_t)tid * S), i1 = min(n, i0 + S);

  for (int64_t i = i0; i < i1; ++i) {
    const T *xr = X + (rows ? rows[i] : i) * ldx;
    double s = 0.0;
    for (int f = 0; f < d; ++f) {
      const double v = ld_f64(xr + f);
      s = fma(v, v, s);
    }
    xn[i] = s;
    closest[i] = __builtin_inf();
  }
  if (tid == 0) sCid[0] = (unsigned long long)first;
  __syncthreads();

  for (int c = 0; c < k; ++c) {
    const int Le = c == 0 ? 1 : L;
    if (c > 0) {
      // thresholds of this round and the prefix sums they are searched in
      if (tid < L) sCid[tid] = (unsigned long long)(n - 1);
      double seg = 0.0;
      for (int64_t i = i0; i < i1; ++i) seg += closest[i];
      double v = seg;
      for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
      }
      if (lane == 63) sTot[wv] = v;
      __syncthreads();
      if (wv == 0) {
        double w = lane < kPpWaves ? sTot[lane] : 0.0;
        for (int o = 1; o < kPpWaves; o <<= 1) {
          const double t = __shfl_up(w, o, 64);
          if (lane >= o) w += t;
        }
        if (lane < kPpWaves) sTot[lane] = w;
      }
      __syncthreads();
      if (wv > 0) v += sTot[wv - 1];
      sIncl[tid] = v;
      __syncthreads();
      const double incl = v;
      const double excl = tid > 0 ? sIncl[tid - 1] : 0.0;
      const double pot = sPot;
      for (int t = 0; t < L; ++t) {
        const double r = U[(int64_t)(c - 1) * L + t] * pot;
        const bool hit = i1 > i0 && (tid == 0 ? r <= incl : (excl < r && r <= incl));
        if (hit) {
          double run = excl;
          int64_t id = i1 - 1;
          for (int64_t i = i0; i < i1; ++i) {
            run += closest[i];
            if (run >= r) {
              id = i;
              break;
            }
          }
          // rounding can make two neighbouring prefix sums straddle r; the first wins
          atomicMin(&sCid[t], (unsigned long long)id);
        }
      }
      __syncthreads();
    }
    // candidate rows to LDS
    for (int e = tid; e < Le * d; e += kPpThreads) {
      const int t = e / d, f = e % d;
      const int64_t ci = (int64_t)sCid[t];
      sCand[t * d + f] = ld_f64(X + (rows ? rows[ci] : ci) * ldx + f);
    }
    if (tid < Le) sCn[tid] = xn[sCid[tid]];
    __syncthreads();
    // distances of own points to every candidate, and the potentials they leave
    double part[kPpMaxTrials];
#pragma unroll
    for (int t = 0; t < kPpMaxTrials; ++t) part[t] = 0.0;
    for (int64_t i = i0; i < i1; ++i) {
      const T *xr = X + (rows ? rows[i] : i) * ldx;
      double acc[kPpMaxTrials];
#pragma unroll
      for (int t = 0; t < kPpMaxTrials; ++t) acc[t] = 0.0;
      for (int f = 0; f < d; ++f) {
        const double xv = ld_f64(xr + f);
#pragma unroll
        for (int t = 0; t < kPpMaxTrials; ++t)
          if (t < Le) acc[t] = fma(xv, sCand[t * d + f], acc[t]);
      }
      const double ci = closest[i], ni = xn[i];
#pragma unroll
      for (int t = 0; t < kPpMaxTrials; ++t) {
        if (t < Le) {
          double dist = (-2.0 * acc[t] + sCn[t]) + ni;
          dist = dist > 0.0 ? dist : 0.0;
          const double dm = ci < dist ? ci : dist;
          dmin[(int64_t)t * n + i] = dm;
          part[t] += dm;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < kPpMaxTrials; ++t) {
      if (t < Le) {
        const double s = wave_sum_f64(part[t]);
        if (lane == 0) sPart[wv][t] = s;
      }
    }
    __syncthreads();
    if (tid == 0) {
      int best = 0;
      double bp = 0.0;
      for (int t = 0; t < Le; ++t) {
        double p = 0.0;
        for (int w = 0; w < kPpWaves; ++w) p += sPart[w][t];
        if (t == 0 || p < bp) {
          bp = p;
          best = t;
        }
      }
      sBest = best;
      sPot = bp;
      out[c] = (int64_t)sCid[best];
    }
    __syncthreads();
    const int best = sBest;
    for (int64_t i = i0; i < i1; ++i) closest[i] = dmin[(int64_t)best * n + i];
  }
}

inline bool km_dtype_ok(int x_dtype) { return x_dtype == SH_F32 || x_dtype == SH_F64; }

}  // namespace sh

using namespace sh;

extern "C" size_t sh_kmeans_pp_workspace(int64_t n, int32_t n_trials) {
  if (n < 1 || n_trials < 1) return 0;
  return (size_t)(2 + n_trials) * (size_t)n * sizeof(double);
}

extern "C" int sh_kmeans_pp(const void *X, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                            const int64_t *rows, int32_t n_clusters, int32_t n_trials,
                            int64_t first, const double *uniforms, int64_t *centers_out,
                            void *ws, size_t ws_bytes, sh_stream_t stream) {
  SH_ARG(km_dtype_ok(x_dtype), "sh_kmeans_pp: x_dtype must be SH_F32 or SH_F64");
  SH_ARG(n >= 1 && d >= 1 && d <= kPpMaxD && ldx >= d,
         "sh_kmeans_pp: need n >= 1, 1 <= d <= %d, ldx >= d", kPpMaxD);
  SH_ARG(n_clusters >= 1 && n_clusters <= n, "sh_kmeans_pp: need 1 <= n_clusters <= n");
  SH_ARG(n_trials >= 1 && n_trials <= kPpMaxTrials, "sh_kmeans_pp: need 1 <= n_trials <= %d",
         kPpMaxTrials);
  SH_ARG(first >= 0 && first < n, "sh_kmeans_pp: first center out of range");
  SH_ARG(X && centers_out && ws && (uniforms || n_clusters == 1), "sh_kmeans_pp: NULL pointer");
  const size_t need = sh_kmeans_pp_workspace(n, n_trials);
  if (ws_bytes < need) {
    set_error("sh_kmeans_pp: workspace %zu bytes < %zu", ws_bytes, need);
    return SH_EWORKSPACE;
  }
  double *xn = (double *)ws;
  double *closest = xn + n;
  double *dmin = closest + n;
  if (x_dtype == SH_F32)
    hipLaunchKernelGGL(k_km_pp<float>, dim3(1), dim3(kPpThreads), 0, (hipStream_t)stream,
                       (const float *)X, ldx, rows, n, d, n_clusters, n_trials, first,
                       uniforms, centers_out, xn, closest, dmin);
  else
    hipLaunchKernelGGL(k_km_pp<double>, dim3(1), dim3(kPpThreads), 0, (hipStream_t)stream,
                       (const double *)X, ldx, rows, n, d, n_clusters, n_trials, first,
                       uniforms, centers_out, xn, closest, dmin);
  SH_LAUNCH_CHECK("k_km_pp");
  return SH_OK;
}

extern "C" int sh_kmeans_assign(const void *X, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                                const double *centers, int32_t n_clusters,
                                double *center_norms, int32_t *labels, sh_stream_t stream) {
  SH_ARG(km_dtype_ok(x_dtype), "sh_kmeans_assign: x_dtype must be SH_F32 or SH_F64");
  SH_ARG(n >= 0 && d >= 1 && ldx >= d && n_clusters >= 1,
         "sh_kmeans_assign: need n >= 0, d >= 1, ldx >= d, n_clusters >= 1");
  if (n == 0) return SH_OK;
  SH_ARG(X && centers && center_norms && labels, "sh_kmeans_assign: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_km_row_norms, dim3((uint32_t)cdiv(n_clusters, 256)), dim3(256), 0, s,
                     centers, n_clusters, d, center_norms);
  SH_LAUNCH_CHECK("k_km_row_norms");
  const dim3 grid((uint32_t)cdiv(n, kAsRows));
  if (x_dtype == SH_F32)
    hipLaunchKernelGGL(k_km_assign<float>, grid, dim3(256), 0, s, (const float *)X, n, d, ldx,
                       centers, n_clusters, center_norms, labels);
  else
    hipLaunchKernelGGL(k_km_assign<double>, grid, dim3(256), 0, s, (const double *)X, n, d,
                       ldx, centers, n_clusters, center_norms, labels);
  SH_LAUNCH_CHECK("k_km_assign");
  return SH_OK;
}

extern "C" int sh_kmeans_update(const void *X, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                                const int32_t *labels, double *centers, int32_t n_clusters,
                                double *weight_sums, int32_t *counts, sh_stream_t stream) {
  SH_ARG(km_dtype_ok(x_dtype), "sh_kmeans_update: x_dtype must be SH_F32 or SH_F64");
  SH_ARG(n >= 0 && d >= 1 && d <= 1024 && ldx >= d && n_clusters >= 1,
         "sh_kmeans_update: need n >= 0, 1 <= d <= 1024, ldx >= d, n_clusters >= 1");
  SH_ARG(X && labels && centers && weight_sums && counts, "sh_kmeans_update: NULL pointer");
  const dim3 block((uint32_t)(64 * cdiv(d, 64)));
  hipStream_t s = (hipStream_t)stream;
  if (x_dtype == SH_F32)
    hipLaunchKernelGGL(k_km_update<float>, dim3((uint32_t)n_clusters), block, 0, s,
                       (const float *)X, n, d, ldx, labels, centers, weight_sums, counts);
  else
    hipLaunchKernelGGL(k_km_update<double>, dim3((uint32_t)n_clusters), block, 0, s,
                       (const double *)X, n, d, ldx, labels, centers, weight_sums, counts);
  SH_LAUNCH_CHECK("k_km_update");
  return SH_OK;
}

extern "C" int sh_kmeans_gather(const void *X, int x_dtype, int32_t d, int64_t ldx,
                                const int64_t *src_rows, const int32_t *dst_rows, int64_t m,
                                double *centers, sh_stream_t stream) {
  SH_ARG(km_dtype_ok(x_dtype), "sh_kmeans_gather: x_dtype must be SH_F32 or SH_F64");
  SH_ARG(m >= 0 && d >= 1 && ldx >= d, "sh_kmeans_gather: need m >= 0, d >= 1, ldx >= d");
  if (m == 0) return SH_OK;
  SH_ARG(X && src_rows && centers, "sh_kmeans_gather: NULL pointer");
  const dim3 grid((uint32_t)cdiv(m * d, 256));
  if (x_dtype == SH_F32)
    hipLaunchKernelGGL(k_km_gather<float>, grid, dim3(256), 0, (hipStream_t)stream,
                       (const float *)X, d, ldx, src_rows, dst_rows, m, centers);
  else
    hipLaunchKernelGGL(k_km_gather<double>, grid, dim3(256), 0, (hipStream_t)stream,
                       (const double *)X, d, ldx, src_rows, dst_rows, m, centers);
  SH_LAUNCH_CHECK("k_km_gather");
  return SH_OK;
}
