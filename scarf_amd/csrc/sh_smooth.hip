// Fuzzy kNN edge weights: umap-learn's smooth_knn_dist + compute_membership_strengths
// as scarf applies them per row chunk (knn_utils.py:89-159), in float32 like the
// numba originals.
//
// Per chunk of `chunk_size` rows:
//   mean_c   = mean of all the chunk's distances (floor for rows with rho == 0)
//   rho_i    = local_connectivity-th non-zero distance (interpolated)
//   sigma_i  = bisection (<= n_iter steps) of sum_{j>=1} f(d_ij - rho_i) = log2(k) * bw,
//              f(x) = exp(-x / sigma) for x > 0 else 1, stop when |psum - target| < 1e-5;
//              floored at 1e-3 * mean(row) (rho > 0) or 1e-3 * mean_c
//   w_ij     = 0 if idx_ij equals the row's position inside the chunk, 1 if
//              d - rho <= 0 or sigma == 0, else exp(-(d - rho) / sigma)
// Then every zero weight becomes global_min: 1, lowered by each chunk that holds a
// zero weight to that chunk's smallest non-zero weight (knn_utils.py:144-158).
#include "sh_common.h"

namespace sh {

constexpr float kSmoothTol = 1e-5f;
constexpr float kMinKDistScale = 1e-3f;

// Deterministic FP64 mean of each chunk's n_c * k distances.  A chunk's entries are cut
// into mean_parts(chunk_size) contiguous parts, one workgroup each (20 chunks at 1M rows
// would otherwise leave the GPU nearly idle); each part sums in a fixed tree order and the
// parts are added in order, so the result depends only on the inputs.
constexpr int kMeanParts = 64;
inline int64_t mean_parts(int64_t cs) { return std::min<int64_t>(kMeanParts, cdiv(cs, 128)); }

__global__ void __launch_bounds__(256) k_chunk_mean_part(const float *__restrict__ dist,
                                                         int64_t n, int32_t k, int64_t cs,
                                                         int32_t parts,
                                                         double *__restrict__ part_out) {
  __shared__ double red[256];
  const int64_t c = blockIdx.x;
  const int64_t a = c * cs * k, b = min(n, (c + 1) * cs) * (int64_t)k;
  const int64_t len = (b - a + parts - 1) / parts;
  const int64_t pa = min(b, a + (int64_t)blockIdx.y * len), pb = min(b, pa + len);
  double s = 0.0;
  for (int64_t t = pa + threadIdx.x; t < pb; t += 256) s += (double)dist[t];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part_out[c * parts + blockIdx.y] = red[0];
}

__global__ void __launch_bounds__(256) k_chunk_mean_sum(const double *__restrict__ part,
                                                        int64_t n, int32_t k, int64_t cs,
                                                        int32_t parts, int64_t nc,
                                                        double *__restrict__ mean_out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int64_t cnt = (min(n, (c + 1) * cs) - c * cs) * (int64_t)k;
  double s = 0.0;
  for (int p = 0; p < parts; ++p) s += part[c * parts + p];
  mean_out[c] = cnt > 0 ? s / (double)cnt : 0.0;
}

template <int KM>
__global__ void __launch_bounds__(256) k_smooth_rows(
    const float *__restrict__ dist, const int64_t *__restrict__ idx, int64_t n, int32_t k,
    int64_t cs, float lc, float bw, int32_t n_iter, const double *__restrict__ chunk_mean,
    float *__restrict__ sigma_out, float *__restrict__ rho_out, float *__restrict__ w_out,
    uint32_t *__restrict__ chunk_minw, int32_t *__restrict__ chunk_zero) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c = i / cs;
  const int64_t local = i - c * cs;
  float d[KM];
  double rs = 0.0;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    d[j] = j < k ? dist[i * k + j] : 0.f;
    if (j < k) rs += (double)d[j];
  }
  const float row_mean = (float)(rs / (double)k);
  const float mean_c = (float)chunk_mean[c];
  // rho: local_connectivity-th non-zero distance
  int nnz = 0;
#pragma unroll
  for (int j = 0; j < KM; ++j)
    if (j < k && d[j] > 0.f) ++nnz;
  float rho = 0.f;
  if ((float)nnz >= lc) {
    const int index = (int)floorf(lc);
    const float interp = lc - (float)index;
    float nz_prev = 0.f, nz_at = 0.f;  // non-zero distances index-1 and index
    int seen = 0;
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      if (j < k && d[j] > 0.f) {
        if (seen == index - 1) nz_prev = d[j];
        if (seen == index) nz_at = d[j];
        ++seen;
      }
    }
    if (index > 0) {
      rho = nz_prev;
      if (interp > kSmoothTol) rho += interp * (nz_at - nz_prev);
    } else {
      float first = 0.f;
      bool got = false;
#pragma unroll
      for (int j = 0; j < KM; ++j)
        if (!got && j < k && d[j] > 0.f) {
          first = d[j];
          got = true;
        }
      rho = interp * first;
    }
  } else if (nnz > 0) {
    float mx = 0.f;
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (j < k && d[j] > 0.f) mx = fmaxf(mx, d[j]);
    rho = mx;
  }
  const double target = log2((double)k) * (double)bw;
  float lo = 0.f, hi = INFINITY, mid = 1.f;
  for (int it = 0; it < n_iter; ++it) {
    float psum = 0.f;
#pragma unroll
    for (int j = 1; j < KM; ++j) {
      if (j < k) {
        const float dd = d[j] - rho;
        psum = dd > 0.f ? psum + expf(-(dd / mid)) : psum + 1.f;
      }
    }
    if (fabs((double)psum - target) < (double)kSmoothTol) break;
    if ((double)psum > target) {
      hi = mid;
      mid = (lo + hi) / 2.f;
    } else {
      lo = mid;
      if (hi == INFINITY)
        mid = mid * 2.f;
      else
        mid = (lo + hi) / 2.f;
    }
  }
  float sigma = mid;
  if (rho > 0.f) {
    if (sigma < kMinKDistScale * row_mean) sigma = kMinKDistScale * row_mean;
  } else {
    if (sigma < kMinKDistScale * mean_c) sigma = kMinKDistScale * mean_c;
  }
  sigma_out[i] = sigma;
  rho_out[i] = rho;
  uint32_t minw = 0xffffffffu;
  int zero = 0;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    if (j < k) {
      float w;
      const int64_t nb = idx[i * k + j];
      if (nb == local)
        w = 0.f;
      else if (d[j] - rho <= 0.f || sigma == 0.f)
        w = 1.f;
      else
        w = expf(-((d[j] - rho) / sigma));
      w_out[i * k + j] = w;
      if (w == 0.f)
        zero = 1;
      else
        minw = min(minw, float_ord(w));
    }
  }
  // one atomic per wave when all 64 rows are present and share a chunk (min and or do not
  // depend on order); lanes past n returned above, so a partial wave takes the per-row path
  const bool full = __ballot(1) == ~0ull;
  const int64_t c_first = __shfl(c, 0, 64);
  if (full && __ballot(c != c_first) == 0) {
    const bool any_zero = __ballot(zero) != 0;
    uint32_t m = minw;
    for (int o = 32; o >= 1; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) {
      if (any_zero) atomicOr(&chunk_zero[c], 1);
      if (m != 0xffffffffu) atomicMin(&chunk_minw[c], m);
    }
  } else {
    if (zero) atomicOr(&chunk_zero[c], 1);
    if (minw != 0xffffffffu) atomicMin(&chunk_minw[c], minw);
  }
}

__global__ void k_global_min(const uint32_t *__restrict__ chunk_minw,
                             const int32_t *__restrict__ chunk_zero, int64_t n_chunks,
                             float *__restrict__ gmin) {
  // one thread: chunks are few (n / chunk_size) and the scan is in chunk order
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float g = 1.f;
  for (int64_t c = 0; c < n_chunks; ++c)
    if (chunk_zero[c] && chunk_minw[c] != 0xffffffffu) {
      const float m = ord_float(chunk_minw[c]);
      if (m < g) g = m;
    }
  *gmin = g;
}

__global__ void k_fix_zero(float *__restrict__ w, int64_t m, const float *__restrict__ gmin) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m && w[t] == 0.f) w[t] = *gmin;
}

__global__ void k_smooth_init(uint32_t *__restrict__ chunk_minw, int32_t *__restrict__ chunk_zero,
                              int64_t n_chunks) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_chunks) {
    chunk_minw[t] = 0xffffffffu;
    chunk_zero[t] = 0;
  }
}

}  // namespace sh

using namespace sh;

extern "C" size_t sh_smooth_knn_workspace(int64_t n, int64_t chunk_size) {
  if (chunk_size <= 0) return 0;
  const int64_t nc = cdiv(n, chunk_size);
  return align_up(nc * sizeof(double), 256) + align_up(nc * sizeof(uint32_t), 256) +
         align_up(nc * sizeof(int32_t), 256) + 256 +
         align_up(nc * mean_parts(chunk_size) * sizeof(double), 256);
}

extern "C" int sh_smooth_knn(const float *dist, const int64_t *idx, int64_t n, int32_t k,
                             int64_t chunk_size, float local_connectivity, float bandwidth,
                             int32_t n_iter, float *sigma_out, float *rho_out, float *w_out,
                             void *workspace, size_t workspace_bytes, sh_stream_t stream) {
  SH_ARG(n >= 0 && k >= 1 && k <= 64, "sh_smooth_knn: k=%d must be in [1, 64]", k);
  SH_ARG(chunk_size >= 1, "sh_smooth_knn: chunk_size must be >= 1");
  SH_ARG(dist && idx && sigma_out && rho_out && w_out, "sh_smooth_knn: NULL pointer");
  if (n == 0) return SH_OK;
  const size_t need = sh_smooth_knn_workspace(n, chunk_size);
  if (!workspace || workspace_bytes < need) {
    set_error("sh_smooth_knn: workspace %zu < %zu bytes", workspace_bytes, need);
    return SH_EWORKSPACE;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t nc = cdiv(n, chunk_size);
  char *p = (char *)workspace;
  double *cmean = (double *)p;
  p += align_up(nc * sizeof(double), 256);
  uint32_t *cmin = (uint32_t *)p;
  p += align_up(nc * sizeof(uint32_t), 256);
  int32_t *czero = (int32_t *)p;
  p += align_up(nc * sizeof(int32_t), 256);
  float *gmin = (float *)p;
  p += 256;
  double *cpart = (double *)p;
  hipLaunchKernelGGL(k_smooth_init, dim3(cdiv(nc, 256)), dim3(256), 0, st, cmin, czero, nc);
  SH_LAUNCH_CHECK("k_smooth_init");
  const int32_t parts = (int32_t)mean_parts(chunk_size);
  hipLaunchKernelGGL(k_chunk_mean_part, dim3(nc, parts), dim3(256), 0, st, dist, n, k,
                     chunk_size, parts, cpart);
  SH_LAUNCH_CHECK("k_chunk_mean_part");
  hipLaunchKernelGGL(k_chunk_mean_sum, dim3(cdiv(nc, 256)), dim3(256), 0, st, cpart, n, k,
                     chunk_size, parts, nc, cmean);
  SH_LAUNCH_CHECK("k_chunk_mean_sum");
  dim3 grid(cdiv(n, 256)), block(256);
#define SH_SMOOTH(KM)                                                                           \
  hipLaunchKernelGGL(k_smooth_rows<KM>, grid, block, 0, st, dist, idx, n, k, chunk_size,        \
                     local_connectivity, bandwidth, n_iter, cmean, sigma_out, rho_out, w_out,   \
                     cmin, czero)
  if (k <= 16)
    SH_SMOOTH(16);
  else if (k <= 32)
    SH_SMOOTH(32);
  else
    SH_SMOOTH(64);
#undef SH_SMOOTH
  SH_LAUNCH_CHECK("k_smooth_rows");
  hipLaunchKernelGGL(k_global_min, dim3(1), dim3(64), 0, st, cmin, czero, nc, gmin);
  SH_LAUNCH_CHECK("k_global_min");
  hipLaunchKernelGGL(k_fix_zero, dim3(cdiv(n * k, 256)), dim3(256), 0, st, w_out, n * k, gmin);
  SH_LAUNCH_CHECK("k_fix_zero");
  return SH_OK;
}
