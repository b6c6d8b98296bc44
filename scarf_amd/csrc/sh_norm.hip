// Library-size normalisation, per-feature statistics, dense z tiles and the sparse
// z-score projection (HBM-bound CSR passes).
//
// Reference semantics (scarf tree): assay.py:41-64 (norm_lib_size[_log]),
// assay.py:812-828 (renormalize_subset library size, 0 -> 1),
// graph_datastore.py:767-796 + utils.py:143-154 (mean/std + clean_array),
// ann.py:129-138 and 191-192 (z-score and reducer).
#include <float.h>

#include "sh_common.h"

namespace sh {

// One wave per row: library size over the mapped columns; 0 -> 1.
__global__ void __launch_bounds__(256) k_libsize(const uint32_t *__restrict__ vals,
                                                 const int32_t *__restrict__ cols,
                                                 const int64_t *__restrict__ rowptr,
                                                 int64_t n_rows,
                                                 const int32_t *__restrict__ col_map,
                                                 int32_t n_in, double *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n_rows) return;
  const int64_t a = rowptr[row], b = rowptr[row + 1];
  uint64_t s = 0;
  for (int64_t t = a + lane; t < b; t += 64) {
    if (map_col(col_map, cols[t], n_in) >= 0) s += vals[t];
  }
  s = wave_sum_u64(s);
  if (lane == 0) out[row] = s == 0 ? 1.0 : (double)s;
}

// Per-feature FP64 sums of x and x^2 over a slab of rows, for columns [c0, c0+cb).
// Slabs never straddle a chunk: slab s covers rows [c*chunk + j*slab_rows, ...) of
// chunk c = s / spc, j = s % spc, so every chunk of the matrix is done in one launch.
// One wave per slab.  The slab's non-zeros are a contiguous range of the CSR arrays and
// are streamed flat, 64 * kStatsU at a time with the next batch's loads issued before
// the current batch is added, so each wave keeps 2 * kStatsU coalesced loads in flight
// whatever the row lengths.  Each lane finds the row of its entry by walking the slab's
// row pointers (staged in LDS with the library sizes; entries only move forward).  The
// wave owns its FP64 accumulators in LDS; entries of one column from different rows
// meet in one ds_add, whose lanes the LDS applies in a fixed order, so the sums are
// bitwise reproducible.
constexpr int kStatsU = 4;
constexpr int kStatsSlabRows = 384;  // rows per slab (row pointers + library sizes: 6 KB)
__global__ void __launch_bounds__(64) k_norm_stats(
    const uint32_t *__restrict__ vals, const int32_t *__restrict__ cols,
    const int64_t *__restrict__ rowptr, int64_t n_rows, int64_t chunk_rows, int spc,
    int64_t slab_rows, const int32_t *__restrict__ col_map, int32_t n_in, int32_t d,
    int32_t cb, double sf, int log_t, const double *__restrict__ libsize,
    double *__restrict__ part /* [n_slabs][2][d] */) {
  extern __shared__ double acc[];  // [2][cb] then row pointers [slab_rows + 1], libsize
  const int lane = threadIdx.x;
  const int32_t c0 = blockIdx.y * cb;
  const int32_t cend = min(d, c0 + cb);
  const int32_t ncol = cend - c0;
  int64_t *srp = reinterpret_cast<int64_t *>(acc + 2 * cb);
  double *sls = acc + 2 * cb + (slab_rows + 1);
  const int64_t chunk = blockIdx.x / spc, j_in = blockIdx.x % spc;
  const int64_t c_end = min(n_rows, (chunk + 1) * chunk_rows);
  const int64_t r0 = chunk * chunk_rows + j_in * slab_rows;
  // the last chunk can be shorter than spc slabs: its trailing slabs are empty
  const int64_t nrow = max<int64_t>(0, min(c_end, r0 + slab_rows) - r0);
  for (int i = lane; i < 2 * cb; i += 64) acc[i] = 0.0;
  for (int64_t i = lane; i <= nrow; i += 64) srp[i] = nrow > 0 ? rowptr[r0 + i] : 0;
  for (int64_t i = lane; i < nrow; i += 64) sls[i] = libsize[r0 + i];
  __syncthreads();
  const int64_t e0 = srp[0], e1 = srp[nrow];
  int row = 0;  // slab-relative row of this lane's current entry
  uint32_t v[kStatsU], vn[kStatsU];
  int32_t c[kStatsU], cn[kStatsU];
#pragma unroll
  for (int u = 0; u < kStatsU; ++u) {
    const int64_t e = e0 + lane + 64 * u;
    c[u] = e < e1 ? cols[e] : -1;
    v[u] = e < e1 ? vals[e] : 0u;
  }
  for (int64_t base = e0; base < e1; base += 64 * kStatsU) {
    const int64_t nb = base + 64 * kStatsU;
#pragma unroll
    for (int u = 0; u < kStatsU; ++u) {  // next batch in flight during this one
      const int64_t e = nb + lane + 64 * u;
      cn[u] = e < e1 ? cols[e] : -1;
      vn[u] = e < e1 ? vals[e] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kStatsU; ++u) {
      const int64_t e = base + lane + 64 * u;
      if (e < e1) {
        while (srp[row + 1] <= e) ++row;
        const int f = map_col(col_map, c[u], n_in);
        if (f >= c0 && f < cend) {
          const double x = norm_value(v[u], sf, sls[row], log_t);
          atomicAdd(&acc[f - c0], x);
          atomicAdd(&acc[cb + f - c0], x * x);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kStatsU; ++u) {
      c[u] = cn[u];
      v[u] = vn[u];
    }
  }
  __syncthreads();
  double *dst = part + (size_t)blockIdx.x * 2 * d;
  for (int j = lane; j < ncol; j += 64) {
    dst[c0 + j] = acc[j];
    dst[d + c0 + j] = acc[cb + j];
  }
}

// One thread per (chunk, feature): the chunk's slabs summed in slab order.
__global__ void k_reduce_slabs(const double *__restrict__ part, int spc, int64_t n_chunks,
                               int32_t d, double *__restrict__ sum_out,
                               double *__restrict__ sq_out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_chunks * d) return;
  const int64_t c = t / d;
  const int f = (int)(t % d);
  double s = 0.0, q = 0.0;
  for (int i = 0; i < spc; ++i) {
    const size_t base = (size_t)(c * spc + i) * 2 * d;
    s += part[base + f];
    q += part[base + d + f];
  }
  sum_out[t] = s;
  sq_out[t] = q;
}

// clean_array (utils.py:143-154): nan_to_num maps NaN -> 0 and +-inf -> +-DBL_MAX
// first, so the later "+-inf -> 0" step never fires; then 0 -> fill.
__device__ __forceinline__ double clean(double v, double fill) {
  if (!(v == v)) v = 0.0;
  else if (isinf(v)) v = v > 0 ? DBL_MAX : -DBL_MAX;
  return v == 0.0 ? fill : v;
}

// Chan et al. pairwise merge of per-chunk (n, mean, M2), chunk order fixed.
__global__ void k_stats_merge(const double *__restrict__ sums, const double *__restrict__ sqs,
                              const int64_t *__restrict__ counts, int32_t n_chunks, int32_t d,
                              double *__restrict__ mu, double *__restrict__ sigma,
                              double *__restrict__ chunk_mean) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d) return;
  double n = 0.0, mean = 0.0, m2 = 0.0;
  for (int c = 0; c < n_chunks; ++c) {
    double nc = (double)counts[c];
    if (nc <= 0) {
      if (chunk_mean) chunk_mean[(size_t)c * d + f] = 0.0;
      continue;
    }
    double s = sums[(size_t)c * d + f], q = sqs[(size_t)c * d + f];
    double mc = s / nc;
    double m2c = q - s * mc;
    if (m2c < 0) m2c = 0;
    if (chunk_mean) chunk_mean[(size_t)c * d + f] = mc;
    double nn = n + nc;
    double delta = mc - mean;
    mean += delta * (nc / nn);
    m2 += m2c + delta * delta * (n * nc / nn);
    n = nn;
  }
  double var = n > 0 ? m2 / n : 0.0;
  mu[f] = clean(n > 0 ? mean : 0.0, 0.0);
  sigma[f] = clean(sqrt(var), 1.0);
}

// One workgroup per row batch: the dense row is assembled in LDS, then streamed out.
template <typename T>
__global__ void __launch_bounds__(256) k_densify(
    const uint32_t *__restrict__ vals, const int32_t *__restrict__ cols,
    const int64_t *__restrict__ rowptr, const int64_t *__restrict__ rows, int64_t n_rows,
    const int32_t *__restrict__ col_map, int32_t n_in, int32_t d, double sf, int log_t,
    const double *__restrict__ libsize, const double *__restrict__ center,
    const double *__restrict__ scale, double clip, T *__restrict__ out, int64_t ld,
    int rows_per_block) {
  extern __shared__ __attribute__((aligned(16))) unsigned char dbuf_raw[];
  T *buf = reinterpret_cast<T *>(dbuf_raw);  // [ld]
  const int64_t rb = (int64_t)blockIdx.x * rows_per_block;
  for (int rr = 0; rr < rows_per_block; ++rr) {
    const int64_t i = rb + rr;
    if (i >= n_rows) break;
    const int64_t r = rows ? rows[i] : i;
    for (int64_t f = threadIdx.x; f < ld; f += blockDim.x) {
      double v = f == d ? 1.0 : 0.0;  // column d (padding) holds ones: Z^T Z gets column sums
      if (f < d) {
        v = (0.0 - center[f]) * scale[f];
        if (clip > 0.0) v = fmin(fmax(v, -clip), clip);
      }
      buf[f] = (T)v;
    }
    __syncthreads();
    const int64_t a = rowptr[r], b = rowptr[r + 1];
    const double l = libsize[r];
    for (int64_t t = a + threadIdx.x; t < b; t += blockDim.x) {
      int f = map_col(col_map, cols[t], n_in);
      if (f >= 0) {
        double x = norm_value(vals[t], sf, l, log_t);
        double v = (x - center[f]) * scale[f];
        if (clip > 0.0) v = fmin(fmax(v, -clip), clip);
        buf[f] = (T)v;
      }
    }
    __syncthreads();
    float4 *o4 = reinterpret_cast<float4 *>(out + i * ld);
    const float4 *b4 = reinterpret_cast<const float4 *>(buf);
    const int64_t n16 = ld * (int64_t)sizeof(T) / 16;
    for (int64_t f = threadIdx.x; f < n16; f += blockDim.x) o4[f] = b4[f];
    __syncthreads();
  }
}

// Sparse projection E[r] = c + sum over the row's non-zeros of (z(x) - z(0)) W[f]
// (c, the projection of an all-zero row, is precomputed by the caller).  One wave per
// row.  A batch of 64 non-zeros is normalised lane-parallel (x in FP32 from log1pf, the
// z-delta, the mapped feature); the wave's four 16-lane groups then take every fourth
// entry, and lane c of a group loads columns 4c..4c+3 of that entry's W row with one
// 16-byte load (W rows padded to a multiple of 4 floats) and adds delta * W into four
// FP32 accumulators.  Each group thus gathers a whole W row per load instruction; the
// groups' sums are added in a fixed order at the end of the row.  DPL = 64-column
// passes (D <= 64 * DPL).
constexpr int kProjGroup = 16;
template <int DPL>
__global__ void __launch_bounds__(256) k_project(
    const uint32_t *__restrict__ vals, const int32_t *__restrict__ cols,
    const int64_t *__restrict__ rowptr, int64_t n_rows, const int32_t *__restrict__ col_map,
    int32_t n_in, float sf, int log_t, const double *__restrict__ libsize,
    const float *__restrict__ mu, const float *__restrict__ inv_sigma, float clip,
    const float *__restrict__ W, int64_t ldw, int32_t D, const float *__restrict__ cvec,
    float *__restrict__ E, int64_t lde) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n_rows) return;
  const int grp = lane / kProjGroup, c4 = lane % kProjGroup;
  const int64_t a = rowptr[row], b = rowptr[row + 1];
  double l;
  if (libsize) {
    l = libsize[row];
  } else {  // library size over the mapped columns (the mapping target's)
    uint64_t s = 0;
    for (int64_t t = a + lane; t < b; t += 64)
      if (map_col(col_map, cols[t], n_in) >= 0) s += vals[t];
    s = wave_sum_u64(s);
    l = s == 0 ? 1.0 : (double)s;
  }
  const float scale = sf / (float)l;
  const int64_t ld4 = ldw / 4;
  const float4 *W4 = reinterpret_cast<const float4 *>(W);
  bool col_ok[DPL];
#pragma unroll
  for (int p = 0; p < DPL; ++p) col_ok[p] = 4 * (c4 + kProjGroup * p) < D;
  float4 acc[DPL];
#pragma unroll
  for (int p = 0; p < DPL; ++p) acc[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int64_t t0 = a; t0 < b; t0 += 64) {
    const int64_t t = t0 + lane;
    int f = -1;
    float delta = 0.f;
    if (t < b) {
      f = map_col(col_map, cols[t], n_in);
      if (f >= 0) {
        const float v = (float)vals[t] * scale;
        const float xv = log_t ? log1pf(v) : v;
        if (clip > 0.f) {
          const float zs = (xv - mu[f]) * inv_sigma[f];
          const float z0 = (0.f - mu[f]) * inv_sigma[f];
          delta = fminf(fmaxf(zs, -clip), clip) - fminf(fmaxf(z0, -clip), clip);
        } else {
          delta = xv * inv_sigma[f];
        }
      }
    }
    // The trip count is the same for the whole wave, so every lane takes part in every
    // shuffle (a shuffle from a lane that has left the loop reads 0); lanes past the row's
    // end hold f = -1 and add nothing.
    const int cnt = (int)min<int64_t>(64, b - t0);
#pragma unroll 4
    for (int j0 = 0; j0 < cnt; j0 += 64 / kProjGroup) {
      const int j = j0 + grp;
      const int fj = __shfl(f, j, 64);
      const float dj = __shfl(delta, j, 64);
      if (fj >= 0) {
#pragma unroll
        for (int p = 0; p < DPL; ++p) {
          if (col_ok[p]) {
            const float4 w = W4[(int64_t)fj * ld4 + c4 + kProjGroup * p];
            acc[p].x = fmaf(dj, w.x, acc[p].x);
            acc[p].y = fmaf(dj, w.y, acc[p].y);
            acc[p].z = fmaf(dj, w.z, acc[p].z);
            acc[p].w = fmaf(dj, w.w, acc[p].w);
          }
        }
      }
    }
  }
  // groups 0 + 1 and 2 + 3, then the two pairs
#pragma unroll
  for (int p = 0; p < DPL; ++p) {
    acc[p].x += __shfl_xor(acc[p].x, 16, 64);
    acc[p].y += __shfl_xor(acc[p].y, 16, 64);
    acc[p].z += __shfl_xor(acc[p].z, 16, 64);
    acc[p].w += __shfl_xor(acc[p].w, 16, 64);
    acc[p].x += __shfl_xor(acc[p].x, 32, 64);
    acc[p].y += __shfl_xor(acc[p].y, 32, 64);
    acc[p].z += __shfl_xor(acc[p].z, 32, 64);
    acc[p].w += __shfl_xor(acc[p].w, 32, 64);
    const int col = 4 * (c4 + kProjGroup * p);
    if (grp == 0 && col_ok[p]) {
      float *o = E + row * lde + col;
      o[0] = acc[p].x + cvec[col];
      if (col + 1 < D) o[1] = acc[p].y + cvec[col + 1];
      if (col + 2 < D) o[2] = acc[p].z + cvec[col + 2];
      if (col + 3 < D) o[3] = acc[p].w + cvec[col + 3];
    }
  }
}

}  // namespace sh

using namespace sh;

struct StatsPlan {
  int64_t n_chunks, spc, slab_rows;
};
static StatsPlan stats_plan(int64_t n_rows, int64_t chunk_rows) {
  StatsPlan p;
  if (chunk_rows <= 0 || chunk_rows > n_rows) chunk_rows = n_rows > 0 ? n_rows : 1;
  p.n_chunks = cdiv(n_rows > 0 ? n_rows : 1, chunk_rows);
  // slabs of at most kStatsSlabRows rows: one wave each
  int64_t spc = cdiv(chunk_rows, kStatsSlabRows);
  if (spc < 1) spc = 1;
  p.spc = spc;
  p.slab_rows = cdiv(chunk_rows, spc);
  return p;
}

extern "C" size_t sh_norm_stats_workspace(int64_t n_rows, int64_t chunk_rows, int32_t d) {
  StatsPlan p = stats_plan(n_rows, chunk_rows);
  return (size_t)(p.n_chunks * p.spc) * 2 * (size_t)(d > 0 ? d : 1) * sizeof(double);
}

extern "C" int sh_norm_stats(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                             int64_t n_rows, int64_t chunk_rows, const int32_t *col_map,
                             int32_t n_cols_in, int32_t d, double sf, int log_transform,
                             const double *libsize_in, double *libsize_out, double *sum_out,
                             double *sumsq_out, void *workspace, size_t workspace_bytes,
                             sh_stream_t stream) {
  hipStream_t st = (hipStream_t)stream;
  SH_ARG(n_rows >= 0 && d > 0 && n_cols_in > 0, "sh_norm_stats: bad sizes n_rows=%lld d=%d n_cols_in=%d",
         (long long)n_rows, d, n_cols_in);
  SH_ARG(col_map || d == n_cols_in, "sh_norm_stats: col_map NULL requires d == n_cols_in");
  SH_ARG(rowptr && sum_out && sumsq_out, "sh_norm_stats: NULL pointer");
  SH_ARG(sf > 0, "sh_norm_stats: size factor must be > 0");
  if (n_rows == 0) {
    SH_HIP(hipMemsetAsync(sum_out, 0, sizeof(double) * d, st));
    SH_HIP(hipMemsetAsync(sumsq_out, 0, sizeof(double) * d, st));
    return SH_OK;
  }
  const StatsPlan plan = stats_plan(n_rows, chunk_rows);
  const int64_t chunk = (chunk_rows > 0 && chunk_rows < n_rows) ? chunk_rows : n_rows;
  const size_t need = sh_norm_stats_workspace(n_rows, chunk_rows, d);
  if (workspace_bytes < need || !workspace) {
    set_error("sh_norm_stats: workspace %zu < %zu bytes", workspace_bytes, need);
    return SH_EWORKSPACE;
  }
  const double *ls = libsize_in;
  if (!ls) {
    SH_ARG(libsize_out, "sh_norm_stats: libsize_out required when libsize_in is NULL");
    hipLaunchKernelGGL(k_libsize, dim3(cdiv(n_rows, 4)), dim3(256), 0, st, vals, cols, rowptr,
                       n_rows, col_map, n_cols_in, libsize_out);
    SH_LAUNCH_CHECK("k_libsize");
    ls = libsize_out;
  } else if (libsize_out && libsize_out != libsize_in) {
    SH_HIP(hipMemcpyAsync(libsize_out, libsize_in, sizeof(double) * n_rows,
                          hipMemcpyDeviceToDevice, st));
  }
  // accumulators of at most 2 x 2048 doubles (32 KiB) + 6 KiB of row data per wave: four
  // or five waves per CU; wider matrices take column blocks (the CSR is re-read per block)
  const int32_t cb = d <= 2048 ? d : (int32_t)cdiv(d, cdiv(d, 2048));
  const int n_cb = (int)cdiv(d, cb);
  const size_t lds =
      (2 * (size_t)cb + 2 * (size_t)plan.slab_rows + 1) * sizeof(double);
  double *part = (double *)workspace;
  hipLaunchKernelGGL(k_norm_stats, dim3((unsigned)(plan.n_chunks * plan.spc), n_cb),
                     dim3(64), lds, st, vals, cols, rowptr, n_rows, chunk,
                     (int)plan.spc, plan.slab_rows, col_map, n_cols_in, d, cb, sf, log_transform,
                     ls, part);
  SH_LAUNCH_CHECK("k_norm_stats");
  hipLaunchKernelGGL(k_reduce_slabs, dim3(cdiv(plan.n_chunks * d, 256)), dim3(256), 0, st, part,
                     (int)plan.spc, plan.n_chunks, d, sum_out, sumsq_out);
  SH_LAUNCH_CHECK("k_reduce_slabs");
  return SH_OK;
}

extern "C" int sh_stats_merge(const double *sums, const double *sumsqs, const int64_t *counts,
                              int32_t n_chunks, int32_t d, double *mu_out, double *sigma_out,
                              double *chunk_mean_out, sh_stream_t stream) {
  SH_ARG(n_chunks > 0 && d > 0, "sh_stats_merge: bad sizes");
  SH_ARG(sums && sumsqs && counts && mu_out && sigma_out, "sh_stats_merge: NULL pointer");
  hipLaunchKernelGGL(k_stats_merge, dim3(cdiv(d, 256)), dim3(256), 0, (hipStream_t)stream, sums,
                     sumsqs, counts, n_chunks, d, mu_out, sigma_out, chunk_mean_out);
  SH_LAUNCH_CHECK("k_stats_merge");
  return SH_OK;
}

extern "C" int sh_densify(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                          const int64_t *rows, int64_t n_rows, const int32_t *col_map,
                          int32_t n_cols_in, int32_t d, double sf, int log_transform,
                          const double *libsize, const double *center, const double *scale,
                          double clip, int out_f64, void *out, int64_t ld, sh_stream_t stream) {
  SH_ARG(n_rows >= 0 && d > 0 && ld >= d && ld % 4 == 0, "sh_densify: bad sizes d=%d ld=%lld", d,
         (long long)ld);
  SH_ARG(col_map || d == n_cols_in, "sh_densify: col_map NULL requires d == n_cols_in");
  SH_ARG(libsize && center && scale && out && rowptr, "sh_densify: NULL pointer");
  const size_t esz = out_f64 ? sizeof(double) : sizeof(float);
  SH_ARG(ld * esz <= 64 * 1024, "sh_densify: ld=%lld exceeds the 64 KiB row buffer",
         (long long)ld);
  if (n_rows == 0) return SH_OK;
  const int rpb = 2;
  if (out_f64)
    hipLaunchKernelGGL(k_densify<double>, dim3(cdiv(n_rows, rpb)), dim3(256), ld * esz,
                       (hipStream_t)stream, vals, cols, rowptr, rows, n_rows, col_map, n_cols_in,
                       d, sf, log_transform, libsize, center, scale, clip, (double *)out, ld, rpb);
  else
    hipLaunchKernelGGL(k_densify<float>, dim3(cdiv(n_rows, rpb)), dim3(256), ld * esz,
                       (hipStream_t)stream, vals, cols, rowptr, rows, n_rows, col_map, n_cols_in,
                       d, sf, log_transform, libsize, center, scale, clip, (float *)out, ld, rpb);
  SH_LAUNCH_CHECK("k_densify");
  return SH_OK;
}

extern "C" int sh_project(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                          int64_t n_rows, const int32_t *col_map, int32_t n_cols_in, int32_t d,
                          double sf, int log_transform, const double *libsize, const float *mu,
                          const float *inv_sigma, float clip, const float *W, int64_t ldw,
                          int32_t D, const float *c, float *E_out, int64_t lde,
                          sh_stream_t stream) {
  SH_ARG(n_rows >= 0 && d > 0 && D > 0 && D <= 128 && ldw >= D && lde >= D,
         "sh_project: bad sizes d=%d D=%d ldw=%lld lde=%lld", d, D, (long long)ldw,
         (long long)lde);
  SH_ARG(col_map || d == n_cols_in, "sh_project: col_map NULL requires d == n_cols_in");
  SH_ARG(mu && inv_sigma && W && c && E_out && rowptr, "sh_project: NULL pointer");
  if (n_rows == 0) return SH_OK;
  SH_ARG(ldw % 4 == 0 && ((uintptr_t)W % 16) == 0,
         "sh_project: W rows must be 16-byte aligned (ldw a multiple of 4)");
  dim3 grid(cdiv(n_rows, 4)), block(256);
  if (D <= 64)
    hipLaunchKernelGGL(k_project<1>, grid, block, 0, (hipStream_t)stream, vals, cols, rowptr,
                       n_rows, col_map, n_cols_in, (float)sf, log_transform, libsize, mu,
                       inv_sigma, clip, W, ldw, D, c, E_out, lde);
  else
    hipLaunchKernelGGL(k_project<2>, grid, block, 0, (hipStream_t)stream, vals, cols, rowptr,
                       n_rows, col_map, n_cols_in, (float)sf, log_transform, libsize, mu,
                       inv_sigma, clip, W, ldw, D, c, E_out, lde);
  SH_LAUNCH_CHECK("k_project");
  return SH_OK;
}
