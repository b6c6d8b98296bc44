// Feature x feature Gram matrix Z^T Z of a dense FP64 or FP32 chunk on MFMA (gfx950).
//
// This replaces the SVD at the heart of IncrementalPCA.partial_fit
// (sklearn decomposition/_incremental_pca.py:345-363, driven from scarf ann.py:241,250):
// the right singular vectors of the augmented chunk matrix are the eigenvectors of its
// Gram matrix, so each chunk contributes Z_c^T Z_c, merged on the host in FP64.
//
// Precision: FP64 Z on v_mfma_f64_16x16x4f64 is the default.  Rounding z to FP32
// perturbs the data by ~6e-8 relative, which the small eigen-gaps of the trailing
// (noise) components amplify past 1e-4; the FP32 path (v_mfma_f32_32x32x2f32) is kept
// for callers that need only the leading components.
//
// Layout: Z is row-major [n x ld], ld a multiple of 128, padded columns zero.
// Work split: 128 x 128 output tiles over the upper triangle (bi <= bj) times S row
// slabs.  Each workgroup (4 waves, 2 x 2) accumulates its tile over its slab with
// v_mfma_f32_32x32x2f32 (each wave: 2 x 2 MFMA tiles = 64 x 64) and writes an FP32
// partial; a second kernel sums the slabs in FP64, in slab order (deterministic).
// Workgroups that share a slab are placed on the same XCD (blockIdx % 8), so the
// slab's rows are fetched from HBM once per XCD L2 rather than once per tile.
#include "sh_common.h"

namespace sh {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kGT = 128;  // output tile edge
constexpr int kGK = 16;   // rows per LDS stage

__global__ void __launch_bounds__(256) k_gram(const float *__restrict__ Z, int64_t n, int64_t ld,
                                              int nb, int T, int S, int64_t rows_per_slab,
                                              float *__restrict__ part) {
  __shared__ float As[kGK][kGT];
  __shared__ float Bs[kGK][kGT];
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int j = b >> 3;
  const int slab = xcd + 8 * (j / T);
  const int t = j % T;
  if (slab >= S) return;
  int bi, bj;
  tile_coords(t, nb, bi, bj);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int64_t r_begin = (int64_t)slab * rows_per_slab;
  const int64_t r_end = min(n, r_begin + rows_per_slab);

  floatx16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][c][v] = 0.f;

  // each thread stages 2 float4 of A and 2 of B per stage
  const float *Acol = Z + (int64_t)bi * kGT;
  const float *Bcol = Z + (int64_t)bj * kGT;
  float4 ra[2], rb[2];
  auto load_stage = [&](int64_t r0) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      int e = tid + 256 * p;  // 0..511
      int row = e >> 5, c4 = e & 31;
      int64_t r = r0 + row;
      if (r < r_end) {
        ra[p] = *reinterpret_cast<const float4 *>(Acol + r * ld + c4 * 4);
        rb[p] = *reinterpret_cast<const float4 *>(Bcol + r * ld + c4 * 4);
      } else {
        ra[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        rb[p] = ra[p];
      }
    }
  };
  auto store_stage = [&]() {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      int e = tid + 256 * p;
      int row = e >> 5, c4 = e & 31;
      *reinterpret_cast<float4 *>(&As[row][c4 * 4]) = ra[p];
      *reinterpret_cast<float4 *>(&Bs[row][c4 * 4]) = rb[p];
    }
  };

  int64_t r0 = r_begin;
  if (r0 < r_end) load_stage(r0);
  while (r0 < r_end) {
    store_stage();
    __syncthreads();
    const int64_t rn = r0 + kGK;
    if (rn < r_end) load_stage(rn);
    const int kr = lane >> 5, cl = lane & 31;
#pragma unroll
    for (int kk = 0; kk < kGK; kk += 2) {
      float a0 = As[kk + kr][wm * 64 + cl];
      float a1 = As[kk + kr][wm * 64 + 32 + cl];
      float b0 = Bs[kk + kr][wn * 64 + cl];
      float b1 = Bs[kk + kr][wn * 64 + 32 + cl];
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
    r0 = rn;
  }
  // D layout of a 32x32 MFMA tile: lane l, value v -> row 8*(v/4) + 4*(l/32) + v%4,
  // column l%32.
  float *dst = part + ((int64_t)slab * T + t) * (kGT * kGT);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        int m = wm * 64 + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        int nn = wn * 64 + c * 32 + (lane & 31);
        dst[m * kGT + nn] = acc[a][c][v];
      }
}

// FP64 variant: v_mfma_f64_16x16x4f64, each wave a 64 x 64 block = 4 x 4 MFMA tiles
// of 16 x 16.  A/B operand: lane l supplies row k = l/16 of the 4-row step and
// column l%16; D: lane l holds rows 4*i + l/16 (i < 4, interleaved) of column l%16.
typedef double doublex4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_gram64(const double *__restrict__ Z, int64_t n,
                                                int64_t ld, int nb, int T, int S,
                                                int64_t rows_per_slab, double *__restrict__ part) {
  __shared__ double As[kGK][kGT];
  __shared__ double Bs[kGK][kGT];
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int j = b >> 3;
  const int slab = xcd + 8 * (j / T);
  const int t = j % T;
  if (slab >= S) return;
  int bi, bj;
  tile_coords(t, nb, bi, bj);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int64_t r_begin = (int64_t)slab * rows_per_slab;
  const int64_t r_end = min(n, r_begin + rows_per_slab);
  doublex4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = doublex4{0.0, 0.0, 0.0, 0.0};
  const double *Acol = Z + (int64_t)bi * kGT;
  const double *Bcol = Z + (int64_t)bj * kGT;
  double2 ra[4], rb[4];
  // stage = 16 rows x 128 doubles = 1024 double2 per operand; 4 per thread
#define SH_G64_LOAD(r0)                                                            \
  _Pragma("unroll") for (int p = 0; p < 4; ++p) {                                  \
    const int e = tid + 256 * p;                                                   \
    const int row = e >> 6, c2 = e & 63;                                           \
    const int64_t r = (r0) + row;                                                  \
    if (r < r_end) {                                                               \
      ra[p] = *reinterpret_cast<const double2 *>(Acol + r * ld + c2 * 2);          \
      rb[p] = *reinterpret_cast<const double2 *>(Bcol + r * ld + c2 * 2);          \
    } else {                                                                       \
      ra[p] = make_double2(0.0, 0.0);                                              \
      rb[p] = ra[p];                                                               \
    }                                                                              \
  }
  int64_t r0 = r_begin;
  if (r0 < r_end) {
    SH_G64_LOAD(r0)
  }
  while (r0 < r_end) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int e = tid + 256 * p;
      const int row = e >> 6, c2 = e & 63;
      *reinterpret_cast<double2 *>(&As[row][c2 * 2]) = ra[p];
      *reinterpret_cast<double2 *>(&Bs[row][c2 * 2]) = rb[p];
    }
    __syncthreads();
    const int64_t rn = r0 + kGK;
    if (rn < r_end) {
      SH_G64_LOAD(rn)
    }
    const int kr = lane >> 4, cl = lane & 15;
#pragma unroll
    for (int kk = 0; kk < kGK; kk += 4) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk + kr][wm * 64 + a * 16 + cl];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = Bs[kk + kr][wn * 64 + c * 16 + cl];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
    __syncthreads();
    r0 = rn;
  }
#undef SH_G64_LOAD
  double *dst = part + ((int64_t)slab * T + t) * (kGT * kGT);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = wm * 64 + a * 16 + 4 * i + (lane >> 4);
        const int nn = wn * 64 + c * 16 + (lane & 15);
        dst[m * kGT + nn] = acc[a][c][i];
      }
}

// Sums the slab partials of the upper-triangle tiles into G (d x d).  With a ones
// column at index d (ones_col), entry (i, d) of the extended Gram is the column sum i.
template <typename PT>
__global__ void __launch_bounds__(256) k_gram_reduce(const PT *__restrict__ part, int T, int S,
                                                     int nb, int32_t d, int ones_col,
                                                     int accumulate, double *__restrict__ G,
                                                     double *__restrict__ colsum) {
  const int t = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;  // element within the tile
  if (e >= kGT * kGT) return;
  int bi, bj;
  tile_coords(t, nb, bi, bj);
  const int m = e / kGT, nn = e % kGT;
  const int i = bi * kGT + m, j = bj * kGT + nn;
  const bool col = ones_col && j == d && i < d;
  if (!col && (i >= d || j >= d)) return;
  if (bi == bj && m > nn) return;  // diagonal tiles: upper half only
  double s = 0.0;
  for (int sl = 0; sl < S; ++sl) s += (double)part[((int64_t)sl * T + t) * (kGT * kGT) + e];
  if (col) {
    colsum[i] = accumulate ? colsum[i] + s : s;
    return;
  }
  if (accumulate) s += G[(int64_t)i * d + j];
  G[(int64_t)i * d + j] = s;
  G[(int64_t)j * d + i] = s;
}

// Column sums in FP64: R row parts of partial sums, then a fixed-order reduction.
constexpr int kColParts = 64;
template <typename ZT>
__global__ void k_colsum_part(const ZT *__restrict__ Z, int64_t n, int64_t ld, int32_t d,
                              double *__restrict__ part) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (f >= d) return;
  const int64_t per = (n + kColParts - 1) / kColParts;
  const int64_t a = p * per, b = min(n, a + per);
  double s = 0.0;
  for (int64_t r = a; r < b; ++r) s += (double)Z[r * ld + f];
  part[(int64_t)p * d + f] = s;
}
__global__ void k_colsum_reduce(const double *__restrict__ part, int32_t d, int accumulate,
                                double *__restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d) return;
  double s = 0.0;
  for (int p = 0; p < kColParts; ++p) s += part[(int64_t)p * d + f];
  out[f] = accumulate ? out[f] + s : s;
}

struct GramPlan {
  int nb, T, S;
  int64_t rows_per_slab;
  size_t part_bytes, col_bytes;
};

static GramPlan gram_plan(int64_t n, int32_t d, int z_f64) {
  GramPlan p;
  p.nb = (int)cdiv(d, kGT);
  p.T = p.nb * (p.nb + 1) / 2;
  // enough workgroups to fill 256 CUs a few times over, slabs of >= 256 rows
  int S = (int)cdiv(2048, p.T);
  S = (int)align_up((size_t)(S < 8 ? 8 : S), 8);
  int64_t max_s = cdiv(n, 256);
  if (max_s < 8) max_s = 8;
  if (S > max_s) S = (int)align_up((size_t)max_s, 8);
  p.S = S;
  p.rows_per_slab = align_up((size_t)cdiv(n > 0 ? n : 1, S), kGK);
  p.part_bytes = (size_t)S * p.T * kGT * kGT * (z_f64 ? sizeof(double) : sizeof(float));
  p.col_bytes = (size_t)kColParts * d * sizeof(double);
  return p;
}

}  // namespace sh

using namespace sh;

extern "C" size_t sh_gram_workspace(int64_t n, int32_t d, int z_f64) {
  if (d <= 0) return 0;
  GramPlan p = gram_plan(n, d + 1, z_f64);  // room for the ones column
  return p.part_bytes + p.col_bytes;
}

extern "C" int sh_gram(const void *Z, int z_f64, int64_t n, int32_t d, int64_t ld, int ones_col,
                       int accumulate, double *G_out, double *colsum_out, void *workspace,
                       size_t workspace_bytes, sh_stream_t stream) {
  hipStream_t st = (hipStream_t)stream;
  SH_ARG(n >= 0 && d > 0, "sh_gram: bad sizes n=%lld d=%d", (long long)n, d);
  SH_ARG(ld % kGT == 0 && ld >= d, "sh_gram: ld=%lld must be a multiple of 128 and >= d",
         (long long)ld);
  SH_ARG(Z && G_out, "sh_gram: NULL pointer");
  SH_ARG(!ones_col || (ld > d && colsum_out), "sh_gram: ones_col needs ld > d and colsum_out");
  const int32_t de = ones_col ? d + 1 : d;  // Gram extent including the ones column
  GramPlan p = gram_plan(n, de, z_f64);
  SH_ARG((int64_t)p.nb * kGT <= ld, "sh_gram: ld too small for the tile grid");
  if (!workspace || workspace_bytes < p.part_bytes + p.col_bytes) {
    set_error("sh_gram: workspace %zu < %zu bytes", workspace_bytes, p.part_bytes + p.col_bytes);
    return SH_EWORKSPACE;
  }
  if (z_f64) {
    double *part = (double *)workspace;
    hipLaunchKernelGGL(k_gram64, dim3(p.T * p.S), dim3(256), 0, st, (const double *)Z, n, ld,
                       p.nb, p.T, p.S, p.rows_per_slab, part);
    SH_LAUNCH_CHECK("k_gram64");
    hipLaunchKernelGGL(k_gram_reduce<double>, dim3(kGT * kGT / 256, p.T), dim3(256), 0, st, part,
                       p.T, p.S, p.nb, d, ones_col, accumulate, G_out, colsum_out);
  } else {
    float *part = (float *)workspace;
    hipLaunchKernelGGL(k_gram, dim3(p.T * p.S), dim3(256), 0, st, (const float *)Z, n, ld, p.nb,
                       p.T, p.S, p.rows_per_slab, part);
    SH_LAUNCH_CHECK("k_gram");
    hipLaunchKernelGGL(k_gram_reduce<float>, dim3(kGT * kGT / 256, p.T), dim3(256), 0, st, part,
                       p.T, p.S, p.nb, d, ones_col, accumulate, G_out, colsum_out);
  }
  SH_LAUNCH_CHECK("k_gram_reduce");
  if (colsum_out && !ones_col) {
    double *cp = (double *)((char *)workspace + p.part_bytes);
    if (z_f64)
      hipLaunchKernelGGL(k_colsum_part<double>, dim3(cdiv(d, 256), kColParts), dim3(256), 0, st,
                         (const double *)Z, n, ld, d, cp);
    else
      hipLaunchKernelGGL(k_colsum_part<float>, dim3(cdiv(d, 256), kColParts), dim3(256), 0, st,
                         (const float *)Z, n, ld, d, cp);
    SH_LAUNCH_CHECK("k_colsum_part");
    hipLaunchKernelGGL(k_colsum_reduce, dim3(cdiv(d, 256)), dim3(256), 0, st, cp, d, accumulate,
                       colsum_out);
    SH_LAUNCH_CHECK("k_colsum_reduce");
  }
  return SH_OK;
}
