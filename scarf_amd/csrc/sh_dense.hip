// Dense blocks of normalised values, as scarf's callers hand them to AnnStream:
//   transform_z(a) = (a - mu) / sigma                      (ann.py:191-192)
//   reducer(a)     = transform_z(a) . loadings, or a . loadings when scaling is off
//                                                          (ann.py:126-138)
// Callers: mapping_datastore.py:204 (transform_ann(reducer(a))), knn_utils.py:42,
// metrics.py:312-315.  Both are FP64, as numpy computes them; z = (x - mu) / sigma with an
// IEEE division, so transform_z equals numpy's bit for bit.  The optional clip is this
// build's switch (scarf has none).
//
// k_dense_project: one workgroup of 256 threads per 64 rows.  The block walks the
// features in tiles of 32: it loads the 64 x 32 tile of x (each row's 32 features are
// 256 contiguous bytes), z-scores it into LDS and stages the 32 x D tile of the
// loadings beside it; thread (r, g) then accumulates row r's outputs g, g + 4, ... in
// FP64 registers, features in increasing order.  The block is HBM-bound on x
// (8 d bytes per row read once); the loadings tile is re-read from L2 by every block.
#include "sh_common.h"

namespace sh {

constexpr int kDenseRows = 64;
constexpr int kDenseTile = 32;
constexpr int kDenseMaxD = 128;
constexpr int kDenseGroups = 4;  // column groups: 256 threads = 64 rows x 4

template <typename T>
__device__ __forceinline__ double zscore(T x, const double *center, const double *sigma,
                                         int f, double clip) {
  double z = (double)x;
  if (center) z = (z - center[f]) / sigma[f];
  if (clip > 0.0) z = fmin(fmax(z, -clip), clip);
  return z;
}

template <typename T>
__global__ void __launch_bounds__(256) k_dense_z(const T *__restrict__ x, int64_t n, int32_t d,
                                                 int64_t ldx, const double *__restrict__ center,
                                                 const double *__restrict__ sigma, double clip,
                                                 double *__restrict__ z, int64_t ldz) {
  const int64_t row = blockIdx.y;
  for (int32_t f = blockIdx.x * 256 + threadIdx.x; f < d; f += gridDim.x * 256)
    z[row * ldz + f] = zscore(x[row * ldx + f], center, sigma, f, clip);
}

template <typename T, int NPER>  // NPER outputs per thread: D <= 4 * NPER
__global__ void __launch_bounds__(256) k_dense_project(
    const T *__restrict__ x, int64_t n, int32_t d, int64_t ldx, const double *__restrict__ center,
    const double *__restrict__ sigma, double clip, const double *__restrict__ W, int32_t D,
    double *__restrict__ out, int64_t ldo) {
  __shared__ double zt[kDenseRows][kDenseTile + 1];
  __shared__ double wt[kDenseTile][kDenseMaxD];
  const int t = threadIdx.x;
  const int r = t & 63, g = t >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * kDenseRows;
  double acc[NPER];
#pragma unroll
  for (int i = 0; i < NPER; ++i) acc[i] = 0.0;
  for (int32_t f0 = 0; f0 < d; f0 += kDenseTile) {
    const int nf = d - f0 < kDenseTile ? d - f0 : kDenseTile;
    // x tile: 64 rows x 32 features; thread t loads feature t % 32 of rows t / 32 + 8 j
    for (int e = t; e < kDenseRows * kDenseTile; e += 256) {
      const int rr = e >> 5, ff = e & 31;
      const int64_t row = row0 + rr;
      double v = 0.0;
      if (row < n && ff < nf) v = zscore(x[row * ldx + f0 + ff], center, sigma, f0 + ff, clip);
      zt[rr][ff] = v;
    }
    for (int e = t; e < kDenseTile * D; e += 256) {
      const int ff = e / D, c = e - ff * D;
      wt[ff][c] = ff < nf ? W[(int64_t)(f0 + ff) * D + c] : 0.0;
    }
    __syncthreads();
    for (int ff = 0; ff < nf; ++ff) {
      const double zv = zt[r][ff];
#pragma unroll
      for (int i = 0; i < NPER; ++i) {
        const int c = g + kDenseGroups * i;
        if (c < D) acc[i] = fma(zv, wt[ff][c], acc[i]);
      }
    }
    __syncthreads();
  }
  const int64_t row = row0 + r;
  if (row < n) {
#pragma unroll
    for (int i = 0; i < NPER; ++i) {
      const int c = g + kDenseGroups * i;
      if (c < D) out[row * ldo + c] = acc[i];
    }
  }
}

template <typename T>
int launch_project(const T *x, int64_t n, int32_t d, int64_t ldx, const double *center,
                   const double *sigma, double clip, const double *W, int32_t D, double *out,
                   int64_t ldo, hipStream_t st) {
  const dim3 grid((uint32_t)cdiv(n, kDenseRows));
  if (D <= 4 * 4)
    hipLaunchKernelGGL((k_dense_project<T, 4>), grid, dim3(256), 0, st, x, n, d, ldx, center,
                       sigma, clip, W, D, out, ldo);
  else if (D <= 4 * 13)
    hipLaunchKernelGGL((k_dense_project<T, 13>), grid, dim3(256), 0, st, x, n, d, ldx, center,
                       sigma, clip, W, D, out, ldo);
  else
    hipLaunchKernelGGL((k_dense_project<T, 32>), grid, dim3(256), 0, st, x, n, d, ldx, center,
                       sigma, clip, W, D, out, ldo);
  SH_LAUNCH_CHECK("k_dense_project");
  return SH_OK;
}

}  // namespace sh

using namespace sh;

extern "C" int sh_dense_z(const void *x, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                          const double *center, const double *sigma, double clip, double *z,
                          int64_t ldz, sh_stream_t stream) {
  SH_ARG(n >= 0 && d >= 1 && ldx >= d && ldz >= d, "sh_dense_z: bad sizes");
  SH_ARG(x_dtype == SH_F32 || x_dtype == SH_F64, "sh_dense_z: x must be FP32 or FP64");
  SH_ARG((center == nullptr) == (sigma == nullptr), "sh_dense_z: center and sigma go together");
  if (n == 0) return SH_OK;
  SH_ARG(x && z, "sh_dense_z: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const uint32_t gx = (uint32_t)(cdiv(d, 256) < 8 ? cdiv(d, 256) : 8);
  // one grid row per matrix row, in batches within the grid's y range
  for (int64_t a = 0; a < n; a += 65535) {
    const int64_t b = n - a < 65535 ? n - a : 65535;
    const dim3 grid(gx, (uint32_t)b);
    if (x_dtype == SH_F64)
      hipLaunchKernelGGL(k_dense_z<double>, grid, dim3(256), 0, st, (const double *)x + a * ldx,
                         b, d, ldx, center, sigma, clip, z + a * ldz, ldz);
    else
      hipLaunchKernelGGL(k_dense_z<float>, grid, dim3(256), 0, st, (const float *)x + a * ldx, b,
                         d, ldx, center, sigma, clip, z + a * ldz, ldz);
    SH_LAUNCH_CHECK("k_dense_z");
  }
  return SH_OK;
}

extern "C" int sh_dense_project(const void *x, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                                const double *center, const double *sigma, double clip,
                                const double *W, int32_t D, double *out, int64_t ldo,
                                sh_stream_t stream) {
  SH_ARG(n >= 0 && d >= 1 && ldx >= d && D >= 1 && D <= kDenseMaxD && ldo >= D,
         "sh_dense_project: bad sizes (1 <= D <= %d)", kDenseMaxD);
  SH_ARG(x_dtype == SH_F32 || x_dtype == SH_F64, "sh_dense_project: x must be FP32 or FP64");
  SH_ARG((center == nullptr) == (sigma == nullptr),
         "sh_dense_project: center and sigma go together");
  if (n == 0) return SH_OK;
  SH_ARG(x && W && out, "sh_dense_project: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  if (x_dtype == SH_F64)
    return launch_project((const double *)x, n, d, ldx, center, sigma, clip, W, D, out, ldo, st);
  return launch_project((const float *)x, n, d, ldx, center, sigma, clip, W, D, out, ldo, st);
}
