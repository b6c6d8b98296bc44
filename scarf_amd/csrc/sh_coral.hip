// CORAL alignment (scarf mapping_utils.py:15-95): feature covariances of the standardised
// query and reference, C = cov(x, ddof = 1) + I (mapping_utils.py:15-19), and the
// alignment A = C_s^{-1/2} C_t^{1/2} (mapping_utils.py:22-38).
//
// scarf takes the fractional powers with scipy's fractional_matrix_power on the host.  Here
// both come from the coupled Newton-Schulz iteration on FP64 MFMA GEMMs (sh_dgemm):
//   Y0 = C / c, Z0 = I, P = Z Y, M = (3I - P) / 2, Y <- Y M, Z <- M Z
// converges quadratically to (C/c)^{1/2} and (C/c)^{-1/2} when the spectrum of C/c lies in
// (0, 1]: c = ||C||_F bounds the largest eigenvalue and C >= I keeps the smallest away
// from zero.  Every reduction runs in a fixed order, so the result is deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sh_common.h"

namespace sh {

constexpr int kCoBlock = 256;
constexpr int kCoParts = 1024;  // partial sums of the Frobenius norms

// C_ij = (G_ij - n m_i m_j) / (n - 1) with m = colsum / n, +inf -> 0 (mapping_utils.py:17),
// + ridge on the diagonal.  Each pair is computed once, as (min, max), so C is exactly
// symmetric.
__global__ void k_cov_from_gram(const double *__restrict__ G, int64_t ldg,
                                const double *__restrict__ cs, double n, int32_t d,
                                double ridge, double *__restrict__ C) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)d * d) return;
  const int r = (int)(t / d), c = (int)(t % d);
  const int i = min(r, c), j = max(r, c);
  const double mi = cs[i] / n, mj = cs[j] / n;
  double v = (G[(int64_t)i * ldg + j] - n * mi * mj) / (n - 1.0);
  if (v == __builtin_inf()) v = 0.0;
  if (r == c) v += ridge;
  C[t] = v;
}

// partial[b] = sum of squares of x over block b's grid-stride share (fixed order)
__global__ void k_sumsq_partial(const double *__restrict__ x, int64_t count, int32_t d,
                                int identity_minus, double *__restrict__ partial) {
  __shared__ double red[kCoBlock / 64];
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * kCoBlock + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * kCoBlock) {
    double v = x[t];
    if (identity_minus) v = ((t / d) == (t % d) ? 1.0 : 0.0) - v;
    s = fma(v, v, s);
  }
  s = wave_sum_f64(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < kCoBlock / 64; ++w) b += red[w];
    partial[blockIdx.x] = b;
  }
}

__global__ void k_sum_partials(const double *__restrict__ partial, int32_t np,
                               double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < np; ++b) s += partial[b];
  *out = s;
}

__global__ void k_ns_init(const double *__restrict__ C, int32_t d, double inv_c,
                          double *__restrict__ Y, double *__restrict__ Z) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)d * d) return;
  Y[t] = C[t] * inv_c;
  Z[t] = (t / d) == (t % d) ? 1.0 : 0.0;
}

__global__ void k_ns_mid(const double *__restrict__ P, int32_t d, double *__restrict__ M) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)d * d) return;
  M[t] = ((t / d) == (t % d) ? 1.5 : 0.0) - 0.5 * P[t];
}

__global__ void k_scale2(const double *__restrict__ Y, const double *__restrict__ Z, int64_t count,
                         double sy, double sz, double *__restrict__ out_y,
                         double *__restrict__ out_z) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  if (out_y) out_y[t] = Y[t] * sy;
  if (out_z) out_z[t] = Z[t] * sz;
}

}  // namespace sh

using namespace sh;

extern "C" int sh_cov_from_gram(const double *G, int64_t ldg, const double *colsum, int64_t n,
                                int32_t d, double ridge, double *C, sh_stream_t stream) {
  SH_ARG(d >= 1 && ldg >= d && n >= 1, "sh_cov_from_gram: need d >= 1, ldg >= d, n >= 1");
  SH_ARG(G && colsum && C, "sh_cov_from_gram: NULL pointer");
  hipLaunchKernelGGL(k_cov_from_gram, dim3((uint32_t)cdiv((int64_t)d * d, 256)), dim3(256), 0,
                     (hipStream_t)stream, G, ldg, colsum, (double)n, d, ridge, C);
  SH_LAUNCH_CHECK("k_cov_from_gram");
  return SH_OK;
}

extern "C" size_t sh_spd_sqrt_workspace(int32_t d) {
  if (d < 1) return 0;
  return 4 * align_up((size_t)d * d * sizeof(double), 256) + (kCoParts + 1) * sizeof(double);
}

// sqrt_out = C^{1/2}, inv_sqrt_out = C^{-1/2} (either may be NULL).  Iterates until
// ||I - Z Y||_F <= tol * sqrt(d), or until it stops falling once below 1e-6, at most
// max_iter times; *iters and *resid report the count and the last residual.  Synchronises
// the stream once per iteration to read the residual.
extern "C" int sh_spd_sqrt(const double *C, int32_t d, double *sqrt_out, double *inv_sqrt_out,
                           int32_t max_iter, double tol, int32_t *iters, double *resid,
                           void *ws, size_t ws_bytes, sh_stream_t stream) {
  SH_ARG(d >= 1 && max_iter >= 1, "sh_spd_sqrt: need d >= 1, max_iter >= 1");
  SH_ARG(C && ws && iters && resid, "sh_spd_sqrt: NULL pointer");
  const size_t need = sh_spd_sqrt_workspace(d);
  if (ws_bytes < need) {
    set_error("sh_spd_sqrt: workspace %zu bytes < %zu", ws_bytes, need);
    return SH_EWORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t d2 = (int64_t)d * d;
  const size_t mb = align_up((size_t)d2 * sizeof(double), 256);
  double *buf[4];
  for (int b = 0; b < 4; ++b) buf[b] = (double *)((char *)ws + b * mb);
  double *partial = (double *)((char *)ws + 4 * mb);
  double *scalar = partial + kCoParts;
  const int np = (int)std::min<int64_t>(kCoParts, std::max<int64_t>(1, cdiv(d2, kCoBlock)));
  const dim3 g1((uint32_t)cdiv(d2, 256));
  auto sumsq = [&](const double *x, int identity_minus, double *out) -> int {
    hipLaunchKernelGGL(k_sumsq_partial, dim3((uint32_t)np), dim3(kCoBlock), 0, s, x, d2, d,
                       identity_minus, partial);
    SH_LAUNCH_CHECK("k_sumsq_partial");
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, np, scalar);
    SH_LAUNCH_CHECK("k_sum_partials");
    SH_HIP(hipMemcpyAsync(out, scalar, sizeof(double), hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    return SH_OK;
  };
  double fro2 = 0.0;
  int rc = sumsq(C, 0, &fro2);
  if (rc != SH_OK) return rc;
  if (!(fro2 > 0.0) || !std::isfinite(fro2)) {
    set_error("sh_spd_sqrt: ||C||_F is %g", std::sqrt(fro2));
    return SH_EINVAL;
  }
  const double c = std::sqrt(fro2);
  double *Y = buf[0], *Z = buf[1], *P = buf[2], *M = buf[3];
  hipLaunchKernelGGL(k_ns_init, g1, dim3(256), 0, s, C, d, 1.0 / c, Y, Z);
  SH_LAUNCH_CHECK("k_ns_init");
  const double stop = tol * std::sqrt((double)d);
  double r_prev = INFINITY, r = INFINITY;
  int it = 0;
  for (; it < max_iter; ++it) {
    rc = sh_dgemm(0, d, d, d, 1.0, Z, d, Y, d, 0.0, P, d, stream);
    if (rc != SH_OK) return rc;
    double r2 = 0.0;
    rc = sumsq(P, 1, &r2);
    if (rc != SH_OK) return rc;
    r = std::sqrt(r2);
    if (!std::isfinite(r)) {
      set_error("sh_spd_sqrt: the iteration diverged (C not positive definite?)");
      return SH_EINVAL;
    }
    if (r <= stop || (r_prev < 1e-6 && r >= r_prev)) break;
    r_prev = r;
    hipLaunchKernelGGL(k_ns_mid, g1, dim3(256), 0, s, P, d, M);
    SH_LAUNCH_CHECK("k_ns_mid");
    // Y <- Y M into P (free now), Z <- M Z into Y's old buffer
    rc = sh_dgemm(0, d, d, d, 1.0, Y, d, M, d, 0.0, P, d, stream);
    if (rc != SH_OK) return rc;
    rc = sh_dgemm(0, d, d, d, 1.0, M, d, Z, d, 0.0, Y, d, stream);
    if (rc != SH_OK) return rc;
    double *ny = P, *nz = Y, *np_ = Z;
    Y = ny;
    Z = nz;
    P = np_;
  }
  *iters = it;
  *resid = r;
  hipLaunchKernelGGL(k_scale2, g1, dim3(256), 0, s, Y, Z, d2, std::sqrt(c), 1.0 / std::sqrt(c),
                     sqrt_out, inv_sqrt_out);
  SH_LAUNCH_CHECK("k_scale2");
  return SH_OK;
}
