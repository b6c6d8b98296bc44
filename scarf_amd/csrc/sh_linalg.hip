// Top eigenpairs of the PCA merge matrix on the GPU: block Lanczos (block size 8) with
// full classical Gram-Schmidt reorthogonalisation done twice, and CholeskyQR2 for each
// new block.  The host only sees the small block-tridiagonal matrix (Ritz values and
// vectors of at most a few thousand rows) and decides when the wanted Ritz pairs have
// converged; every d-sized operation runs here.
//
// Storage: M is d x d row-major and symmetric; the Krylov basis V is [kmax][d] (each
// basis vector contiguous), block j = rows [8j, 8j + 8).  Per block step j:
//   W  = M V_j                       (k_symm_part + k_sum_parts: split-K, fixed order)
//   H  = V_{0..j}^T W, W -= V H      (twice: k_cgs_proj, k_cgs_update_part, k_sum_parts;
//                                      each pass subtracts its own coefficients)
//   A_j = sum over both passes of H's block-j rows (diagonal block of the Lanczos matrix)
//   V_{j+1} R_j = W                  (k_cholqr2, one workgroup; R_j below the diagonal)
// All reductions have a fixed order, so a run is bit-reproducible.
#include <algorithm>

#include "sh_common.h"

namespace sh {

typedef double doublex4 __attribute__((ext_vector_type(4)));
constexpr int kLP = 8;         // Lanczos block size
constexpr int kSymSeg = 512;   // max rows of M per split-K segment (LDS: 8 x 512 doubles)
constexpr int kCgsRows = 16;   // basis vectors per workgroup in k_cgs_proj (4 per wave)

// ---- FP64 GEMM on MFMA ------------------------------------------------------------
// C[m x n] = alpha * op(A) * B + beta * C, op(A) = A [m x k] or A^T (A stored k x m), all
// row-major.  64 x 64 output tile per workgroup (4 waves in 2 x 2, each 32 x 32 = 2 x 2
// v_mfma_f64_16x16x4f64 tiles), K staged through LDS 16 at a time.
constexpr int kDT = 64;
constexpr int kDK = 16;

__global__ void __launch_bounds__(256) k_dgemm(int transA, int64_t m, int64_t n, int64_t k,
                                               const double *__restrict__ A, int64_t lda,
                                               const double *__restrict__ B, int64_t ldb,
                                               double alpha, double beta, double *__restrict__ C,
                                               int64_t ldc) {
  __shared__ double As[kDK][kDT + 1];
  __shared__ double Bs[kDK][kDT + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int64_t m0 = (int64_t)blockIdx.y * kDT, n0 = (int64_t)blockIdx.x * kDT;
  doublex4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = doublex4{0.0, 0.0, 0.0, 0.0};
  for (int64_t k0 = 0; k0 < k; k0 += kDK) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int e = tid + 256 * p;
      // op(A) tile: for A^T the 64 consecutive threads walk a row of A (coalesced);
      // for A they walk 16 consecutive k of one row
      int kk, i;
      if (transA) {
        kk = e >> 6;
        i = e & 63;
      } else {
        kk = e & 15;
        i = e >> 4;
      }
      const int64_t gk = k0 + kk, gi = m0 + i;
      double av = 0.0;
      if (gk < k && gi < m) av = transA ? A[gk * lda + gi] : A[gi * lda + gk];
      As[kk][i] = av;
      const int bk = e >> 6, bj = e & 63;
      const int64_t gbk = k0 + bk, gj = n0 + bj;
      Bs[bk][bj] = (gbk < k && gj < n) ? B[gbk * ldb + gj] : 0.0;
    }
    __syncthreads();
    const int kr = lane >> 4, cl = lane & 15;
#pragma unroll
    for (int kk = 0; kk < kDK; kk += 4) {
      double av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[kk + kr][wm * 32 + a * 16 + cl];
#pragma unroll
      for (int c = 0; c < 2; ++c) bv[c] = Bs[kk + kr][wn * 32 + c * 16 + cl];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
          acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout of v_mfma_f64_16x16x4f64: lane l, value i -> row 4*i + l/16, column l%16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = m0 + wm * 32 + a * 16 + 4 * i + (lane >> 4);
        const int64_t col = n0 + wn * 32 + c * 16 + (lane & 15);
        if (row < m && col < n) {
          double v = alpha * acc[a][c][i];
          if (beta != 0.0) v += beta * C[row * ldc + col];
          C[row * ldc + col] = v;
        }
      }
}

// ---- merge matrix of one partial fit ------------------------------------------------
// M = (G - n bm bm^T) + Vt^T diag(s2) Vt + mc mc^T, every term formed so that M[i][j] and
// M[j][i] round identically (G is stored symmetric).
__global__ void k_merge_matrix(const double *__restrict__ G, int32_t d, double n,
                               const double *__restrict__ bm, const double *__restrict__ Vt,
                               const double *__restrict__ s2, int32_t nc,
                               const double *__restrict__ mc, double *__restrict__ M) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)d * d) return;
  const int i = (int)(e / d), j = (int)(e % d);
  double v = G[e] - n * (bm[i] * bm[j]);
  if (Vt) {
    double r = 0.0;
    for (int c = 0; c < nc; ++c) r += s2[c] * (Vt[(int64_t)c * d + i] * Vt[(int64_t)c * d + j]);
    v += r;
  }
  if (mc) v += mc[i] * mc[j];
  M[e] = v;
}

// ---- block Lanczos pieces -------------------------------------------------------------
// part[s][q][i] = sum_{j in segment s} M[j][i] V[q][j]   (M symmetric: columns coalesced)
__global__ void __launch_bounds__(256) k_symm_part(const double *__restrict__ M, int32_t d,
                                                   const double *__restrict__ V, int32_t seg,
                                                   double *__restrict__ part) {
  __shared__ double vs[kLP][kSymSeg];
  const int s = blockIdx.y;
  const int j0 = s * seg, j1 = min(d, j0 + seg);
  for (int e = threadIdx.x; e < kLP * seg; e += 256) {
    const int q = e / seg, jj = e % seg;
    vs[q][jj] = (j0 + jj < d) ? V[(int64_t)q * d + j0 + jj] : 0.0;
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= d) return;
  double acc[kLP];
#pragma unroll
  for (int q = 0; q < kLP; ++q) acc[q] = 0.0;
  for (int j = j0; j < j1; ++j) {
    const double mv = M[(int64_t)j * d + i];
#pragma unroll
    for (int q = 0; q < kLP; ++q) acc[q] += mv * vs[q][j - j0];
  }
#pragma unroll
  for (int q = 0; q < kLP; ++q) part[((int64_t)s * kLP + q) * d + i] = acc[q];
}

// out[e] = sign * sum_s part[s][e] (+ out[e] when accumulate), s in order
__global__ void k_sum_parts(const double *__restrict__ part, int32_t S, int64_t n, double sign,
                            int accumulate, double *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int p = 0; p < S; ++p) s += part[(int64_t)p * n + e];
  out[e] = accumulate ? out[e] + sign * s : sign * s;
}

// H[c][q] = V[c] . W[q] for c < k (this pass's coefficients; 4 basis vectors per wave).
// Rows c of block jb (the newest) are also added into T_blk (the diagonal block of the
// Lanczos matrix, summed over both passes; accumulate = 0 on the first pass).
__global__ void __launch_bounds__(256) k_cgs_proj(const double *__restrict__ V, int32_t k,
                                                  int32_t d, const double *__restrict__ W,
                                                  double *__restrict__ H, int32_t jb,
                                                  int accumulate, double *__restrict__ T_blk) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c0 = blockIdx.x * kCgsRows + w * 4;
  if (c0 >= k) return;
  double acc[4][kLP];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int q = 0; q < kLP; ++q) acc[a][q] = 0.0;
  for (int i = lane; i < d; i += 64) {
    double wv[kLP];
#pragma unroll
    for (int q = 0; q < kLP; ++q) wv[q] = W[(int64_t)q * d + i];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double v = (c0 + a < k) ? V[(int64_t)(c0 + a) * d + i] : 0.0;
#pragma unroll
      for (int q = 0; q < kLP; ++q) acc[a][q] += v * wv[q];
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int q = 0; q < kLP; ++q) {
      const double s = wave_sum_f64(acc[a][q]);
      const int c = c0 + a;
      if (lane == 0 && c < k) {
        H[(int64_t)c * kLP + q] = s;
        if (c / kLP == jb) {
          double *t = T_blk + (c - jb * kLP) * kLP + q;
          *t = accumulate ? *t + s : s;
        }
      }
    }
}

// part[s][q][i] = sum_{c in segment s} V[c][i] H[c][q]
__global__ void __launch_bounds__(256) k_cgs_update_part(const double *__restrict__ V, int32_t k,
                                                         int32_t d, const double *__restrict__ H,
                                                         int32_t seg,
                                                         double *__restrict__ part) {
  __shared__ double hs[kSymSeg][kLP];
  const int s = blockIdx.y;
  const int c0 = s * seg, c1 = min(k, c0 + seg);
  for (int e = threadIdx.x; e < (c1 - c0) * kLP; e += 256) hs[e / kLP][e % kLP] = H[(int64_t)c0 * kLP + e];
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= d) return;
  double acc[kLP];
#pragma unroll
  for (int q = 0; q < kLP; ++q) acc[q] = 0.0;
  for (int c = c0; c < c1; ++c) {
    const double v = V[(int64_t)c * d + i];
#pragma unroll
    for (int q = 0; q < kLP; ++q) acc[q] += v * hs[c - c0][q];
  }
#pragma unroll
  for (int q = 0; q < kLP; ++q) part[((int64_t)s * kLP + q) * d + i] = acc[q];
}

// Upper Cholesky factor R (S = R^T R) of an 8 x 8 SPD matrix and its inverse.
// Returns false when a pivot is not safely positive (condition of W beyond ~1e7).
__device__ bool chol8(const double S[kLP][kLP], double R[kLP][kLP], double Ri[kLP][kLP]) {
  double smax = 0.0;
  for (int q = 0; q < kLP; ++q) smax = fmax(smax, S[q][q]);
  if (!(smax > 0.0)) return false;
  for (int r = 0; r < kLP; ++r)
    for (int c = 0; c < kLP; ++c) R[r][c] = 0.0;
  for (int r = 0; r < kLP; ++r) {
    double dg = S[r][r];
    for (int t = 0; t < r; ++t) dg -= R[t][r] * R[t][r];
    if (!(dg > 1e-14 * smax)) return false;
    const double p = sqrt(dg);
    R[r][r] = p;
    for (int c = r + 1; c < kLP; ++c) {
      double v = S[r][c];
      for (int t = 0; t < r; ++t) v -= R[t][r] * R[t][c];
      R[r][c] = v / p;
    }
  }
  // inverse of an upper-triangular matrix, column by column
  for (int r = 0; r < kLP; ++r)
    for (int c = 0; c < kLP; ++c) Ri[r][c] = 0.0;
  for (int c = 0; c < kLP; ++c) {
    Ri[c][c] = 1.0 / R[c][c];
    for (int r = c - 1; r >= 0; --r) {
      double v = 0.0;
      for (int t = r + 1; t <= c; ++t) v += R[r][t] * Ri[t][c];
      Ri[r][c] = -v / R[r][r];
    }
  }
  return true;
}

// CholeskyQR2 of W [8][d] in one workgroup: Vout R = W with Vout orthonormal rows.
// Each thread owns rows i = tid, tid + 256, ... and only touches those, so the only
// cross-thread data are the 8 x 8 Gram and factor (in LDS).
__global__ void __launch_bounds__(256) k_cholqr2(const double *__restrict__ W, int32_t d,
                                                 double *__restrict__ Vout,
                                                 double *__restrict__ R_out,
                                                 int32_t *__restrict__ status) {
  __shared__ double red[4][kLP * (kLP + 1) / 2];
  __shared__ double S[kLP][kLP], R[kLP][kLP], Ri[kLP][kLP], Rt[kLP][kLP];
  __shared__ int ok;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int pass = 0; pass < 2; ++pass) {
    const double *src = pass == 0 ? W : Vout;
    double acc[kLP * (kLP + 1) / 2];
#pragma unroll
    for (int e = 0; e < kLP * (kLP + 1) / 2; ++e) acc[e] = 0.0;
    for (int i = tid; i < d; i += 256) {
      double x[kLP];
#pragma unroll
      for (int q = 0; q < kLP; ++q) x[q] = src[(int64_t)q * d + i];
      int e = 0;
#pragma unroll
      for (int a = 0; a < kLP; ++a)
#pragma unroll
        for (int b = a; b < kLP; ++b) acc[e++] += x[a] * x[b];
    }
#pragma unroll
    for (int e = 0; e < kLP * (kLP + 1) / 2; ++e) {
      const double s = wave_sum_f64(acc[e]);
      if (lane == 0) red[w][e] = s;
    }
    __syncthreads();
    if (tid == 0) {
      int e = 0;
      for (int a = 0; a < kLP; ++a)
        for (int b = a; b < kLP; ++b, ++e) {
          const double v = red[0][e] + red[1][e] + red[2][e] + red[3][e];
          S[a][b] = v;
          S[b][a] = v;
        }
      ok = chol8(S, R, Ri) ? 1 : 0;
      if (ok) {
        if (pass == 0) {
          for (int a = 0; a < kLP; ++a)
            for (int b = 0; b < kLP; ++b) Rt[a][b] = R[a][b];
        } else {  // total factor R2 * R1 (both upper triangular)
          double P[kLP][kLP];
          for (int a = 0; a < kLP; ++a)
            for (int b = 0; b < kLP; ++b) {
              double v = 0.0;
              for (int t = a; t <= b; ++t) v += R[a][t] * Rt[t][b];
              P[a][b] = v;
            }
          for (int a = 0; a < kLP; ++a)
            for (int b = 0; b < kLP; ++b) Rt[a][b] = P[a][b];
        }
      }
    }
    __syncthreads();
    if (!ok) {
      if (tid == 0) *status = 1;
      return;
    }
    // rows of this thread: x <- x R^{-1}
    for (int i = tid; i < d; i += 256) {
      double x[kLP], y[kLP];
#pragma unroll
      for (int q = 0; q < kLP; ++q) x[q] = src[(int64_t)q * d + i];
#pragma unroll
      for (int q = 0; q < kLP; ++q) {
        double v = 0.0;
#pragma unroll
        for (int r = 0; r <= q; ++r) v += x[r] * Ri[r][q];
        y[q] = v;
      }
#pragma unroll
      for (int q = 0; q < kLP; ++q) Vout[(int64_t)q * d + i] = y[q];
    }
    __syncthreads();
  }
  if (tid < kLP * kLP) R_out[tid] = Rt[tid / kLP][tid % kLP];
}

struct LanczosPlan {
  int32_t sym_seg, sym_S;  // split-K of M V
  int32_t cgs_seg, cgs_S;  // split of the basis in the update
  size_t part_elems;       // doubles of the split-K partials
};

static LanczosPlan lanczos_plan(int32_t d, int32_t kmax) {
  LanczosPlan p;
  const int64_t bx = cdiv(d, 256);
  // about 512 workgroups for M V, segments of at most kSymSeg rows
  int64_t S = std::max<int64_t>(cdiv(d, kSymSeg), std::min<int64_t>(64, cdiv(512, bx)));
  p.sym_seg = (int32_t)cdiv(d, S);
  p.sym_S = (int32_t)cdiv(d, p.sym_seg);
  int64_t Sc = std::max<int64_t>(cdiv(kmax, kSymSeg), std::min<int64_t>(64, cdiv(256, bx)));
  p.cgs_seg = (int32_t)std::max<int64_t>(1, cdiv(kmax, Sc));
  p.cgs_S = (int32_t)cdiv(kmax, p.cgs_seg);
  p.part_elems = (size_t)std::max(p.sym_S, p.cgs_S) * kLP * d;
  return p;
}

}  // namespace sh

using namespace sh;

extern "C" int sh_dgemm(int transA, int64_t m, int64_t n, int64_t k, double alpha,
                        const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
                        double *C, int64_t ldc, sh_stream_t stream) {
  SH_ARG(m >= 0 && n >= 0 && k >= 0, "sh_dgemm: negative size");
  SH_ARG(A && B && C, "sh_dgemm: NULL pointer");
  SH_ARG(lda >= (transA ? m : k) && ldb >= n && ldc >= n,
         "sh_dgemm: leading dimension too small");
  SH_ARG(cdiv(n, kDT) <= 0x7fffffff && cdiv(m, kDT) <= 65535, "sh_dgemm: matrix too large");
  if (m == 0 || n == 0) return SH_OK;
  dim3 grid((uint32_t)cdiv(n, kDT), (uint32_t)cdiv(m, kDT));
  hipLaunchKernelGGL(k_dgemm, grid, dim3(256), 0, (hipStream_t)stream, transA, m, n, k, A, lda,
                     B, ldb, alpha, beta, C, ldc);
  SH_LAUNCH_CHECK("k_dgemm");
  return SH_OK;
}

extern "C" int sh_merge_matrix(const double *G, int32_t d, double n, const double *bm,
                               const double *Vt, const double *s2, int32_t nc, const double *mc,
                               double *M, sh_stream_t stream) {
  SH_ARG(d > 0 && G && bm && M, "sh_merge_matrix: bad arguments");
  SH_ARG(!Vt || (s2 && nc > 0), "sh_merge_matrix: Vt needs s2 and nc > 0");
  const int64_t n2 = (int64_t)d * d;
  hipLaunchKernelGGL(k_merge_matrix, dim3((uint32_t)cdiv(n2, 256)), dim3(256), 0,
                     (hipStream_t)stream, G, d, n, bm, Vt, s2, nc, mc, M);
  SH_LAUNCH_CHECK("k_merge_matrix");
  return SH_OK;
}

extern "C" size_t sh_lanczos_workspace(int32_t d, int32_t kmax) {
  if (d <= 0 || kmax <= 0) return 0;
  LanczosPlan p = lanczos_plan(d, kmax);
  // W [8][d], H [kmax][8], partials
  return align_up(((size_t)kLP * d + (size_t)kmax * kLP + p.part_elems) * sizeof(double), 256);
}

extern "C" int sh_lanczos_steps(const double *M, int32_t d, double *V, int32_t kmax,
                                int32_t j0, int32_t n_steps, double *T_diag, double *T_sub,
                                int32_t *status, void *ws, size_t ws_bytes, sh_stream_t stream) {
  SH_ARG(M && V && T_diag && T_sub && status, "sh_lanczos_steps: NULL pointer");
  SH_ARG(d >= kLP && kmax >= kLP && kmax % kLP == 0, "sh_lanczos_steps: bad d/kmax");
  SH_ARG(j0 >= 0 && n_steps >= 0 && (int64_t)(j0 + n_steps + 1) * kLP <= kmax,
         "sh_lanczos_steps: steps exceed the basis capacity kmax");
  SH_ARG(ws_bytes >= sh_lanczos_workspace(d, kmax), "sh_lanczos_steps: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  LanczosPlan p = lanczos_plan(d, kmax);
  double *W = (double *)ws;
  double *H = W + (size_t)kLP * d;
  double *part = H + (size_t)kmax * kLP;
  const int64_t nW = (int64_t)kLP * d;
  const dim3 bx((uint32_t)cdiv(d, 256));
  for (int32_t j = j0; j < j0 + n_steps; ++j) {
    const int32_t k = (j + 1) * kLP;  // basis vectors so far
    const double *Vj = V + (int64_t)j * kLP * d;
    hipLaunchKernelGGL(k_symm_part, dim3(bx.x, p.sym_S), dim3(256), 0, st, M, d, Vj, p.sym_seg,
                       part);
    hipLaunchKernelGGL(k_sum_parts, dim3((uint32_t)cdiv(nW, 256)), dim3(256), 0, st, part,
                       p.sym_S, nW, 1.0, 0, W);
    // two passes of classical Gram-Schmidt; each subtracts only its own coefficients
    for (int pass = 0; pass < 2; ++pass) {
      hipLaunchKernelGGL(k_cgs_proj, dim3((uint32_t)cdiv(k, kCgsRows)), dim3(256), 0, st, V, k,
                         d, W, H, j, pass, T_diag + (int64_t)j * kLP * kLP);
      const int32_t seg = p.cgs_seg;
      const int32_t S = (int32_t)cdiv(k, seg);
      hipLaunchKernelGGL(k_cgs_update_part, dim3(bx.x, S), dim3(256), 0, st, V, k, d, H, seg,
                         part);
      hipLaunchKernelGGL(k_sum_parts, dim3((uint32_t)cdiv(nW, 256)), dim3(256), 0, st, part, S,
                         nW, -1.0, 1, W);
    }
    hipLaunchKernelGGL(k_cholqr2, dim3(1), dim3(256), 0, st, W, d, V + (int64_t)k * d,
                       T_sub + (int64_t)j * kLP * kLP, status);
  }
  SH_LAUNCH_CHECK("sh_lanczos_steps");
  return SH_OK;
}
