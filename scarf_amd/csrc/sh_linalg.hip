// Top eigenpairs of the PCA merge matrix on the GPU: block Lanczos (block size 8) with
// full classical Gram-Schmidt reorthogonalisation done twice, and CholeskyQR2 for each
// new block.  The host only sees the small block-tridiagonal matrix (Ritz values and
// vectors of at most a few thousand rows) and decides when the wanted Ritz pairs have
// converged; every d-sized operation runs here.
//
// Storage: M is d x d row-major and symmetric; the Krylov basis V is [kmax][d] (each
// basis vector contiguous), block j = rows [8j, 8j + 8).  Per block step j:
//   W  = M V_j                       (k_symm_rows: one wave per 4 rows of M)
//   H  = V_{0..j}^T W, W -= V H      (twice: k_cgs_proj, k_cgs_update; each pass
//                                      subtracts its own coefficients)
//   A_j = sum over both passes of H's block-j rows (diagonal block of the Lanczos matrix)
//   V_{j+1} R_j = W                  (k_cholqr2, one workgroup; R_j below the diagonal)
// All reductions have a fixed order, so a run is bit-reproducible.
#include <algorithm>

#include "sh_common.h"

namespace sh {

typedef double doublex4 __attribute__((ext_vector_type(4)));
constexpr int kLP = 8;         // Lanczos block size

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
// W[q][i] = sum_j M[i][j] V[q][j]: one wave per row of M (read once, coalesced), 16 rows
// per workgroup.  The block V (8 x d) passes through LDS 1024 columns at a time (64 KiB),
// so each workgroup reads it once instead of once per row; each lane has its 16 loads of
// the row's chunk in flight at once.
constexpr int kSymWaves = 16, kSymCols = 1024;
__global__ void __launch_bounds__(kSymWaves * 64) k_symm_rows(const double *__restrict__ M,
                                                              int32_t d,
                                                              const double *__restrict__ V,
                                                              double *__restrict__ W) {
  __shared__ double sV[kLP][kSymCols];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * kSymWaves + w;
  const bool row_ok = i < d;
  const double *mrow = M + (row_ok ? i : 0) * d;
  double acc[kLP];
#pragma unroll
  for (int q = 0; q < kLP; ++q) acc[q] = 0.0;
  for (int c0 = 0; c0 < d; c0 += kSymCols) {
    __syncthreads();
    for (int e = threadIdx.x; e < kLP * kSymCols; e += kSymWaves * 64) {
      const int q = e / kSymCols, c = e % kSymCols;
      sV[q][c] = c0 + c < d ? V[(int64_t)q * d + c0 + c] : 0.0;
    }
    __syncthreads();
    double m[kSymCols / 64];
#pragma unroll
    for (int t = 0; t < kSymCols / 64; ++t) {
      const int c = c0 + lane + 64 * t;
      m[t] = row_ok && c < d ? mrow[c] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < kSymCols / 64; ++t)
#pragma unroll
      for (int q = 0; q < kLP; ++q) acc[q] += m[t] * sV[q][lane + 64 * t];
  }
#pragma unroll
  for (int q = 0; q < kLP; ++q) {
    const double sw = wave_sum_f64_dpp(acc[q]);
    if (lane == 0 && row_ok) W[(int64_t)q * d + i] = sw;
  }
}

// H[c][q] = V[c] . W[q] for c < k (this pass's coefficients; one workgroup per basis
// vector, its 4 waves over interleaved quarters of the features, summed in a fixed order).
// Rows c of block jb (the newest) are also added into T_blk (the diagonal block of the
// Lanczos matrix, summed over both passes; accumulate = 0 on the first pass).
__global__ void __launch_bounds__(256) k_cgs_proj(const double *__restrict__ V, int32_t k,
                                                  int32_t d, const double *__restrict__ W,
                                                  double *__restrict__ H, int32_t jb,
                                                  int accumulate, double *__restrict__ T_blk) {
  __shared__ double red[4][kLP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = blockIdx.x;
  double acc[kLP];
#pragma unroll
  for (int q = 0; q < kLP; ++q) acc[q] = 0.0;
  const double *vc = V + (int64_t)c * d;
#pragma unroll 2
  for (int i = threadIdx.x; i < d; i += 256) {
    const double v = vc[i];
#pragma unroll
    for (int q = 0; q < kLP; ++q) acc[q] += v * W[(int64_t)q * d + i];
  }
#pragma unroll
  for (int q = 0; q < kLP; ++q) {
    const double sw = wave_sum_f64_dpp(acc[q]);
    if (lane == 0) red[w][q] = sw;
  }
  __syncthreads();
  if (threadIdx.x < kLP) {
    const int q = threadIdx.x;
    const double sum = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
    H[(int64_t)c * kLP + q] = sum;
    if (c / kLP == jb) {
      double *t = T_blk + (c - jb * kLP) * kLP + q;
      *t = accumulate ? *t + sum : sum;
    }
  }
}

// W[q][i] -= sum_{c < k} V[c][i] H[c][q]: a workgroup owns 64 features; its 16 waves take
// every 16th basis vector and their partial sums are added in a fixed order (deterministic).
constexpr int kUpdWaves = 16;
__global__ void __launch_bounds__(kUpdWaves * 64) k_cgs_update(const double *__restrict__ V,
                                                               int32_t k, int32_t d,
                                                               const double *__restrict__ H,
                                                               double *__restrict__ W) {
  __shared__ double part[kUpdWaves][kLP][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  double acc[kLP];
#pragma unroll
  for (int q = 0; q < kLP; ++q) acc[q] = 0.0;
  if (i < d) {
#pragma unroll 2
    for (int c = w; c < k; c += kUpdWaves) {
      const double v = V[(int64_t)c * d + i];
#pragma unroll
      for (int q = 0; q < kLP; ++q) acc[q] += v * H[(int64_t)c * kLP + q];
    }
  }
#pragma unroll
  for (int q = 0; q < kLP; ++q) part[w][q][lane] = acc[q];
  __syncthreads();
  // threads 0..511 finish the 8 x 64 outputs
  if (threadIdx.x < kLP * 64) {
    const int q = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int ii = blockIdx.x * 64 + l;
    if (ii < d) {
      double sum = part[0][q][l];
#pragma unroll
      for (int ww = 1; ww < kUpdWaves; ++ww) sum += part[ww][q][l];
      W[(int64_t)q * d + ii] -= sum;
    }
  }
}

// In-place upper Cholesky factor of an 8 x 8 SPD matrix held in registers (A = S on entry,
// R with S = R^T R on exit; the strict lower part is left as is).  False when a pivot is
// not safely positive (condition of W beyond ~1e7).
__device__ __forceinline__ bool chol8_inplace(double (&A)[kLP][kLP]) {
  double smax = 0.0;
#pragma unroll
  for (int q = 0; q < kLP; ++q) smax = fmax(smax, A[q][q]);
  bool ok = smax > 0.0;
#pragma unroll
  for (int r = 0; r < kLP; ++r) {
    double dg = A[r][r];
#pragma unroll
    for (int t = 0; t < r; ++t) dg -= A[t][r] * A[t][r];
    ok = ok && dg > 1e-14 * smax;
    const double p = sqrt(fmax(dg, 1e-300));
    const double ip = 1.0 / p;
    A[r][r] = p;
#pragma unroll
    for (int c = r + 1; c < kLP; ++c) {
      double v = A[r][c];
#pragma unroll
      for (int t = 0; t < r; ++t) v -= A[t][r] * A[t][c];
      A[r][c] = v * ip;
    }
  }
  return ok;
}

// One CholeskyQR factorisation from the per-wave partial Grams in red (all threads call
// it): the Gram summed over waves in a fixed order (36 threads), its Cholesky factor R
// (thread 0) into Rc, and R^{-1} into Ri (8 threads, one column each).  ok_s is 0 when a
// pivot is not safely positive.  Inlined: as a call, the caller's live rows were saved
// and restored through scratch memory around it.
template <int NW>
__device__ __forceinline__ void qr_factor(const double (*red)[kLP * (kLP + 1) / 2],
                                          double *gsum, double (*Rc)[kLP], double (*Ri)[kLP],
                                          int *ok_s) {
  constexpr int NE = kLP * (kLP + 1) / 2;
  const int tid = threadIdx.x;
  if (tid < NE) {
    double v = red[0][tid];
#pragma unroll
    for (int ww = 1; ww < NW; ++ww) v += red[ww][tid];
    gsum[tid] = v;
  }
  __syncthreads();
  if (tid == 0) {
    double A[kLP][kLP];
    int e = 0;
#pragma unroll
    for (int a = 0; a < kLP; ++a)
#pragma unroll
      for (int b = a; b < kLP; ++b, ++e) {
        A[a][b] = gsum[e];
        A[b][a] = gsum[e];
      }
    *ok_s = chol8_inplace(A) ? 1 : 0;
#pragma unroll
    for (int a = 0; a < kLP; ++a)
#pragma unroll
      for (int b = 0; b < kLP; ++b) Rc[a][b] = b >= a ? A[a][b] : 0.0;
  }
  __syncthreads();
  if (tid < kLP) {  // column c of R^{-1} by back substitution
    const int c = tid;
    double col[kLP];
#pragma unroll
    for (int r = 0; r < kLP; ++r) col[r] = 0.0;
    col[c] = 1.0 / Rc[c][c];
#pragma unroll
    for (int r = kLP - 2; r >= 0; --r) {
      if (r < c) {
        double v = 0.0;
#pragma unroll
        for (int t = 1; t < kLP; ++t)
          if (t > r && t <= c) v += Rc[r][t] * col[t];
        col[r] = -v / Rc[r][r];
      }
    }
#pragma unroll
    for (int r = 0; r < kLP; ++r) Ri[r][c] = col[r];
  }
  __syncthreads();
}

// CholeskyQR2 of W [8][d] in one workgroup: Vout R = W with Vout orthonormal rows and
// R = R2 R1.  Thread t owns rows t, t + 512, ... (at most kQrRows of them, d <= 2048):
// they are loaded once, kept in registers through both passes and stored once; the only
// cross-thread data are the 8 x 8 Gram (per-wave partials in LDS) and the factors (LDS).
constexpr int kQrThreads = 512, kQrWaves = kQrThreads / 64, kQrRows = 4;
constexpr int kQrMaxD = kQrThreads * kQrRows;
__global__ void __launch_bounds__(kQrThreads) k_cholqr2(const double *__restrict__ W, int32_t d,
                                                        double *__restrict__ Vout,
                                                        double *__restrict__ R_out,
                                                        int32_t *__restrict__ status) {
  constexpr int NE = kLP * (kLP + 1) / 2;
  __shared__ double red[kQrWaves][NE];
  __shared__ double gsum[NE];
  __shared__ double Rc_s[2][kLP][kLP], Ri_s[kLP][kLP];
  __shared__ int ok_s;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double x[kQrRows][kLP];
#pragma unroll
  for (int r = 0; r < kQrRows; ++r) {
    const int i = tid + r * kQrThreads;
#pragma unroll
    for (int q = 0; q < kLP; ++q) x[r][q] = i < d ? W[(int64_t)q * d + i] : 0.0;
  }
  for (int pass = 0; pass < 2; ++pass) {
    double acc[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) acc[e] = 0.0;
#pragma unroll
    for (int r = 0; r < kQrRows; ++r) {
      int e = 0;
#pragma unroll
      for (int a = 0; a < kLP; ++a)
#pragma unroll
        for (int b = a; b < kLP; ++b) acc[e++] += x[r][a] * x[r][b];
    }
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const double sw = wave_sum_f64_dpp(acc[e]);
      if (lane == 0) red[w][e] = sw;
    }
    __syncthreads();
    qr_factor<kQrWaves>(red, gsum, Rc_s[pass], Ri_s, &ok_s);
    if (!ok_s) {
      if (tid == 0) *status = 1;
      return;
    }
    // x <- x R^{-1} (rows past d are zero and stay zero)
#pragma unroll
    for (int r = 0; r < kQrRows; ++r) {
      double y[kLP];
#pragma unroll
      for (int q = 0; q < kLP; ++q) {
        double v = 0.0;
#pragma unroll
        for (int t = 0; t <= q; ++t) v += x[r][t] * Ri_s[t][q];
        y[q] = v;
      }
#pragma unroll
      for (int q = 0; q < kLP; ++q) x[r][q] = y[q];
    }
    __syncthreads();  // Ri_s and red are rewritten by the next pass
  }
#pragma unroll
  for (int r = 0; r < kQrRows; ++r) {
    const int i = tid + r * kQrThreads;
    if (i < d) {
#pragma unroll
      for (int q = 0; q < kLP; ++q) Vout[(int64_t)q * d + i] = x[r][q];
    }
  }
  if (tid < kLP * kLP) {  // R = R2 R1 (upper triangular)
    const int a = tid / kLP, b = tid % kLP;
    double v = 0.0;
    for (int t = a; t <= b; ++t) v += Rc_s[1][a][t] * Rc_s[0][t][b];
    R_out[tid] = v;
  }
}

// The same for d > 2048: rows streamed from global memory in both passes.
__global__ void __launch_bounds__(kQrThreads) k_cholqr2_wide(const double *__restrict__ W,
                                                             int32_t d,
                                                             double *__restrict__ Vout,
                                                             double *__restrict__ R_out,
                                                             int32_t *__restrict__ status) {
  constexpr int NE = kLP * (kLP + 1) / 2;
  __shared__ double red[kQrWaves][NE];
  __shared__ double gsum[NE];
  __shared__ double Rc_s[2][kLP][kLP], Ri_s[kLP][kLP];
  __shared__ int ok_s;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int pass = 0; pass < 2; ++pass) {
    const double *src = pass == 0 ? W : Vout;
    double acc[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) acc[e] = 0.0;
    for (int i = tid; i < d; i += kQrThreads) {
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
    for (int e = 0; e < NE; ++e) {
      const double sw = wave_sum_f64_dpp(acc[e]);
      if (lane == 0) red[w][e] = sw;
    }
    __syncthreads();
    qr_factor<kQrWaves>(red, gsum, Rc_s[pass], Ri_s, &ok_s);
    if (!ok_s) {
      if (tid == 0) *status = 1;
      return;
    }
    for (int i = tid; i < d; i += kQrThreads) {
      double x[kLP], y[kLP];
#pragma unroll
      for (int q = 0; q < kLP; ++q) x[q] = src[(int64_t)q * d + i];
#pragma unroll
      for (int q = 0; q < kLP; ++q) {
        double v = 0.0;
#pragma unroll
        for (int t = 0; t <= q; ++t) v += x[t] * Ri_s[t][q];
        y[q] = v;
      }
#pragma unroll
      for (int q = 0; q < kLP; ++q) Vout[(int64_t)q * d + i] = y[q];
    }
    __syncthreads();
  }
  if (tid < kLP * kLP) {
    const int a = tid / kLP, b = tid % kLP;
    double v = 0.0;
    for (int t = a; t <= b; ++t) v += Rc_s[1][a][t] * Rc_s[0][t][b];
    R_out[tid] = v;
  }
}

// ---- block subspace iteration pieces --------------------------------------------------
// A block of b vectors is stored as the columns of Y (d x ldy row-major, ldy a multiple of
// 16, columns past b zero).  One Chebyshev filter step is Y2 = alpha M Y1 + beta Y1 +
// gamma Y0.  The product is split over slices of the k axis so that a d ~ 2000 matrix
// still fills the GPU: k_symm_block_part computes one 64-row x (16 NT)-column partial per
// (row tile, k slice, column group) on v_mfma_f64_16x16x4f64, and k_symm_block_sum adds
// the slices in slice order (bit-reproducible) and applies the three-term update.
// M is symmetric, so the A operand M[row][k] is read as M[k][row]: each lane quarter
// reads 16 consecutive doubles of one row of M.  The k-slice of Y1 passes through LDS 32
// rows at a time, double-buffered, shared by the workgroup's four waves (16 rows each).
constexpr int kSbRows = 64;
constexpr int kSbK = 32;
constexpr int kSbMaxNT = 8;  // 16-column tiles per column group (128 columns)

template <int NT>
__global__ void __launch_bounds__(256) k_symm_block_part(
    const double *__restrict__ M, int32_t d, const double *__restrict__ Y, int64_t ldy,
    int32_t kslice, double *__restrict__ part /* [slices][d][ldy] */) {
  __shared__ double ys[2][kSbK][16 * NT];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int kr = lane >> 4, cl = lane & 15;
  const int64_t row = (int64_t)blockIdx.x * kSbRows + 16 * w + cl;  // this lane's A row
  const int64_t kb = (int64_t)blockIdx.y * kslice;
  const int64_t ke = min<int64_t>(d, kb + kslice);
  const int c0 = blockIdx.z * 16 * NT;
  constexpr int CW = 16 * NT;           // columns of the group
  constexpr int PAIRS = kSbK * CW / 2;  // double2 per stage
  auto stage = [&](int buf, int64_t k0) {
    for (int e = tid; e < PAIRS; e += 256) {
      const int kk = e / (CW / 2), c2 = 2 * (e % (CW / 2));
      const int64_t k = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (k < ke) v = *reinterpret_cast<const double2 *>(Y + k * ldy + c0 + c2);
      *reinterpret_cast<double2 *>(&ys[buf][kk][c2]) = v;
    }
  };
  auto a_chunk = [&](int64_t k0, double *a) {
#pragma unroll
    for (int s = 0; s < kSbK / 4; ++s) {
      const int64_t k = k0 + 4 * s + kr;
      a[s] = (k < ke && row < d) ? M[k * d + row] : 0.0;
    }
  };
  doublex4 acc[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) acc[j] = doublex4{0.0, 0.0, 0.0, 0.0};
  double a[kSbK / 4], an[kSbK / 4];
  int buf = 0;
  if (kb < ke) {
    stage(0, kb);
    a_chunk(kb, a);
  }
  __syncthreads();
  for (int64_t k0 = kb; k0 < ke; k0 += kSbK) {
    const int64_t kn = k0 + kSbK;
    if (kn < ke) {  // next chunk: LDS stage and A values in flight during this one
      stage(buf ^ 1, kn);
      a_chunk(kn, an);
    }
#pragma unroll
    for (int s = 0; s < kSbK / 4; ++s) {
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const double bv = ys[buf][4 * s + kr][16 * j + cl];
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bv, acc[j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int s = 0; s < kSbK / 4; ++s) a[s] = an[s];
    __syncthreads();
    buf ^= 1;
  }
  // D layout of v_mfma_f64_16x16x4f64: lane l, value i -> row 4 i + l / 16, column l % 16
  double *P = part + (int64_t)blockIdx.y * d * ldy;
  const int64_t rbase = (int64_t)blockIdx.x * kSbRows + 16 * w;
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t r = rbase + 4 * i + kr;
      if (r < d) P[r * ldy + c0 + 16 * j + cl] = acc[j][i];
    }
}

// Y2[r][c] = alpha * sum over slices (in order) + beta Y1[r][c] + gamma Y0[r][c], for the
// first ncols columns (the others are left untouched).
__global__ void __launch_bounds__(256) k_symm_block_sum(
    const double *__restrict__ part, int32_t slices, int32_t d, int64_t ldy, int32_t ncols,
    double alpha, const double *__restrict__ Y1, double beta, const double *__restrict__ Y0,
    double gamma, double *__restrict__ Y2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)d * ncols) return;
  const int64_t r = e / ncols, c = e % ncols;
  const int64_t o = r * ldy + c;
  double s = 0.0;
  for (int t = 0; t < slices; ++t) s += part[(int64_t)t * d * ldy + o];
  double v = alpha * s;
  if (beta != 0.0) v += beta * Y1[o];
  if (gamma != 0.0) v += gamma * Y0[o];
  Y2[o] = v;
}

// Tall-skinny products C[na x nb] = A^T B of two column blocks (A, B: d x lda / ldb
// row-major): the rows are split into slices, each workgroup computes one 64 x 64 output
// block of one slice (four waves of 32 x 32, v_mfma_f64_16x16x4f64, operands read
// straight from memory: each lane quarter reads 16 consecutive doubles of one row), and
// k_block_gram_sum adds the slices in slice order.
constexpr int kBgSlices = 128;
__global__ void __launch_bounds__(256) k_block_gram_part(
    const double *__restrict__ A, int64_t lda, int32_t na, const double *__restrict__ B,
    int64_t ldb, int32_t nb, int32_t d, int32_t rows_per_slice,
    double *__restrict__ part /* [slices][na][nb] */) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int kr = lane >> 4, cl = lane & 15;
  const int bx = (nb + 63) / 64;
  const int i0 = (int)(blockIdx.y / bx) * 64 + (w >> 1) * 32;  // wave's first output row
  const int j0 = (int)(blockIdx.y % bx) * 64 + (w & 1) * 32;   // and column
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_slice;
  const int64_t r1 = min<int64_t>(d, r0 + rows_per_slice);
  doublex4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = doublex4{0.0, 0.0, 0.0, 0.0};
  for (int64_t k0 = r0; k0 < r1; k0 += 4) {
    const int64_t r = k0 + kr;
    double av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int i = i0 + 16 * a + cl;
      av[a] = (r < r1 && i < na) ? A[r * lda + i] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int j = j0 + 16 * c + cl;
      bv[c] = (r < r1 && j < nb) ? B[r * ldb + j] : 0.0;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
  }
  double *P = part + (int64_t)blockIdx.x * na * nb;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + 16 * a + 4 * q + kr, j = j0 + 16 * c + cl;
        if (i < na && j < nb) P[(int64_t)i * nb + j] = acc[a][c][q];
      }
}

__global__ void k_block_gram_sum(const double *__restrict__ part, int32_t slices, int32_t n,
                                 double *__restrict__ C) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int t = 0; t < slices; ++t) s += part[(int64_t)t * n + e];
  C[e] = s;
}

// res[c] = || XM[:, c] - theta[c] X[:, c] || for c < b: one workgroup per column, the
// squares summed per thread in row order and then over the threads in a fixed tree.
__global__ void __launch_bounds__(256) k_block_residual(const double *__restrict__ X,
                                                        const double *__restrict__ XM,
                                                        int32_t d, int64_t ldy,
                                                        const double *__restrict__ theta,
                                                        double *__restrict__ res) {
  __shared__ double red[4];
  const int c = blockIdx.x, tid = threadIdx.x;
  const double th = theta[c];
  double s = 0.0;
  for (int64_t r = tid; r < d; r += 256) {
    const double v = XM[r * ldy + c] - th * X[r * ldy + c];
    s += v * v;
  }
  s = wave_sum_f64_dpp(s);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) res[c] = sqrt((red[0] + red[1]) + (red[2] + red[3]));
}

}  // namespace sh

using namespace sh;

// k-slice length of sh_symm_block: about 512 workgroups over (row tiles x slices x column
// groups), slices a multiple of the LDS stage.
static int64_t symm_block_kslice(int32_t d, int32_t ncols) {
  const int64_t tiles_r = cdiv(d, kSbRows);
  const int64_t groups = cdiv(ncols / 16, kSbMaxNT);
  int64_t slices = cdiv(512, tiles_r * groups);
  slices = std::max<int64_t>(1, std::min<int64_t>(slices, cdiv(d, kSbK)));
  return align_up((size_t)cdiv(d, slices), kSbK);
}

extern "C" size_t sh_symm_block_workspace(int32_t d, int32_t ncols, int64_t ldy) {
  if (d <= 0 || ncols <= 0 || ldy < ncols) return 0;
  const int64_t slices = cdiv(d, symm_block_kslice(d, ncols));
  return align_up((size_t)slices * d * ldy * sizeof(double), 256);
}

extern "C" int sh_symm_block(const double *M, int32_t d, const double *Y1, const double *Y0,
                             int32_t ncols, int64_t ldy, double alpha, double beta, double gamma,
                             double *Y2, void *ws, size_t ws_bytes, sh_stream_t stream) {
  SH_ARG(M && Y1 && Y2 && ws, "sh_symm_block: NULL pointer");
  SH_ARG(gamma == 0.0 || Y0, "sh_symm_block: gamma != 0 needs Y0");
  SH_ARG(d > 0 && ncols > 0 && ncols % 16 == 0 && ldy >= ncols && ldy % 16 == 0,
         "sh_symm_block: bad sizes d=%d ncols=%d ldy=%lld", d, ncols, (long long)ldy);
  const int tiles = ncols / 16;
  const int groups = (int)cdiv(tiles, kSbMaxNT);
  SH_ARG(tiles % groups == 0 && groups <= 65535,
         "sh_symm_block: ncols=%d must split into equal groups of at most 128 columns", ncols);
  SH_ARG(((uintptr_t)Y1 % 16) == 0, "sh_symm_block: Y1 must be 16-byte aligned");
  SH_ARG(Y2 != Y1 && (Y0 == nullptr || Y2 != Y0), "sh_symm_block: Y2 must not alias Y1 or Y0");
  SH_ARG(ws_bytes >= sh_symm_block_workspace(d, ncols, ldy), "sh_symm_block: workspace too small");
  const int nt = tiles / groups;
  const int64_t kslice = symm_block_kslice(d, ncols);
  const int64_t slices = cdiv(d, kslice);
  hipStream_t st = (hipStream_t)stream;
  double *part = (double *)ws;
  dim3 grid((uint32_t)cdiv(d, kSbRows), (uint32_t)slices, (uint32_t)groups);
#define SH_SB_CASE(N)                                                                      \
  case N:                                                                                  \
    hipLaunchKernelGGL(k_symm_block_part<N>, grid, dim3(256), 0, st, M, d, Y1, ldy,        \
                       (int32_t)kslice, part);                                             \
    break;
  switch (nt) {
    SH_SB_CASE(1) SH_SB_CASE(2) SH_SB_CASE(3) SH_SB_CASE(4)
    SH_SB_CASE(5) SH_SB_CASE(6) SH_SB_CASE(7) SH_SB_CASE(8)
    default:
      SH_ARG(false, "sh_symm_block: internal tile count %d", nt);
  }
#undef SH_SB_CASE
  const int64_t n = (int64_t)d * ncols;
  hipLaunchKernelGGL(k_symm_block_sum, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, part,
                     (int32_t)slices, d, ldy, ncols, alpha, Y1, beta, Y0, gamma, Y2);
  SH_LAUNCH_CHECK("sh_symm_block");
  return SH_OK;
}

static int32_t block_gram_rows(int32_t d) {
  return (int32_t)align_up((size_t)cdiv(d, kBgSlices), 4);
}

extern "C" size_t sh_block_gram_workspace(int32_t d, int32_t na, int32_t nb) {
  if (d <= 0 || na <= 0 || nb <= 0) return 0;
  const int64_t slices = cdiv(d, block_gram_rows(d));
  return align_up((size_t)slices * na * nb * sizeof(double), 256);
}

extern "C" int sh_block_gram(const double *A, int64_t lda, int32_t na, const double *B,
                             int64_t ldb, int32_t nb, int32_t d, double *C, void *ws,
                             size_t ws_bytes, sh_stream_t stream) {
  SH_ARG(A && B && C && ws, "sh_block_gram: NULL pointer");
  SH_ARG(d > 0 && na > 0 && nb > 0 && lda >= na && ldb >= nb, "sh_block_gram: bad sizes");
  SH_ARG(ws_bytes >= sh_block_gram_workspace(d, na, nb), "sh_block_gram: workspace too small");
  const int32_t rps = block_gram_rows(d);
  const int64_t slices = cdiv(d, rps);
  const int64_t blocks = cdiv(na, 64) * cdiv(nb, 64);
  SH_ARG(blocks <= 65535, "sh_block_gram: output too large");
  hipStream_t st = (hipStream_t)stream;
  double *part = (double *)ws;
  hipLaunchKernelGGL(k_block_gram_part, dim3((uint32_t)slices, (uint32_t)blocks), dim3(256), 0,
                     st, A, lda, na, B, ldb, nb, d, rps, part);
  const int64_t n = (int64_t)na * nb;
  hipLaunchKernelGGL(k_block_gram_sum, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, part,
                     (int32_t)slices, (int32_t)n, C);
  SH_LAUNCH_CHECK("sh_block_gram");
  return SH_OK;
}

extern "C" int sh_block_residual(const double *X, const double *XM, int32_t d, int32_t b,
                                 int64_t ldy, const double *theta, double *res,
                                 sh_stream_t stream) {
  SH_ARG(X && XM && theta && res, "sh_block_residual: NULL pointer");
  SH_ARG(d > 0 && b > 0 && ldy >= b, "sh_block_residual: bad sizes");
  hipLaunchKernelGGL(k_block_residual, dim3((uint32_t)b), dim3(256), 0, (hipStream_t)stream,
                     X, XM, d, ldy, theta, res);
  SH_LAUNCH_CHECK("k_block_residual");
  return SH_OK;
}

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
  // W [8][d] and the pass coefficients H [kmax][8]
  return align_up(((size_t)kLP * d + (size_t)kmax * kLP) * sizeof(double), 256);
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
  double *W = (double *)ws;
  double *H = W + (size_t)kLP * d;
  for (int32_t j = j0; j < j0 + n_steps; ++j) {
    const int32_t k = (j + 1) * kLP;  // basis vectors so far
    const double *Vj = V + (int64_t)j * kLP * d;
    hipLaunchKernelGGL(k_symm_rows, dim3((uint32_t)cdiv(d, kSymWaves)),
                       dim3(kSymWaves * 64), 0, st, M, d, Vj, W);
    // two passes of classical Gram-Schmidt; each subtracts only its own coefficients
    for (int pass = 0; pass < 2; ++pass) {
      hipLaunchKernelGGL(k_cgs_proj, dim3((uint32_t)k), dim3(256), 0, st, V, k, d, W, H, j,
                         pass, T_diag + (int64_t)j * kLP * kLP);
      hipLaunchKernelGGL(k_cgs_update, dim3((uint32_t)cdiv(d, 64)), dim3(kUpdWaves * 64), 0, st,
                         V, k, d, H, W);
    }
    if (d <= kQrMaxD)
      hipLaunchKernelGGL(k_cholqr2, dim3(1), dim3(kQrThreads), 0, st, W, d,
                         V + (int64_t)k * d, T_sub + (int64_t)j * kLP * kLP, status);
    else
      hipLaunchKernelGGL(k_cholqr2_wide, dim3(1), dim3(kQrThreads), 0, st, W, d,
                         V + (int64_t)k * d, T_sub + (int64_t)j * kLP * kLP, status);
  }
  SH_LAUNCH_CHECK("sh_lanczos_steps");
  return SH_OK;
}
