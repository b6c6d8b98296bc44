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
#include <algorithm>

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

// blockIdx.y takes centers [y*cs, (y+1)*cs): with few rows the centers are split over
// workgroups and each split's best (score, center) goes to pv/pi for k_km_assign_reduce.
template <typename T>
__global__ __launch_bounds__(256) void k_km_assign(const T *__restrict__ X, int64_t n,
                                                   int32_t d, int64_t ldx,
                                                   const double *__restrict__ C, int32_t k,
                                                   const double *__restrict__ cn, int32_t cs,
                                                   int32_t *__restrict__ labels,
                                                   double *__restrict__ pv,
                                                   int32_t *__restrict__ pi) {
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
  const int cb = blockIdx.y * cs, ce = min(k, cb + cs);
  for (int c0 = cb; c0 < ce; c0 += kAsCenters) {
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
        sC[f][p] = (cc < ce && ff < d) ? C[(int64_t)cc * d + ff] : 0.0;
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
      if (cc < ce) {
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
    if (row < n) {
      if (pv) {
        pv[(int64_t)blockIdx.y * n + row] = v;
        pi[(int64_t)blockIdx.y * n + row] = id;
      } else {
        labels[row] = id == INT32_MAX ? 0 : id;
      }
    }
  }
}

__global__ void k_km_assign_reduce(const double *__restrict__ pv, const int32_t *__restrict__ pi,
                                   int32_t S, int64_t n, int32_t *__restrict__ labels) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double v = pv[row];
  int32_t id = pi[row];
  for (int y = 1; y < S; ++y) {
    const double v2 = pv[(int64_t)y * n + row];
    const int32_t i2 = pi[(int64_t)y * n + row];
    if (v2 < v || (v2 == v && i2 < id)) {
      v = v2;
      id = i2;
    }
  }
  labels[row] = id == INT32_MAX ? 0 : id;
}

// center splits for n rows: enough workgroups to cover the chip when the rows are few
inline int32_t assign_splits(int64_t n, int32_t k, int32_t *cs) {
  const int64_t rt = cdiv(n, kAsRows), ct = cdiv(k, kAsCenters);
  int64_t S = std::min<int64_t>(ct, std::max<int64_t>(1, cdiv(1024, rt)));
  *cs = (int32_t)(cdiv(ct, S) * kAsCenters);
  return (int32_t)cdiv(k, *cs);
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
// One launch per greedy round (the rounds are sequential by construction).  Every block
// scores its 256 points against the round's candidates: d_t(i) = min(closest_i,
// max(-2 x_i.c_t + ||c_t||^2 + ||x_i||^2, 0)), kept per candidate in the workspace, and
// the block's partial potentials sum_i d_t(i).  The last block to arrive (an atomic count,
// with agent-scope fences on both sides: blocks may sit on different XCDs and L2s) sums
// the partials in a fixed order, keeps the first candidate of smallest potential, and
// draws the next round's candidates: each threshold u_t * potential (the host's uniform
// draws) goes to the first point whose running sum of closest distances reaches it
// (np.searchsorted over the cumulative sum; past the end -> the last point).  Nothing
// waits on another block.
constexpr int kPpBlock = 256, kPpWaves = kPpBlock / 64;
constexpr int kPpMaxTrials = 16, kPpMaxD = 256;

struct PpState {
  double pot;
  int32_t best;
  uint32_t done;
  int64_t cand[kPpMaxTrials];
};

template <typename T>
__global__ void k_km_pp_init(const T *__restrict__ X, int64_t ldx,
                             const int64_t *__restrict__ rows, int64_t n, int32_t d,
                             int64_t first, double *__restrict__ xn, PpState *st) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const T *xr = X + (rows ? rows[i] : i) * ldx;
    double s = 0.0;
    for (int f = 0; f < d; ++f) {
      const double v = ld_f64(xr + f);
      s = fma(v, v, s);
    }
    xn[i] = s;
  }
  if (i == 0) {
    st->cand[0] = first;
    st->best = 0;
    st->pot = 0.0;
    st->done = 0u;
  }
}

// block-wide sums of v[t], t < Le, in a fixed order; the result is valid in thread t
__device__ __forceinline__ double pp_block_sum(double v, double (*sred)[kPpMaxTrials], int t_slot,
                                               int lane, int wv) {
  const double s = wave_sum_f64(v);
  if (lane == 0) sred[wv][t_slot] = s;
  return s;
}

template <typename T>
__global__ __launch_bounds__(kPpBlock) void k_km_pp_round(
    const T *__restrict__ X, int64_t ldx, const int64_t *__restrict__ rows, int64_t n,
    int32_t d, int32_t k, int32_t L, int32_t c, const double *__restrict__ U,
    const double *__restrict__ xn, double *__restrict__ dmin, double *__restrict__ part,
    PpState *st, int64_t *__restrict__ out) {
  __shared__ double sCand[kPpMaxTrials * kPpMaxD];
  __shared__ double sCn[kPpMaxTrials];
  __shared__ int64_t sCid[kPpMaxTrials];
  __shared__ double sRed[kPpWaves][kPpMaxTrials];
  __shared__ double sScan[kPpWaves];
  __shared__ unsigned long long sNext[kPpMaxTrials];
  __shared__ unsigned long long sBlk[kPpMaxTrials];
  __shared__ double sStart[kPpMaxTrials];
  __shared__ double sPot;
  __shared__ int sBest, sLast;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int Le = c == 0 ? 1 : L;
  const int64_t B = gridDim.x;
  double *cur = dmin + (int64_t)(c & 1) * L * n;
  const double *prev = dmin + (int64_t)((c + 1) & 1) * L * n;

  // ---- score this block's points against the round's candidates
  if (tid < Le) {
    sCid[tid] = st->cand[tid];
    sCn[tid] = xn[st->cand[tid]];
  }
  __syncthreads();
  for (int e = tid; e < Le * d; e += kPpBlock) {
    const int t = e / d, f = e % d;
    const int64_t ci = sCid[t];
    sCand[t * d + f] = ld_f64(X + (rows ? rows[ci] : ci) * ldx + f);
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kPpBlock + tid;
  double dm[kPpMaxTrials];
#pragma unroll
  for (int t = 0; t < kPpMaxTrials; ++t) dm[t] = 0.0;
  if (i < n) {
    const double ci = c == 0 ? __builtin_inf() : prev[(int64_t)st->best * n + i];
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
    const double ni = xn[i];
#pragma unroll
    for (int t = 0; t < kPpMaxTrials; ++t) {
      if (t < Le) {
        double dist = (-2.0 * acc[t] + sCn[t]) + ni;
        dist = dist > 0.0 ? dist : 0.0;
        dm[t] = ci < dist ? ci : dist;
        cur[(int64_t)t * n + i] = dm[t];
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kPpMaxTrials; ++t)
    if (t < Le) pp_block_sum(dm[t], sRed, t, lane, wv);
  __syncthreads();
  if (tid < Le) {
    double s = 0.0;
    for (int w = 0; w < kPpWaves; ++w) s += sRed[w][tid];
    part[(int64_t)blockIdx.x * kPpMaxTrials + tid] = s;
  }
  // ---- arrival: the last block selects
  __threadfence();
  __syncthreads();
  if (tid == 0) {
    const uint32_t before =
        __hip_atomic_fetch_add(&st->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    sLast = before == (uint32_t)(B - 1);
  }
  __syncthreads();
  if (!sLast) return;
  __threadfence();

  // potentials: block partials summed over a fixed thread stride, then across waves
  double ps[kPpMaxTrials];
#pragma unroll
  for (int t = 0; t < kPpMaxTrials; ++t) ps[t] = 0.0;
  for (int64_t b = tid; b < B; b += kPpBlock) {
#pragma unroll
    for (int t = 0; t < kPpMaxTrials; ++t)
      if (t < Le) ps[t] += part[b * kPpMaxTrials + t];
  }
#pragma unroll
  for (int t = 0; t < kPpMaxTrials; ++t)
    if (t < Le) pp_block_sum(ps[t], sRed, t, lane, wv);
  if (tid < kPpMaxTrials) {
    sNext[tid] = (unsigned long long)(n - 1);
    sBlk[tid] = (unsigned long long)B;  // none yet
  }
  __syncthreads();
  if (tid == 0) {
    int best = 0;
    double bp = 0.0;
    for (int t = 0; t < Le; ++t) {
      double p = 0.0;
      for (int w = 0; w < kPpWaves; ++w) p += sRed[w][t];
      if (t == 0 || p < bp) {
        bp = p;
        best = t;
      }
    }
    sBest = best;
    sPot = bp;
    out[c] = sCid[best];
    st->best = best;
    st->pot = bp;
  }
  __syncthreads();
  const int best = sBest;
  if (c + 1 < k) {
    // closest distances are now cur[best]; block sums of them are part[b][best]
    const int64_t Q = (B + kPpBlock - 1) / kPpBlock;
    const int64_t b0 = min(B, (int64_t)tid * Q), b1 = min(B, b0 + Q);
    double seg = 0.0;
    for (int64_t b = b0; b < b1; ++b) seg += part[b * kPpMaxTrials + best];
    double v = seg;
    for (int o = 1; o < 64; o <<= 1) {
      const double t = __shfl_up(v, o, 64);
      if (lane >= o) v += t;
    }
    if (lane == 63) sScan[wv] = v;
    __syncthreads();
    double off = 0.0;
    for (int w = 0; w < wv; ++w) off += sScan[w];
    const double incl = v + off;
    // the previous thread's inclusive sum, exactly: lane 63 of wave w-1 holds
    // sScan[w-1] + off(w-1), the same two operands as off(w)
    const double up = __shfl_up(incl, 1, 64);
    const double excl = lane > 0 ? up : off;
    const double pot = sPot;
    const double *closest = cur + (int64_t)best * n;
    // which block each threshold falls in: the owner of the crossing walks its blocks
    for (int t = 0; t < L; ++t) {
      const double r = U[(int64_t)c * L + t] * pot;
      const bool hit = b1 > b0 && (tid == 0 ? r <= incl : (excl < r && r <= incl));
      if (hit) {
        double run = tid == 0 ? 0.0 : excl;
        int64_t blk = b1 - 1;
        for (int64_t b = b0; b < b1 - 1; ++b) {
          const double nb = run + part[b * kPpMaxTrials + best];
          if (nb >= r) {
            blk = b;
            break;
          }
          run = nb;
        }
        // rounding can let two neighbouring prefix sums straddle r; the first wins
        atomicMin(&sBlk[t], (unsigned long long)blk);
      }
    }
    __syncthreads();
    // the running sum at the start of each chosen block, from its owner
    for (int t = 0; t < L; ++t) {
      const int64_t blk = (int64_t)sBlk[t];
      if (blk >= b0 && blk < b1 && blk < B) {
        double run = tid == 0 ? 0.0 : excl;
        for (int64_t b = b0; b < blk; ++b) run += part[b * kPpMaxTrials + best];
        sStart[t] = run;
      }
    }
    __syncthreads();
    // one wave per threshold: the first point of the block whose running sum reaches it
    for (int t = wv; t < L; t += kPpWaves) {
      const int64_t blk = (int64_t)sBlk[t];
      int64_t id = n - 1;
      if (blk < B) {
        const double r = U[(int64_t)c * L + t] * pot;
        const int64_t q0 = blk * kPpBlock, e = min(n, q0 + kPpBlock);
        const int64_t ql = q0 + 4 * lane;
        double p[4];
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          acc += ql + u < e ? closest[ql + u] : 0.0;
          p[u] = acc;
        }
        double ex = acc;
        for (int o = 1; o < 64; o <<= 1) {
          const double y = __shfl_up(ex, o, 64);
          if (lane >= o) ex += y;
        }
        ex -= acc;  // exclusive lane prefix (only compared)
        const double start = sStart[t];
        int first = -1;
#pragma unroll
        for (int u = 3; u >= 0; --u)
          if (ql + u < e && start + (ex + p[u]) >= r) first = u;
        const uint64_t mask = __ballot(first >= 0);
        if (mask) {
          const int fl = __ffsll((unsigned long long)mask) - 1;
          const int fu = __shfl(first, fl, 64);
          id = q0 + 4 * fl + fu;
        } else {
          id = e - 1;
        }
      }
      if (lane == 0) sNext[t] = (unsigned long long)id;
    }
    __syncthreads();
    if (tid < L) st->cand[tid] = (int64_t)sNext[tid];
  }
  if (tid == 0) st->done = 0u;
}

inline bool km_dtype_ok(int x_dtype) { return x_dtype == SH_F32 || x_dtype == SH_F64; }

inline size_t pp_state_bytes() { return align_up(sizeof(PpState), 256); }

}  // namespace sh

using namespace sh;

extern "C" size_t sh_kmeans_pp_workspace(int64_t n, int32_t n_trials) {
  if (n < 1 || n_trials < 1) return 0;
  const int64_t B = cdiv(n, kPpBlock);
  return pp_state_bytes() +
         (size_t)(n + 2 * (int64_t)n_trials * n + B * kPpMaxTrials) * sizeof(double);
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
  SH_ARG(cdiv(n, kPpBlock) <= (int64_t)0x7fffffff, "sh_kmeans_pp: n too large");
  const size_t need = sh_kmeans_pp_workspace(n, n_trials);
  if (ws_bytes < need) {
    set_error("sh_kmeans_pp: workspace %zu bytes < %zu", ws_bytes, need);
    return SH_EWORKSPACE;
  }
  PpState *st = (PpState *)ws;
  double *xn = (double *)((char *)ws + pp_state_bytes());
  double *dmin = xn + n;
  double *part = dmin + 2 * (int64_t)n_trials * n;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((uint32_t)cdiv(n, kPpBlock));
  if (x_dtype == SH_F32)
    hipLaunchKernelGGL(k_km_pp_init<float>, grid, dim3(kPpBlock), 0, s, (const float *)X, ldx,
                       rows, n, d, first, xn, st);
  else
    hipLaunchKernelGGL(k_km_pp_init<double>, grid, dim3(kPpBlock), 0, s, (const double *)X,
                       ldx, rows, n, d, first, xn, st);
  SH_LAUNCH_CHECK("k_km_pp_init");
  for (int32_t c = 0; c < n_clusters; ++c) {
    if (x_dtype == SH_F32)
      hipLaunchKernelGGL(k_km_pp_round<float>, grid, dim3(kPpBlock), 0, s, (const float *)X,
                         ldx, rows, n, d, n_clusters, n_trials, c, uniforms, xn, dmin, part, st,
                         centers_out);
    else
      hipLaunchKernelGGL(k_km_pp_round<double>, grid, dim3(kPpBlock), 0, s, (const double *)X,
                         ldx, rows, n, d, n_clusters, n_trials, c, uniforms, xn, dmin, part, st,
                         centers_out);
    SH_LAUNCH_CHECK("k_km_pp_round");
  }
  return SH_OK;
}

extern "C" size_t sh_kmeans_assign_workspace(int64_t n, int32_t n_clusters) {
  if (n < 1 || n_clusters < 1) return 0;
  int32_t cs;
  const int32_t S = assign_splits(n, n_clusters, &cs);
  return S > 1 ? align_up((size_t)S * n * sizeof(double), 256) + (size_t)S * n * sizeof(int32_t)
               : 0;
}

extern "C" int sh_kmeans_assign(const void *X, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                                const double *centers, int32_t n_clusters,
                                double *center_norms, int32_t *labels, void *ws,
                                size_t ws_bytes, sh_stream_t stream) {
  SH_ARG(km_dtype_ok(x_dtype), "sh_kmeans_assign: x_dtype must be SH_F32 or SH_F64");
  SH_ARG(n >= 0 && d >= 1 && ldx >= d && n_clusters >= 1,
         "sh_kmeans_assign: need n >= 0, d >= 1, ldx >= d, n_clusters >= 1");
  if (n == 0) return SH_OK;
  SH_ARG(X && centers && center_norms && labels, "sh_kmeans_assign: NULL pointer");
  SH_ARG(cdiv(n, kAsRows) <= (int64_t)0x7fffffff, "sh_kmeans_assign: n too large");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_km_row_norms, dim3((uint32_t)cdiv(n_clusters, 256)), dim3(256), 0, s,
                     centers, n_clusters, d, center_norms);
  SH_LAUNCH_CHECK("k_km_row_norms");
  int32_t cs;
  int32_t S = assign_splits(n, n_clusters, &cs);
  const size_t need = sh_kmeans_assign_workspace(n, n_clusters);
  if (S > 1 && (ws == nullptr || ws_bytes < need)) {  // no room to split: one pass
    S = 1;
    cs = n_clusters;
  }
  double *pv = nullptr;
  int32_t *pi = nullptr;
  if (S > 1) {
    pv = (double *)ws;
    pi = (int32_t *)((char *)ws + align_up((size_t)S * n * sizeof(double), 256));
  }
  const dim3 grid((uint32_t)cdiv(n, kAsRows), (uint32_t)S);
  if (x_dtype == SH_F32)
    hipLaunchKernelGGL(k_km_assign<float>, grid, dim3(256), 0, s, (const float *)X, n, d, ldx,
                       centers, n_clusters, center_norms, cs, labels, pv, pi);
  else
    hipLaunchKernelGGL(k_km_assign<double>, grid, dim3(256), 0, s, (const double *)X, n, d,
                       ldx, centers, n_clusters, center_norms, cs, labels, pv, pi);
  SH_LAUNCH_CHECK("k_km_assign");
  if (S > 1) {
    hipLaunchKernelGGL(k_km_assign_reduce, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, s, pv,
                       pi, S, n, labels);
    SH_LAUNCH_CHECK("k_km_assign_reduce");
  }
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
