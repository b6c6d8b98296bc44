// Mapping scores of reference cells (get_mapping_score, mapping_datastore.py:211-285):
//   ms_g[ref] = multiplier / (n_g * k) * sum_{q in g, j} w(d_qj) [ref == idx_qj]
//   w(d) = 1 / (log1p(d) + 1)  (or fixed_weight), then log1p(ms) if log_transform.
// The sums are kept in unsigned 2^-40 fixed point and accumulated with 64-bit
// integer atomics, which are associative: the result does not depend on launch
// order or batch composition.
#include <mutex>

#include "sh_common.h"

namespace sh {

constexpr double kFix = 1099511627776.0;  // 2^40

__global__ void k_ms_clear(uint64_t *__restrict__ acc, int64_t m, int32_t *__restrict__ cnt,
                           int32_t n_groups) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) acc[t] = 0;
  if (t < n_groups) cnt[t] = 0;
}

__global__ void k_ms_scatter(const int64_t *__restrict__ idx, const float *__restrict__ dist,
                             int64_t nq, int32_t k, const int32_t *__restrict__ groups,
                             int32_t n_groups, int64_t n_ref, int weighted, double fixed_weight,
                             uint64_t *__restrict__ acc, int32_t *__restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq * k) return;
  const int64_t q = t / k;
  const int g = groups ? groups[q] : 0;
  if (g < 0 || g >= n_groups) return;
  // queries per group: one atomic per wave when the wave's queries share a group
  const bool head = t % k == 0;
  const uint64_t hm = __ballot(head);
  if (hm) {
    const int lead = __ffsll((unsigned long long)hm) - 1;
    const int g0 = __shfl(g, lead, 64);
    if (__ballot(head && g != g0) == 0) {
      if ((int)(threadIdx.x & 63) == lead) atomicAdd(&cnt[g0], (int)__popcll(hm));
    } else if (head) {
      atomicAdd(&cnt[g], 1);
    }
  }
  const int64_t r = idx[t];
  if (r < 0 || r >= n_ref) return;
  const double w = weighted ? 1.0 / (log1p((double)dist[t]) + 1.0) : fixed_weight;
  const uint64_t f = (uint64_t)llrint(w * kFix);
  atomicAdd((unsigned long long *)&acc[(int64_t)g * n_ref + r], (unsigned long long)f);
}

// acc (fixed point) -> ms.  With `sums` non-NULL the target rows came in several
// batches: `sums` holds the FP64 sum of the earlier batches' fixed-point totals and
// acc the last batch's; they are added in batch order.
__global__ void k_ms_final(const uint64_t *__restrict__ acc, const double *__restrict__ sums,
                           const int32_t *__restrict__ cnt, int32_t n_groups, int64_t n_ref,
                           int32_t k, double multiplier, int log_t, double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_groups * n_ref) return;
  const int g = (int)(t / n_ref);
  double s = (double)acc[t] / kFix;
  if (sums) s = sums[t] + s;
  const double den = (double)cnt[g] * (double)k;
  double ms = multiplier * s / den;
  out[t] = log_t ? log1p(ms) : ms;
}

// Between batches: sums (+)= acc / 2^40, acc = 0.
__global__ void k_ms_fold(uint64_t *__restrict__ acc, double *__restrict__ sums, int64_t m,
                          int first) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const double s = (double)acc[t] / kFix;
  sums[t] = first ? s : sums[t] + s;
  acc[t] = 0;
}


// ---- label transfer -------------------------------------------------------------------
// numpy's float64 sum of a short row: sequential below 8 elements, else eight running
// accumulators combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) plus a sequential tail
// (numpy's pairwise_sum for n <= 128), so a sum and its comparisons match numpy bit for bit.
__device__ double numpy_sum(const double *v, int n) {
  if (n < 8) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += v[i];
    return s;
  }
  double r[8];
  for (int j = 0; j < 8; ++j) r[j] = v[j];
  int i = 8;
  for (; i + 8 <= n; i += 8)
    for (int j = 0; j < 8; ++j) r[j] += v[i + j];
  double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) s += v[i];
  return s;
}

constexpr int kTcMaxK = 128;

// One thread per target row: weights 1 - d / max(d) over the whole row, the first k - 1
// neighbours vote with their weights for their reference class (classes in order of
// first appearance, sums in neighbour order); the row's class is the single class whose
// share of the vote exceeds `threshold`, else -1.
__global__ void k_target_classes(const int64_t *__restrict__ idx, const double *__restrict__ dist,
                                 int64_t n, int32_t k, const int32_t *__restrict__ groups,
                                 int64_t n_ref, double threshold, int32_t *__restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int64_t *ir = idx + row * k;
  const double *dr = dist + row * k;
  double dmax = dr[0];  // numpy max: a NaN anywhere gives NaN
  for (int j = 1; j < k; ++j) {
    const double v = dr[j];
    if (dmax == dmax && (v != v || v > dmax)) dmax = v;
  }
  double w[kTcMaxK];
  int32_t cls[kTcMaxK];
  double csum[kTcMaxK];
  int nc = 0;
  const int m = k - 1;
  for (int j = 0; j < m; ++j) {
    w[j] = 1.0 - dr[j] / dmax;
    const int64_t r = ir[j];
    const int32_t c = (r >= 0 && r < n_ref) ? groups[r] : -2;
    int slot = -1;
    for (int t = 0; t < nc; ++t)
      if (cls[t] == c) {
        slot = t;
        break;
      }
    if (slot < 0) {
      slot = nc++;
      cls[slot] = c;
      csum[slot] = 0.0;
    }
    csum[slot] += w[j];
  }
  const double s = numpy_sum(w, m);
  int32_t res = -1;
  for (int t = 0; t < nc; ++t) {
    if (csum[t] / s > threshold) {
      if (res == -1) {
        res = cls[t];
      } else {
        res = -1;
        break;
      }
    }
  }
  out[row] = res;
}

}  // namespace sh

using namespace sh;

extern "C" size_t sh_mapping_score_workspace(int64_t n_ref, int32_t n_groups) {
  return align_up((size_t)n_ref * (n_groups > 0 ? n_groups : 1) * sizeof(uint64_t), 256) +
         align_up((size_t)(n_groups > 0 ? n_groups : 1) * sizeof(int32_t), 256);
}

extern "C" int sh_mapping_score(const int64_t *idx, const float *dist, int64_t nq, int32_t k,
                                const int32_t *groups, int32_t n_groups, int64_t n_ref,
                                int weighted, double fixed_weight, double multiplier,
                                int log_transform, double *out, void *workspace,
                                size_t workspace_bytes, sh_stream_t stream) {
  SH_ARG(nq >= 0 && k >= 1 && n_ref >= 1 && n_groups >= 1, "sh_mapping_score: bad sizes");
  SH_ARG(groups || n_groups == 1, "sh_mapping_score: groups NULL requires n_groups == 1");
  SH_ARG(idx && dist && out, "sh_mapping_score: NULL pointer");
  // one (query, neighbour) pair adds at most w_max (1 for weighted with d >= 0) to a
  // reference's 64-bit fixed-point sum; a batch of nb rows adds at most nb * k * w_max
  const double w_max = weighted ? 1.0 : fixed_weight;
  SH_ARG(weighted || (fixed_weight > 0 && fixed_weight * kFix * k < 1.8e19),
         "sh_mapping_score: fixed_weight out of range");
  const int64_t nb = (int64_t)(1.8e19 / (kFix * (double)k * (w_max > 1.0 ? w_max : 1.0)));
  const size_t need = sh_mapping_score_workspace(n_ref, n_groups);
  if (!workspace || workspace_bytes < need) {
    set_error("sh_mapping_score: workspace %zu < %zu bytes", workspace_bytes, need);
    return SH_EWORKSPACE;
  }
  hipStream_t st = (hipStream_t)stream;
  uint64_t *acc = (uint64_t *)workspace;
  int32_t *cnt = (int32_t *)((char *)workspace +
                             align_up((size_t)n_ref * n_groups * sizeof(uint64_t), 256));
  const int64_t m = n_ref * n_groups;
  hipLaunchKernelGGL(k_ms_clear, dim3(cdiv(m > n_groups ? m : n_groups, 256)), dim3(256), 0, st,
                     acc, m, cnt, n_groups);
  SH_LAUNCH_CHECK("k_ms_clear");
  // batches of nb target rows; `out` holds the running FP64 sums between batches
  const double *sums = nullptr;
  for (int64_t a = 0; a < nq; a += nb) {
    const int64_t b = nq - a < nb ? nq - a : nb;
    if (a > 0) {
      hipLaunchKernelGGL(k_ms_fold, dim3(cdiv(m, 256)), dim3(256), 0, st, acc, out, m,
                         (int)(a == nb));
      SH_LAUNCH_CHECK("k_ms_fold");
      sums = out;
    }
    hipLaunchKernelGGL(k_ms_scatter, dim3(cdiv(b * k, 256)), dim3(256), 0, st, idx + a * k,
                       dist + a * k, b, k, groups ? groups + a : nullptr, n_groups, n_ref,
                       weighted, fixed_weight, acc, cnt);
    SH_LAUNCH_CHECK("k_ms_scatter");
  }
  hipLaunchKernelGGL(k_ms_final, dim3(cdiv(m, 256)), dim3(256), 0, st, acc, sums, cnt, n_groups,
                     n_ref, k, multiplier, log_transform, out);
  SH_LAUNCH_CHECK("k_ms_final");
  return SH_OK;
}

// ---- allocation-free variant ----------------------------------------------------------
// The accumulators of sh_mapping_score live in a file-scope device array, so a caller
// that cannot allocate (code under graph capture, a tight serving loop) passes only its
// inputs and outputs.  Each GPU holds its own copy of the array for the life of the
// process; its device address is looked up once per device with hipGetSymbolAddress and
// only that address is handed to kernels.  Every call on one GPU shares the array:
// calls that overlap on one GPU (different streams or host threads) race.
namespace sh {
constexpr int64_t kScratchAccWords = (int64_t)1 << 23;  // n_ref * n_groups <= 8,388,608
constexpr int32_t kScratchMaxGroups = 4096;
constexpr size_t kScratchBytes = (size_t)kScratchAccWords * sizeof(uint64_t) +
                                 (size_t)kScratchMaxGroups * sizeof(int32_t);
static __device__ __align__(256) uint64_t scratch[kScratchBytes / sizeof(uint64_t)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

}  // namespace sh

extern "C" size_t sh_mapping_score_static_max(int32_t n_groups) {
  if (n_groups < 1 || n_groups > sh::kScratchMaxGroups) return 0;
  return (size_t)(sh::kScratchAccWords / n_groups);
}

extern "C" int sh_mapping_score_static(const int64_t *idx, const float *dist, int64_t nq,
                                       int32_t k, const int32_t *groups, int32_t n_groups,
                                       int64_t n_ref, int weighted, double fixed_weight,
                                       double multiplier, int log_transform, double *out,
                                       sh_stream_t stream) {
  using namespace sh;
  SH_ARG(n_ref >= 1 && n_groups >= 1 && n_groups <= kScratchMaxGroups &&
             n_ref * (int64_t)n_groups <= kScratchAccWords,
         "sh_mapping_score_static: n_ref x n_groups = %lld x %d exceeds the static scratch "
         "(%lld accumulators, at most %d groups); use sh_mapping_score with a workspace",
         (long long)n_ref, n_groups, (long long)kScratchAccWords, kScratchMaxGroups);
  int dev = -1;
  SH_HIP(hipGetDevice(&dev));
  SH_ARG(dev >= 0 && dev < kMaxDevices, "sh_mapping_score_static: device %d out of range", dev);
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      SH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = p;
    }
    base = g_scratch_ptr[dev];
  }
  // same layout as the workspace: accumulators, then the group counts (256-aligned)
  const size_t need = sh_mapping_score_workspace(n_ref, n_groups);
  SH_ARG(need <= kScratchBytes, "sh_mapping_score_static: scratch layout exceeds its size");
  return sh_mapping_score(idx, dist, nq, k, groups, n_groups, n_ref, weighted, fixed_weight,
                          multiplier, log_transform, out, base, kScratchBytes, stream);
}

extern "C" int sh_target_classes(const int64_t *idx, const double *dist, int64_t n, int32_t k,
                                 const int32_t *groups, int64_t n_ref, double threshold,
                                 int32_t *out, sh_stream_t stream) {
  SH_ARG(n >= 0 && k >= 2 && k <= kTcMaxK + 1 && n_ref >= 1,
         "sh_target_classes: need n >= 0, 2 <= k <= %d, n_ref >= 1", kTcMaxK + 1);
  if (n == 0) return SH_OK;
  SH_ARG(idx && dist && groups && out, "sh_target_classes: NULL pointer");
  hipLaunchKernelGGL(k_target_classes, dim3((uint32_t)cdiv(n, 128)), dim3(128), 0,
                     (hipStream_t)stream, idx, dist, n, k, groups, n_ref, threshold, out);
  SH_LAUNCH_CHECK("k_target_classes");
  return SH_OK;
}
