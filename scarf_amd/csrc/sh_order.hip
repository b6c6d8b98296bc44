// Locality order of embedding rows and row permutation.
//
// The kNN screen evaluates 32 queries per MFMA tile; a tile only leaves its fast path when
// one of its queries finds a candidate below its running threshold.  Queries whose
// embeddings are close share neighbours and pass in the same tiles, so the screen takes
// its queries in a locality order: the Morton (Z-order) code of the first three
// coordinates (the leading principal components of a PCA embedding), quantised to 10
// bits each over their range, sorted stably.  Candidate pools are independent of the
// query order, so results are unchanged; only the time spent selecting falls.
#include <hipcub/hipcub.hpp>

#include "sh_common.h"

namespace sh {

constexpr int kOrderCoords = 3;
constexpr int kOrderBits = 10;

// per-coordinate min / max as orderable integers (order-independent atomics)
__global__ void k_order_init(uint32_t *mm) {
  const int t = threadIdx.x;
  if (t < kOrderCoords) {
    mm[2 * t] = 0xffffffffu;  // min
    mm[2 * t + 1] = 0u;       // max
  }
}

__global__ void k_order_minmax(const float *__restrict__ E, int64_t n, int32_t D, int64_t lde,
                               int nc, uint32_t *__restrict__ mm) {
  uint32_t lo[kOrderCoords], hi[kOrderCoords];
#pragma unroll
  for (int c = 0; c < kOrderCoords; ++c) {
    lo[c] = 0xffffffffu;
    hi[c] = 0u;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int c = 0; c < kOrderCoords; ++c) {
      if (c < nc) {
        const uint32_t o = float_ord(E[i * lde + c]);
        lo[c] = min(lo[c], o);
        hi[c] = max(hi[c], o);
      }
    }
  }
  // wave, then workgroup (LDS), then one atomic per workgroup and coordinate: the six
  // words are shared by every workgroup, so per-wave atomics would queue on them
  __shared__ uint32_t red[4][2 * kOrderCoords];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kOrderCoords; ++c) {
    for (int m = 32; m >= 1; m >>= 1) {
      lo[c] = min(lo[c], (uint32_t)__shfl_xor((int)lo[c], m, kWave));
      hi[c] = max(hi[c], (uint32_t)__shfl_xor((int)hi[c], m, kWave));
    }
    if ((threadIdx.x & 63) == 0) {
      red[w][2 * c] = lo[c];
      red[w][2 * c + 1] = hi[c];
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * kOrderCoords && (int)threadIdx.x < 2 * nc) {
    const int j = threadIdx.x;
    uint32_t v = red[0][j];
    for (int i = 1; i < 4; ++i) v = (j & 1) ? max(v, red[i][j]) : min(v, red[i][j]);
    if (j & 1)
      atomicMax(&mm[j], v);
    else
      atomicMin(&mm[j], v);
  }
}

__global__ void k_order_keys(const float *__restrict__ E, int64_t n, int64_t lde, int nc,
                             const uint32_t *__restrict__ mm, uint32_t *__restrict__ key,
                             int64_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t q[kOrderCoords] = {0, 0, 0};
  for (int c = 0; c < nc; ++c) {
    const float lo = ord_float(mm[2 * c]), hi = ord_float(mm[2 * c + 1]);
    const float span = hi - lo;
    float t = span > 0.f ? (E[i * lde + c] - lo) / span : 0.f;
    t = fminf(fmaxf(t, 0.f), 1.f);
    q[c] = min((uint32_t)(t * (float)((1 << kOrderBits) - 1) + 0.5f),
               (uint32_t)((1 << kOrderBits) - 1));
  }
  uint32_t code = 0;
#pragma unroll
  for (int b = 0; b < kOrderBits; ++b)
#pragma unroll
    for (int c = 0; c < kOrderCoords; ++c) code |= ((q[c] >> b) & 1u) << (b * kOrderCoords + c);
  key[i] = code;
  idx[i] = i;
}

// dst row i <- src row perm[i] (gather) or dst row perm[i] <- src row i (scatter), in
// 4-byte words
__global__ void k_permute_rows(const uint32_t *__restrict__ src, int64_t src_ld,
                               const int64_t *__restrict__ perm, int64_t n, int64_t words,
                               uint32_t *__restrict__ dst, int64_t dst_ld, int scatter) {
  const int64_t total = n * words;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / words, w = e % words;
    const int64_t p = perm[i];
    if (scatter)
      dst[p * dst_ld + w] = src[i * src_ld + w];
    else
      dst[i * dst_ld + w] = src[p * src_ld + w];
  }
}

static size_t sort_temp_bytes(int64_t n) {
  size_t tmp = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t *)nullptr,
                                     (uint32_t *)nullptr, (const int64_t *)nullptr,
                                     (int64_t *)nullptr, (int)n, 0, kOrderCoords * kOrderBits);
  return tmp;
}

}  // namespace sh

using namespace sh;

extern "C" size_t sh_locality_order_workspace(int64_t n) {
  if (n <= 0 || n > 0x7fffffff) return 0;
  return 256 + align_up((size_t)n * 4, 256) * 2 + align_up((size_t)n * 8, 256) +
         align_up(sort_temp_bytes(n), 256);
}

extern "C" int sh_locality_order(const float *E, int64_t n, int32_t D, int64_t lde,
                                 int64_t *perm_out, void *ws, size_t ws_bytes,
                                 sh_stream_t stream) {
  SH_ARG(n >= 0 && n <= 0x7fffffff, "sh_locality_order: n out of range");
  SH_ARG(D >= 1 && lde >= D, "sh_locality_order: bad D / lde");
  if (n == 0) return SH_OK;
  SH_ARG(E && perm_out && ws, "sh_locality_order: NULL pointer");
  SH_ARG(ws_bytes >= sh_locality_order_workspace(n), "sh_locality_order: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  const int nc = D < kOrderCoords ? D : kOrderCoords;
  unsigned char *p = (unsigned char *)ws;
  uint32_t *mm = (uint32_t *)p;
  p += 256;
  uint32_t *key_in = (uint32_t *)p;
  p += align_up((size_t)n * 4, 256);
  uint32_t *key_out = (uint32_t *)p;
  p += align_up((size_t)n * 4, 256);
  int64_t *idx_in = (int64_t *)p;
  p += align_up((size_t)n * 8, 256);
  size_t tmp_bytes = sort_temp_bytes(n);
  void *tmp = p;
  hipLaunchKernelGGL(k_order_init, dim3(1), dim3(64), 0, st, mm);
  const int64_t blocks = std::min<int64_t>(cdiv(n, 256), 512);
  hipLaunchKernelGGL(k_order_minmax, dim3((uint32_t)blocks), dim3(256), 0, st, E, n, D, lde, nc,
                     mm);
  hipLaunchKernelGGL(k_order_keys, dim3((uint32_t)cdiv(n, 256)), dim3(256), 0, st, E, n, lde,
                     nc, mm, key_in, idx_in);
  SH_LAUNCH_CHECK("k_order_keys");
  SH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key_in, key_out, idx_in, perm_out,
                                            (int)n, 0, kOrderCoords * kOrderBits, st));
  return SH_OK;
}

extern "C" int sh_permute_rows(const void *src, int64_t src_ld_bytes, const int64_t *perm,
                               int64_t n, int64_t row_bytes, void *dst, int64_t dst_ld_bytes,
                               int scatter, sh_stream_t stream) {
  SH_ARG(n >= 0 && row_bytes >= 0, "sh_permute_rows: negative size");
  if (n == 0 || row_bytes == 0) return SH_OK;
  SH_ARG(src && dst && perm, "sh_permute_rows: NULL pointer");
  SH_ARG(row_bytes % 4 == 0 && src_ld_bytes % 4 == 0 && dst_ld_bytes % 4 == 0,
         "sh_permute_rows: row sizes and strides must be multiples of 4 bytes");
  SH_ARG(src_ld_bytes >= row_bytes && dst_ld_bytes >= row_bytes,
         "sh_permute_rows: stride smaller than the row");
  const int64_t words = row_bytes / 4;
  const int64_t blocks = std::min<int64_t>(cdiv(n * words, 256), 65536);
  hipLaunchKernelGGL(k_permute_rows, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream,
                     (const uint32_t *)src, src_ld_bytes / 4, perm, n, words, (uint32_t *)dst,
                     dst_ld_bytes / 4, scatter);
  SH_LAUNCH_CHECK("k_permute_rows");
  return SH_OK;
}
