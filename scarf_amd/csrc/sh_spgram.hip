// Sparse FP64 Gram of z = clip((x - center) * scale) read straight from the CSR counts: the
// Z^T Z and column sums that sh_densify + sh_gram produce for an IncrementalPCA partial fit
// (sklearn decomposition/_incremental_pca.py:345-363, driven from scarf ann.py:229-250),
// without densifying z.
//
// Decomposition.  A zero count gives z_f = c_f = clip(-center_f * scale_f), so
// z_rf = c_f + delta_rf with delta non-zero only at stored entries, and
//   Z^T Z = D + c s^T + s c^T + n c c^T,   D = delta^T delta,   s = delta^T 1,
//   column sums = s + n c.
// D costs sum_r nnz_r^2 products where the dense Gram costs n d^2 (about 1% of it at 8%
// density).
//
// Exact integer accumulation.  Column f carries a power-of-two scale 2^e_f >= max |delta_f|
// over the call's rows (a column-max pass).  A product delta_i delta_j 2^(F - e_i - e_j),
// |.| <= 2^F with F <= 50, is rounded to the nearest integer by one FMA against 1.5 * 2^52,
// and the integers are summed in int64 with LDS atomics.  Integer addition is associative,
// so the result is bit-identical whatever order the waves and workgroups add in.  Each
// product carries one rounding of at most 2^(e_i + e_j - F - 1); F is chosen per call so
// that a row slab's sum stays below 2^61 in magnitude.
//
// Kernels.
//   k_sg_len + hipcub scan: packed-row offsets P (row length + 1 for a ones entry).
//   k_sg_pack: one wave per row; entries bucketed by 128-feature block (a counting sort in
//     wave-private LDS), delta in FP64 and the feature as u16; per-row block offsets BO;
//     column maxima of |delta| (float, rounded up).  Feature d is a ones column (delta = 1):
//     its products with feature f give s_f.
//   k_sg_scale: every packed delta times 2^(F/2 - e_f), once the maxima are known.
//   k_sg_tile: one 1024-thread workgroup per (128 x 128 block tile of the upper triangle, row
//     slab); the tile accumulates in LDS as int64 [128][129].  Each wave takes 64 rows at a
//     time, ranks them by the length of their broadcast segment and multiplies them four
//     at a time: each 16-lane slot holds one row's shorter (at most 16-entry) segment, and
//     each step broadcasts one entry of the other segment to its slot with DPP
//     row_newbcast, so one ds_add_u64 carries up to 64 products.  The next group's loads
//     are in flight while a group multiplies.  A row with more than 16 entries in both
//     segments is cut into 16-entry pieces of its longer segment, which go through the
//     same groups.  Slab tiles go to memory as int64.
//   k_sg_colsum / k_sg_finish: slab tiles summed in FP64 in slab order, scales applied and the
//     rank-one terms added: G (full symmetric) and the column sums.
#include <hipcub/hipcub.hpp>

#include "sh_common.h"

namespace sh {

constexpr int kSgB = 128;          // feature block = tile edge
constexpr int kSgMaxBlocks = 63;   // d + 1 <= 8064 (lane nb of a wave holds a row's total)
constexpr int kSgCntStride = 64;   // per-wave block counters in k_sg_pack
constexpr int kSgLd = kSgB + 1;    // LDS row stride of the tile (int64 words): spreads banks
constexpr int kSgThreads = 1024;
constexpr int kSgWaves = kSgThreads / 64;
constexpr int kSgPackWaves = 4;
constexpr int kSgTileElems = kSgB * kSgB;
constexpr double kSgMagic = 6755399441055744.0;  // 1.5 * 2^52: ulp 1 over [2^52, 2^53)

// Exponent e with 2^e >= v for the float image of a non-negative column maximum (0 -> 0).
__device__ __forceinline__ int sg_col_exp(uint32_t bits) {
  const float v = __uint_as_float(bits);
  if (!(v > 0.f)) return 0;
  int e;
  (void)frexpf(v, &e);  // v = m 2^e, m in [0.5, 1)
  return e;
}
__device__ __forceinline__ bool sg_finite(uint32_t bits) {
  return (bits & 0x7f800000u) != 0x7f800000u;
}

// z of a zero count for feature f.
__device__ __forceinline__ double sg_z0(const double *center, const double *scale, double clip,
                                        int f) {
  double v = (0.0 - center[f]) * scale[f];
  if (clip > 0.0) v = fmin(fmax(v, -clip), clip);
  return v;
}

__device__ __forceinline__ void sg_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), lane);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Row lengths (+1 for the ones entry) for the offset scan; also clears the column maxima.
__global__ void k_sg_len(const int64_t *__restrict__ rowptr, const int64_t *__restrict__ rows,
                         int64_t n, int64_t *__restrict__ len, uint32_t *__restrict__ cm,
                         int ncm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncm) cm[i] = 0u;
  if (i > n) return;
  if (i == n) {
    len[n] = 0;
    return;
  }
  const int64_t r = rows ? rows[i] : i;
  len[i] = rowptr[r + 1] - rowptr[r] + 1;
}

__global__ void __launch_bounds__(64 * kSgPackWaves) k_sg_pack(
    const uint32_t *__restrict__ vals, const int32_t *__restrict__ cols,
    const int64_t *__restrict__ rowptr, const int64_t *__restrict__ rows, int64_t n,
    const int32_t *__restrict__ col_map, int32_t n_in, int32_t d, int nb, double sf, int log_t,
    const double *__restrict__ libsize, const double *__restrict__ center,
    const double *__restrict__ scale, double clip, const int64_t *__restrict__ P,
    int32_t *__restrict__ BO, double *__restrict__ dv, uint16_t *__restrict__ fc,
    uint32_t *__restrict__ cm) {
  extern __shared__ uint32_t sg_pack_lds[];
  const int dw = nb * kSgB;
  uint32_t *cmax = sg_pack_lds;                              // [dw]
  int *cnt = (int *)(sg_pack_lds + dw);                      // [waves][kSgCntStride]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int *my = cnt + w * kSgCntStride;
  for (int f = threadIdx.x; f < dw; f += blockDim.x) cmax[f] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) cmax[d] = __float_as_uint(1.0f);  // the ones column
  const int dblk = d / kSgB;
  const int nb1 = nb + 1;
  const int64_t stride = (int64_t)gridDim.x * kSgPackWaves;
  for (int64_t i = (int64_t)blockIdx.x * kSgPackWaves + w; i < n; i += stride) {
    const int64_t r = rows ? rows[i] : i;
    const int64_t a = rowptr[r], b = rowptr[r + 1];
    const double l = libsize[r];
    const int64_t base = P[i];
    if (lane < nb) my[lane] = lane == dblk ? 1 : 0;
    sg_wave_sync();
    for (int64_t t0 = a; t0 < b; t0 += 64) {
      const int64_t t = t0 + lane;
      if (t < b) {
        const int f = map_col(col_map, cols[t], n_in);
        if (f >= 0) atomicAdd(&my[f / kSgB], 1);
      }
    }
    sg_wave_sync();
    const int c = lane < nb ? my[lane] : 0;
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    const int excl = incl - c;  // lane nb: the row's entry count
    if (lane <= nb) BO[i * nb1 + lane] = excl;
    sg_wave_sync();
    if (lane < nb) my[lane] = excl;  // block cursors
    sg_wave_sync();
    for (int64_t t0 = a; t0 < b; t0 += 64) {
      const int64_t t = t0 + lane;
      if (t < b) {
        const int f = map_col(col_map, cols[t], n_in);
        if (f >= 0) {
          const double x = norm_value(vals[t], sf, l, log_t);
          double v = (x - center[f]) * scale[f];
          if (clip > 0.0) v = fmin(fmax(v, -clip), clip);
          const double delta = v - sg_z0(center, scale, clip, f);
          const int64_t pos = base + atomicAdd(&my[f / kSgB], 1);
          dv[pos] = delta;
          fc[pos] = (uint16_t)f;
          atomicMax(&cmax[f], __float_as_uint(__double2float_ru(fabs(delta))));
        }
      }
    }
    if (lane == 0) {
      const int64_t pos = base + atomicAdd(&my[dblk], 1);
      dv[pos] = 1.0;
      fc[pos] = (uint16_t)d;
    }
    sg_wave_sync();
  }
  __syncthreads();
  for (int f = threadIdx.x; f < dw; f += blockDim.x)
    if (cmax[f]) atomicMax(&cm[f], cmax[f]);
}

// Packed deltas times 2^(H - e_f), H = F / 2, once the column maxima are known: every
// product of two scaled deltas is then delta_i delta_j 2^(F - e_i - e_j), |.| <= 2^F.  Slots
// past a row's entries (input columns outside the map) hold no entry and are skipped or
// rescaled harmlessly; no kernel reads them.
__global__ void k_sg_scale(double *__restrict__ dv, const uint16_t *__restrict__ fc, int64_t total,
                           const uint32_t *__restrict__ cm, int ncm, int H) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int f = fc[t];
    if (f < ncm) dv[t] = ldexp(dv[t], H - sg_col_exp(cm[f]));
  }
}

// Lane b of each 16-lane row broadcast to the row (DPP row_newbcast; all lanes must be active).
template <int B>
__device__ __forceinline__ double sg_bcast64(double v) {
  return __longlong_as_double(
      __builtin_amdgcn_update_dpp(0ll, __double_as_longlong(v), 0x150 + B, 0xF, 0xF, true));
}
template <int B>
__device__ __forceinline__ int sg_bcast32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + B, 0xF, 0xF, true);
}

// A row's two segments in the tile: the "lane" side L (one entry per lane) and the
// "broadcast" side B (one entry per step, broadcast to the lanes).  The cell of a product
// is (J column) * 129 + (K column); mulL / mulB carry each side's byte stride so that either
// segment can take either role.
constexpr int kSgMulJ = kSgLd * 8, kSgMulK = 8;

// A wave's batch of 64 rows, ranked by broadcast length and laid out so that group G
// (0..15) is ranks 4G..4G+3: lane L holds rank 4 (L % 16) + L / 16, so the four rows of
// group G sit in lane G of each 16-lane slot.  After each group the batch rotates by one
// lane within the slots, so the next group's rows are always in lane 0 and one
// row_newbcast:0 hands each slot its row's segments.  Rows outside the group path (both
// segments longer than 16, or one empty) carry nL = nB = 0 here.
struct SgBatch {
  int64_t pL, pB;
  int nL, nB, mulL, mulB;
  int mB;  // largest nB over the lane's group (lanes L, L^16, L^32, L^48)
};

// One group: slot s = lane / 16 holds its row's lane-side entry a = lane % 16 and
// broadcast-side entry a of the first 16-entry chunk.  Values are the raw loads (consumed
// a group later); masking happens in sg_group_products.
struct SgGroup {
  double vL, vB;
  int fL, fB;
  int64_t pB;
  int nL, nB, mulL, mulB;  // the slot row's
  int mB;                  // largest nB of the group (uniform)
};

// lane (i + 1) % 16 of each 16-lane row -> lane i (DPP row_ror:15)
__device__ __forceinline__ int sg_rot32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x12F, 0xF, 0xF, false);
}
__device__ __forceinline__ int64_t sg_rot64(int64_t v) {
  const uint32_t lo = (uint32_t)sg_rot32((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)sg_rot32((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Takes the group in lane 0 of each slot, issues its loads, rotates the batch.
__device__ __forceinline__ void sg_fetch(SgBatch &r, SgGroup &g, const double *__restrict__ dv,
                                         const uint16_t *__restrict__ fc, int a) {
  const int64_t pL = __double_as_longlong(sg_bcast64<0>(__longlong_as_double(r.pL)));
  g.pB = __double_as_longlong(sg_bcast64<0>(__longlong_as_double(r.pB)));
  g.nL = sg_bcast32<0>(r.nL);
  g.nB = sg_bcast32<0>(r.nB);
  g.mulL = sg_bcast32<0>(r.mulL);
  g.mulB = sg_bcast32<0>(r.mulB);
  g.mB = __builtin_amdgcn_readfirstlane(r.mB);
  // unconditional loads (entry 0 always exists) keep the wait counts exact
  const int64_t iL = a < g.nL ? pL + a : 0, iB = a < g.nB ? g.pB + a : 0;
  g.vL = dv[iL];
  g.fL = fc[iL];
  g.vB = dv[iB];
  g.fB = fc[iB];
  r.pL = sg_rot64(r.pL);
  r.pB = sg_rot64(r.pB);
  r.nL = sg_rot32(r.nL);
  r.nB = sg_rot32(r.nB);
  r.mulL = sg_rot32(r.mulL);
  r.mulB = sg_rot32(r.mulB);
  r.mB = sg_rot32(r.mB);
}

// Adds the products of broadcast entry B of the chunk with the lanes' entries: the value
// and byte offset are broadcast from lane B of each 16-lane row.  Lanes whose row has no
// entry B (B >= nb) or no lane entry (act false) add nothing.
template <int B>
__device__ __forceinline__ void sg_kstep(unsigned long long *acc, double vL, double vB, int oL,
                                         int oB, bool act, int nb) {
  const double kb = sg_bcast64<B>(vB);
  const int o = oL + sg_bcast32<B>(oB);
  const double t = fma(vL, kb, kSgMagic);
  const long long q = __double_as_longlong(t) - __double_as_longlong(kSgMagic);
  if (act && B < nb)
    atomicAdd(reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(acc) + o),
              (unsigned long long)q);
}

// Steps 0..n-1 (n uniform, <= 16) of one broadcast chunk.
__device__ __forceinline__ void sg_chunk(unsigned long long *acc, double vL, double vB, int oL,
                                         int oB, bool act, int nb, int n) {
  sg_kstep<0>(acc, vL, vB, oL, oB, act, nb);
  if (n > 1) sg_kstep<1>(acc, vL, vB, oL, oB, act, nb);
  if (n > 2) sg_kstep<2>(acc, vL, vB, oL, oB, act, nb);
  if (n > 3) sg_kstep<3>(acc, vL, vB, oL, oB, act, nb);
  if (n > 4) sg_kstep<4>(acc, vL, vB, oL, oB, act, nb);
  if (n > 5) sg_kstep<5>(acc, vL, vB, oL, oB, act, nb);
  if (n > 6) sg_kstep<6>(acc, vL, vB, oL, oB, act, nb);
  if (n > 7) sg_kstep<7>(acc, vL, vB, oL, oB, act, nb);
  if (n > 8) sg_kstep<8>(acc, vL, vB, oL, oB, act, nb);
  if (n > 9) sg_kstep<9>(acc, vL, vB, oL, oB, act, nb);
  if (n > 10) sg_kstep<10>(acc, vL, vB, oL, oB, act, nb);
  if (n > 11) sg_kstep<11>(acc, vL, vB, oL, oB, act, nb);
  if (n > 12) sg_kstep<12>(acc, vL, vB, oL, oB, act, nb);
  if (n > 13) sg_kstep<13>(acc, vL, vB, oL, oB, act, nb);
  if (n > 14) sg_kstep<14>(acc, vL, vB, oL, oB, act, nb);
  if (n > 15) sg_kstep<15>(acc, vL, vB, oL, oB, act, nb);
}

__device__ __forceinline__ void sg_group_products(unsigned long long *acc, SgGroup &g,
                                                  const double *__restrict__ dv,
                                                  const uint16_t *__restrict__ fc, int a) {
  if (g.mB <= 0) return;
  // the loads must land here, not where they were issued: keeps the arithmetic below
  // from being hoisted next to them
  asm volatile("" : "+v"(g.vL), "+v"(g.vB), "+v"(g.fL), "+v"(g.fB));
  const bool act = a < g.nL;
  const int oL = (g.fL & (kSgB - 1)) * g.mulL;
  sg_chunk(acc, g.vL, a < g.nB ? g.vB : 0.0, oL, (g.fB & (kSgB - 1)) * g.mulB, act, g.nB,
           min(g.mB, 16));
  // broadcast sides longer than 16: further chunks, loaded here
  for (int c = 16; c < g.mB; c += 16) {
    const int nb = g.nB - c;
    const int64_t iB = a < nb ? g.pB + c + a : 0;
    const double vB = dv[iB];
    const int fB = fc[iB];
    sg_chunk(acc, g.vL, a < nb ? vB : 0.0, oL, (fB & (kSgB - 1)) * g.mulB, act, nb,
             min(g.mB - c, 16));
  }
}

// Multiplies a wave's batch of up to 64 rows given in natural lane order (lane = row; nL = 0
// for a lane without a row; every row's lane segment holds at most 16 entries): ranks the
// rows by broadcast length (bitonic sort of (127 - nB, lane) across the wave) so that each
// group of four holds rows of similar length (a group takes as many steps as its longest
// broadcast segment), lays them out by group and runs the groups that hold rows with one
// group's loads in flight while the previous one multiplies.
__device__ __forceinline__ void sg_run_batch(unsigned long long *acc, int64_t rpL, int64_t rpB,
                                             int rnL, int rnB, int rmL, int rmB,
                                             const double *__restrict__ dv,
                                             const uint16_t *__restrict__ fc, int lane) {
  const int a = lane & 15;
  uint32_t key = ((uint32_t)(127 - min(rnL > 0 ? rnB : 0, 127)) << 6) | (uint32_t)lane;
#pragma unroll
  for (int kk = 2; kk <= 64; kk <<= 1) {
#pragma unroll
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      const uint32_t o = __shfl_xor(key, jj, 64);
      const bool up = (lane & kk) == 0, lower = (lane & jj) == 0;
      key = (lower == up) ? min(key, o) : max(key, o);
    }
  }
  // lane 16 s + G takes the row of rank 4 G + s
  const int src = (int)(__shfl(key, 4 * (lane & 15) + (lane >> 4), 64) & 63u);
  SgBatch bt;
  bt.pL = __shfl(rpL, src, 64);
  bt.pB = __shfl(rpB, src, 64);
  bt.nL = __shfl(rnL, src, 64);
  bt.nB = __shfl(rnL > 0 ? rnB : 0, src, 64);
  bt.mulL = __shfl(rmL, src, 64);
  bt.mulB = __shfl(rmB, src, 64);
  const int m = max(bt.nB, __shfl_xor(bt.nB, 16, 64));
  bt.mB = max(m, __shfl_xor(m, 32, 64));
  // rows are ranked first, empty lanes last: only the first ng groups hold rows
  const int ng = (int)((__popcll(__ballot(rnL > 0)) + 3) >> 2);
  if (ng == 0) return;
  SgGroup ga, gb;
  sg_fetch(bt, ga, dv, fc, a);
  for (int G = 1;; G += 2) {
    if (G == ng) {
      sg_group_products(acc, ga, dv, fc, a);
      break;
    }
    sg_fetch(bt, gb, dv, fc, a);
    sg_group_products(acc, ga, dv, fc, a);
    if (G + 1 == ng) {
      sg_group_products(acc, gb, dv, fc, a);
      break;
    }
    sg_fetch(bt, ga, dv, fc, a);
    sg_group_products(acc, gb, dv, fc, a);
  }
}

__global__ void __launch_bounds__(kSgThreads) k_sg_tile(
    const int64_t *__restrict__ P, const int32_t *__restrict__ BO,
    const double *__restrict__ dv, const uint16_t *__restrict__ fc, int64_t n, int nb, int T,
    int S, int64_t rows_per_slab, long long *__restrict__ part) {
  __shared__ unsigned long long acc[kSgB * kSgLd];
  const int b = blockIdx.x;
  int slab, t;
  if (S % 8 == 0) {  // the slab's tiles on one XCD: its rows are fetched once per L2
    const int x = b & 7, j = b >> 3;
    slab = x + 8 * (j / T);
    t = j % T;
  } else {
    slab = b / T;
    t = b % T;
  }
  if (slab >= S) return;
  int bi, bj;
  tile_coords(t, nb, bi, bj);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int e = tid; e < kSgB * kSgLd; e += kSgThreads) acc[e] = 0ull;
  __syncthreads();
  const int nb1 = nb + 1;
  const int64_t r0 = (int64_t)slab * rows_per_slab;
  const int64_t r1 = min(n, r0 + rows_per_slab);

  for (int64_t rb = r0 + (int64_t)w * 64; rb < r1; rb += (int64_t)kSgWaves * 64) {
    const int64_t i = rb + lane;
    int64_t pj = 0, pk = 0;
    int nj = 0, nk = 0;
    if (i < r1) {
      const int64_t p = P[i];
      const int32_t *bo = BO + i * nb1;
      const int j0 = bo[bi], j1 = bo[bi + 1], k0 = bo[bj], k1 = bo[bj + 1];
      pj = p + j0;
      nj = j1 - j0;
      pk = p + k0;
      nk = k1 - k0;
    }
    const bool live = nj > 0 && nk > 0;
    const bool grouped = live && min(nj, nk) <= 16;
    uint64_t single = __ballot(live && !grouped);
    // an atomic instruction costs the same however many lanes it carries, so the longer
    // segment that fits in 16 lanes goes to the lanes and the shorter one is broadcast
    const bool swap = nj > 16 || (nk <= 16 && nk > nj);
    if (__ballot(grouped))
      sg_run_batch(acc, swap ? pk : pj, swap ? pj : pk, grouped ? (swap ? nk : nj) : 0,
                   swap ? nj : nk, swap ? kSgMulK : kSgMulJ, swap ? kSgMulJ : kSgMulK, dv, fc,
                   lane);
    // rows with both segments longer than 16: the longer segment is cut into 16-entry
    // pieces, each a row of its own with the whole shorter segment broadcast, and those
    // rows go through the same ranked groups (64 at a time)
    int64_t vpL = 0, vpB = 0;
    int vnL = 0, vnB = 0, vmL = 0, vmB = 0, nv = 0;
    while (single) {
      const int q = __ffsll((unsigned long long)single) - 1;
      single &= single - 1;
      const int64_t spj = readlane_i64(pj, q), spk = readlane_i64(pk, q);
      const int snj = __builtin_amdgcn_readlane(nj, q), snk = __builtin_amdgcn_readlane(nk, q);
      const bool jl = snj >= snk;
      const int64_t sL = jl ? spj : spk, sB = jl ? spk : spj;
      const int nL = jl ? snj : snk, nB = jl ? snk : snj;
      const int mL = jl ? kSgMulJ : kSgMulK, mB = jl ? kSgMulK : kSgMulJ;
      for (int c = 0; c < nL; c += 16) {
        if (nv == 64) {
          sg_run_batch(acc, vpL, vpB, vnL, vnB, vmL, vmB, dv, fc, lane);
          vnL = 0;
          nv = 0;
        }
        if (lane == nv) {
          vpL = sL + c;
          vnL = min(16, nL - c);
          vpB = sB;
          vnB = nB;
          vmL = mL;
          vmB = mB;
        }
        ++nv;
      }
    }
    if (nv) sg_run_batch(acc, vpL, vpB, vnL, vnB, vmL, vmB, dv, fc, lane);
  }
  __syncthreads();
  long long *dst = part + ((int64_t)slab * T + t) * kSgTileElems;
  for (int e = tid; e < kSgTileElems; e += kSgThreads)
    dst[e] = (long long)acc[(e / kSgB) * kSgLd + (e % kSgB)];
}

__device__ __forceinline__ int sg_tile_index(int bi, int bj, int nb) {
  return bi * nb - bi * (bi - 1) / 2 + (bj - bi);
}

__device__ __forceinline__ double sg_cell(const long long *__restrict__ part, int S, int T, int t,
                                          int e) {
  double s = 0.0;
  for (int sl = 0; sl < S; ++sl) s += (double)part[((int64_t)sl * T + t) * kSgTileElems + e];
  return s;
}

// s_f = column sums of delta (the ones column's products), and the column sums of z.
__global__ void k_sg_colsum(const long long *__restrict__ part, const uint32_t *__restrict__ cm,
                            int S, int T, int nb, int32_t d, int F, double n_rows,
                            const double *__restrict__ center, const double *__restrict__ scale,
                            double clip, int accumulate, double *__restrict__ sbuf,
                            double *__restrict__ colsum) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d) return;
  const int t = sg_tile_index(f / kSgB, d / kSgB, nb);
  const double A = sg_cell(part, S, T, t, (f % kSgB) * kSgB + (d % kSgB));
  double s = ldexp(A, sg_col_exp(cm[f]) + sg_col_exp(cm[d]) - F);
  if (!sg_finite(cm[f])) s = __builtin_nan("");
  sbuf[f] = s;
  const double v = s + n_rows * sg_z0(center, scale, clip, f);
  colsum[f] = accumulate ? colsum[f] + v : v;
}

__global__ void __launch_bounds__(256) k_sg_finish(
    const long long *__restrict__ part, const uint32_t *__restrict__ cm, int S, int T, int nb,
    int32_t d, int F, double n_rows, const double *__restrict__ center,
    const double *__restrict__ scale, double clip, const double *__restrict__ sbuf,
    int accumulate, double *__restrict__ G) {
  const int t = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= kSgTileElems) return;
  int bi, bj;
  tile_coords(t, nb, bi, bj);
  const int m = e / kSgB, nn = e % kSgB;
  const int i = bi * kSgB + m, j = bj * kSgB + nn;
  if (i >= d || j >= d) return;
  if (bi == bj && m > nn) return;  // diagonal tiles: upper half only
  const double A = sg_cell(part, S, T, t, e);
  const double D = ldexp(A, sg_col_exp(cm[i]) + sg_col_exp(cm[j]) - F);
  const double ci = sg_z0(center, scale, clip, i), cj = sg_z0(center, scale, clip, j);
  double g = ((D + ci * sbuf[j]) + sbuf[i] * cj) + (n_rows * ci) * cj;
  if (!sg_finite(cm[i]) || !sg_finite(cm[j])) g = __builtin_nan("");
  if (accumulate) g += G[(int64_t)i * d + j];
  G[(int64_t)i * d + j] = g;
  G[(int64_t)j * d + i] = g;
}

struct SgPlan {
  int nb, T, S, F;
  int64_t rows_per_slab, cap;
  size_t tmp_bytes;
  size_t off_len, off_P, off_tmp, off_BO, off_dv, off_fc, off_cm, off_part, off_s, total;
};

static int ceil_log2(int64_t v) {
  int k = 0;
  while (((int64_t)1 << k) < v) ++k;
  return k;
}

static SgPlan sg_plan(int64_t n, int64_t nnz, int32_t d) {
  SgPlan p;
  p.nb = (int)cdiv((int64_t)d + 1, kSgB);
  p.T = p.nb * (p.nb + 1) / 2;
  const int64_t nn = n > 0 ? n : 1;
  // about 2048 workgroups (one per CU at a time: the tile takes 129 KB of LDS), slabs of at
  // least 1024 rows
  int64_t S = cdiv(2048, p.T);
  if (S >= 8) S = (int64_t)align_up((size_t)S, 8);
  const int64_t max_s = cdiv(nn, 1024);
  if (S > max_s) S = max_s;
  if (S < 1) S = 1;
  p.rows_per_slab = cdiv(nn, S);
  p.S = (int)cdiv(nn, p.rows_per_slab);
  int F = 61 - ceil_log2(p.rows_per_slab);
  F = F > 50 ? 50 : F;
  p.F = F & ~1;  // even: both factors of a product carry 2^(F / 2)
  p.cap = nnz + n;
  size_t tb = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (const int64_t *)nullptr,
                                         (int64_t *)nullptr, (int)(nn + 1));
  p.tmp_bytes = tb;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes, 256);
    return at;
  };
  p.off_len = take((size_t)(nn + 1) * sizeof(int64_t));
  p.off_P = take((size_t)(nn + 1) * sizeof(int64_t));
  p.off_tmp = take(tb);
  p.off_BO = take((size_t)nn * (p.nb + 1) * sizeof(int32_t));
  p.off_dv = take((size_t)(p.cap > 0 ? p.cap : 1) * sizeof(double));
  p.off_fc = take((size_t)(p.cap > 0 ? p.cap : 1) * sizeof(uint16_t));
  p.off_cm = take((size_t)p.nb * kSgB * sizeof(uint32_t));
  p.off_part = take((size_t)p.S * p.T * kSgTileElems * sizeof(long long));
  p.off_s = take((size_t)d * sizeof(double));
  p.total = o;
  return p;
}

}  // namespace sh

using namespace sh;

extern "C" size_t sh_sparse_gram_workspace(int64_t n_rows, int64_t nnz, int32_t d) {
  if (n_rows < 0 || nnz < 0 || d <= 0 || (int64_t)d + 1 > (int64_t)kSgB * kSgMaxBlocks ||
      n_rows >= ((int64_t)1 << 31) - 1)
    return 0;
  return sg_plan(n_rows, nnz, d).total;
}

extern "C" int sh_sparse_gram(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                              const int64_t *rows, int64_t n_rows, int64_t nnz,
                              const int32_t *col_map, int32_t n_cols_in, int32_t d, double sf,
                              int log_transform, const double *libsize, const double *center,
                              const double *scale, double clip, int accumulate, double *G_out,
                              double *colsum_out, void *workspace, size_t workspace_bytes,
                              sh_stream_t stream) {
  SH_ARG(n_rows >= 0 && nnz >= 0 && d > 0 && n_cols_in >= 0,
         "sh_sparse_gram: bad sizes n_rows=%lld nnz=%lld d=%d", (long long)n_rows,
         (long long)nnz, d);
  SH_ARG((int64_t)d + 1 <= (int64_t)kSgB * kSgMaxBlocks,
         "sh_sparse_gram: d=%d exceeds %d features (use sh_densify + sh_gram)", d,
         kSgB * kSgMaxBlocks - 1);
  SH_ARG(n_rows < ((int64_t)1 << 31) - 1, "sh_sparse_gram: n_rows=%lld too large for one call",
         (long long)n_rows);
  SH_ARG(col_map || d == n_cols_in, "sh_sparse_gram: col_map NULL requires d == n_cols_in");
  SH_ARG(G_out && colsum_out && center && scale, "sh_sparse_gram: NULL pointer");
  SH_ARG(n_rows == 0 || (vals && cols && rowptr && libsize),
         "sh_sparse_gram: NULL CSR pointer");
  hipStream_t st = (hipStream_t)stream;
  const SgPlan p = sg_plan(n_rows, nnz, d);
  SH_ARG(workspace && workspace_bytes >= p.total,
         "sh_sparse_gram: workspace of %zu bytes, %zu needed", workspace_bytes, p.total);
  char *ws = (char *)workspace;
  int64_t *len = (int64_t *)(ws + p.off_len);
  int64_t *P = (int64_t *)(ws + p.off_P);
  int32_t *BO = (int32_t *)(ws + p.off_BO);
  double *dv = (double *)(ws + p.off_dv);
  uint16_t *fc = (uint16_t *)(ws + p.off_fc);
  uint32_t *cm = (uint32_t *)(ws + p.off_cm);
  long long *part = (long long *)(ws + p.off_part);
  double *sbuf = (double *)(ws + p.off_s);
  const int ncm = p.nb * kSgB;
  const int64_t nl = (n_rows + 1 > ncm ? n_rows + 1 : ncm);
  hipLaunchKernelGGL(k_sg_len, dim3(cdiv(nl, 256)), dim3(256), 0, st, rowptr, rows, n_rows, len,
                     cm, ncm);
  SH_LAUNCH_CHECK("k_sg_len");
  size_t tb = p.tmp_bytes;
  SH_HIP(hipcub::DeviceScan::ExclusiveSum(ws + p.off_tmp, tb, len, P, (int)(n_rows + 1), st));
  int64_t total = 0;
  SH_HIP(hipMemcpyAsync(&total, P + n_rows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SH_HIP(hipStreamSynchronize(st));
  SH_ARG(total <= p.cap,
         "sh_sparse_gram: the rows hold %lld stored entries, more than nnz=%lld",
         (long long)(total - n_rows), (long long)nnz);
  if (n_rows > 0) {
    const int64_t grid = cdiv(n_rows, kSgPackWaves) < 2048 ? cdiv(n_rows, kSgPackWaves) : 2048;
    const size_t lds = ((size_t)ncm + (size_t)kSgPackWaves * kSgCntStride) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_sg_pack, dim3(grid), dim3(64 * kSgPackWaves), lds, st, vals, cols,
                       rowptr, rows, n_rows, col_map, n_cols_in, d, p.nb, sf, log_transform,
                       libsize, center, scale, clip, P, BO, dv, fc, cm);
    SH_LAUNCH_CHECK("k_sg_pack");
  }
  if (total > 0) {
    const int64_t sgrid = cdiv(total, 256) < 8192 ? cdiv(total, 256) : 8192;
    hipLaunchKernelGGL(k_sg_scale, dim3(sgrid), dim3(256), 0, st, dv, fc, total, cm, ncm,
                       p.F / 2);
    SH_LAUNCH_CHECK("k_sg_scale");
  }
  hipLaunchKernelGGL(k_sg_tile, dim3(p.T * p.S), dim3(kSgThreads), 0, st, P, BO, dv, fc,
                     n_rows, p.nb, p.T, p.S, p.rows_per_slab, part);
  SH_LAUNCH_CHECK("k_sg_tile");
  hipLaunchKernelGGL(k_sg_colsum, dim3(cdiv(d, 256)), dim3(256), 0, st, part, cm, p.S, p.T, p.nb,
                     d, p.F, (double)n_rows, center, scale, clip, accumulate, sbuf, colsum_out);
  SH_LAUNCH_CHECK("k_sg_colsum");
  hipLaunchKernelGGL(k_sg_finish, dim3(kSgTileElems / 256, p.T), dim3(256), 0, st, part, cm, p.S,
                     p.T, p.nb, d, p.F, (double)n_rows, center, scale, clip, sbuf, accumulate,
                     G_out);
  SH_LAUNCH_CHECK("k_sg_finish");
  return SH_OK;
}
