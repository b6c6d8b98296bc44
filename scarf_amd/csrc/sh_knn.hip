// Exact k-nearest-neighbour search (squared L2) on MFMA with an error-bounded
// low-precision screen.  Replaces scarf's hnswlib index (ann.py:14-28, 194-205,
// 290-326) and fix_knn_query (ann.py:31-52).
//
// Screen: the MFMA computes s(q, r) = ||r||^2 - 2 q.r directly: reference rows are
// packed as [r, ||r||^2 (split over 1 or 3 slots), 0...] and query rows as
// [-2q, 1, ..., 0...], so one K-padded dot product gives the ranking key
// (||q||^2 is the same for every reference of a query and is left out).
//   A operand = 32 references, B operand = 32 queries, D = 32 x 32 FP32 tile: lane l
//   holds query l%32 and 16 references 8*(v/4) + 4*(l/32) + v%4.
// The K axis is split between the two lane halves: half h supplies elements
// [h*K/2, (h+1)*K/2) of its row, for both operands, so each lane streams one
// contiguous half-row (the MFMA pairs A and B by the same (half, element) slot,
// so any such split gives the same dot product).
//
// Selection: per (query, reference slice) a candidate buffer of CAP (key, index)
// pairs in global memory and a threshold tau = the pool-th best key so far.  A
// 32 x 32 tile is filtered with a 16-value min tree and one compare per query;
// only tiles with a hit (rare once tau has settled) take the insertion path.  When
// a buffer nears CAP the owning wave sorts it (bitonic across the wave) and keeps
// the best `pool`.  Each wave owns its queries, so no cross-wave synchronisation.
// Optionally each query block first screens a set of seed reference rows (for self
// queries in locality order: the block's own rows) and starts tau at their pool-th key.
#include <float.h>

#include <type_traits>

#include "sh_common.h"
#include "sh_wave_sort.h"

namespace sh {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

int32_t pack_k(int dtype, int32_t D) {
  if (dtype == SH_F32) return (int32_t)align_up((size_t)D + 1, 8);
  if (dtype == SH_BF16) return (int32_t)align_up((size_t)D + 3, 16);
  if (dtype == SH_FP8) return (int32_t)align_up((size_t)D, 64);
  return -1;
}

// Bytes per packed row: K operand elements, plus (FP8) a 16-byte trailer holding the
// row's ||r||^2 (f32) and its E8M0 scale exponent (biased, i32).
int64_t pack_row_bytes(int dtype, int32_t D) {
  const int32_t K = pack_k(dtype, D);
  if (K < 0) return -1;
  if (dtype == SH_F32) return (int64_t)K * 4;
  if (dtype == SH_BF16) return (int64_t)K * 2;
  return (int64_t)K + 16;
}

// float -> OCP FP8 E4M3 (bias 7, max 448, no infinities), round to nearest even,
// saturating at +-448.
__device__ __forceinline__ uint8_t f2e4m3(float x) {
  const uint32_t sign = (__float_as_uint(x) >> 24) & 0x80u;
  const float a = fabsf(x);
  if (!(a == a)) return (uint8_t)(sign | 0x7Fu);
  if (a >= 464.0f) return (uint8_t)(sign | 0x7Eu);
  if (a < 0.015625f) {  // subnormal: multiples of 2^-9 (8 -> the smallest normal)
    const int m = (int)rintf(a * 512.0f);
    return (uint8_t)(sign | (uint32_t)m);
  }
  int e;
  frexpf(a, &e);                              // a = f * 2^e, f in [0.5, 1)
  int E = e - 1 + 7;                          // biased exponent of 1.m * 2^(e-1)
  int m = (int)rintf((ldexpf(a, 1 - e) - 1.0f) * 8.0f);
  if (m == 8) {
    m = 0;
    ++E;
  }
  if (E > 15 || (E == 15 && m == 7)) return (uint8_t)(sign | 0x7Eu);
  return (uint8_t)(sign | ((uint32_t)E << 3) | (uint32_t)m);
}

// FP8 rows: x / 2^e in E4M3 with e the smallest power of two that brings max|x|
// under 448; role 0 packs r and its ||r||^2, role 1 packs -2q.
__global__ void k_pack_fp8(const float *__restrict__ E, int64_t n, int32_t D, int64_t lde,
                           int role, int32_t K, uint8_t *__restrict__ out,
                           float *__restrict__ aux) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float *e = E + r * lde;
  uint8_t *o = out + r * (int64_t)(K + 16);
  double s = 0.0;
  float amax = 0.f;
  for (int t = 0; t < D; ++t) {
    const float x = e[t];
    s += (double)x * (double)x;
    amax = fmaxf(amax, fabsf(role == 0 ? x : 2.0f * x));
  }
  int ex = 0;
  if (amax > 0.f) {
    ex = (int)ceilf(log2f(amax / 448.0f));
    while (ldexpf(amax, -ex) > 448.0f) ++ex;
    ex = max(-126, min(126, ex));
  }
  for (int t = 0; t < D; ++t) {
    const float x = role == 0 ? e[t] : -2.0f * e[t];
    o[t] = f2e4m3(ldexpf(x, -ex));
  }
  for (int t = D; t < K; ++t) o[t] = 0;
  float *tf = reinterpret_cast<float *>(o + K);
  int32_t *ti = reinterpret_cast<int32_t *>(o + K);
  tf[0] = role == 0 ? (float)s : 0.f;
  ti[1] = ex + 127;  // E8M0 biased exponent
  ti[2] = 0;
  ti[3] = 0;
  if (aux) aux[r] = (float)s;
}

// ---------------------------------------------------------------------------------
// Packing: one thread per row (rows are short: K <= 136 elements).
__global__ void k_pack_f32(const float *__restrict__ E, int64_t n, int32_t D, int64_t lde,
                           int role, int32_t K, float *__restrict__ out, float *__restrict__ aux) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float *e = E + r * lde;
  float *o = out + r * K;
  double s = 0.0;
  for (int t = 0; t < D; ++t) {
    float x = e[t];
    s += (double)x * (double)x;
    o[t] = role == 0 ? x : -2.0f * x;
  }
  o[D] = role == 0 ? (float)s : 1.0f;
  for (int t = D + 1; t < K; ++t) o[t] = 0.f;
  if (aux) aux[r] = (float)s;
}

// BF16 rows, up to 128 per workgroup staged through LDS (at most 64 KiB): the FP32 rows are
// read coalesced, each thread packs its own row (the squared norm summed over the row in
// order, as the FP64 value the three-term split needs), and the packed rows are written
// back coalesced.
constexpr int kPackRows = 128;
__global__ void __launch_bounds__(kPackRows) k_pack_bf16(const float *__restrict__ E, int64_t n,
                                                        int32_t D, int64_t lde, int role,
                                                        int32_t K, uint16_t *__restrict__ out,
                                                        float *__restrict__ aux) {
  extern __shared__ float pack_lds[];
  const int RB = blockDim.x;  // rows per workgroup (<= kPackRows)
  float *sE = pack_lds;                                 // [RB][D]
  uint16_t *sO = (uint16_t *)(pack_lds + RB * D);       // [RB][K]
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * RB;
  const int rows = (int)min<int64_t>(RB, n - r0);
  for (int e = tid; e < rows * D; e += RB) {
    const int rr = e / D, t = e % D;
    sE[e] = E[(r0 + rr) * lde + t];
  }
  __syncthreads();
  if (tid < rows) {
    const float *e = sE + tid * D;
    uint16_t *o = sO + tid * K;
    double s = 0.0;
    for (int t = 0; t < D; ++t) {
      float x = e[t];
      s += (double)x * (double)x;
      o[t] = role == 0 ? f2bf(x) : f2bf(-2.0f * x);
    }
    if (role == 0) {
      // ||r||^2 as an unevaluated sum of three bf16 values (24 significant bits)
      float hi = bf2f(f2bf((float)s));
      double r1 = s - (double)hi;
      float mid = bf2f(f2bf((float)r1));
      double r2 = r1 - (double)mid;
      float lo = bf2f(f2bf((float)r2));
      o[D] = f2bf(hi);
      o[D + 1] = f2bf(mid);
      o[D + 2] = f2bf(lo);
    } else {
      o[D] = o[D + 1] = o[D + 2] = f2bf(1.0f);
    }
    for (int t = D + 3; t < K; ++t) o[t] = 0;
    if (aux) aux[r0 + tid] = (float)s;
  }
  __syncthreads();
  // K is even: write 32-bit pairs
  const uint32_t *s32 = (const uint32_t *)sO;
  uint32_t *o32 = (uint32_t *)(out + r0 * K);
  for (int e = tid; e < rows * K / 2; e += RB) o32[e] = s32[e];
}

// ---------------------------------------------------------------------------------
// Screen kernel.
struct OpF32 {
  typedef float Frag;
  static constexpr int EPL = 1;  // elements per lane per MFMA
  static constexpr int ESZ = 4;
  static constexpr bool kScaled = false;
  static __device__ __forceinline__ floatx16 mma(Frag a, Frag b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
};
struct OpBF16 {
  typedef bf16x8 Frag;
  static constexpr int EPL = 8;
  static constexpr int ESZ = 2;
  static constexpr bool kScaled = false;
  static __device__ __forceinline__ floatx16 mma(Frag a, Frag b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
// FP8 E4M3 with per-row E8M0 scales (MX-style, one scale per lane's 32-element block):
// the norms enter through the accumulator, read from each row's 16-byte trailer.
typedef int intx8 __attribute__((ext_vector_type(8)));
struct OpFP8 {
  typedef intx8 Frag;
  static constexpr int EPL = 32;
  static constexpr int ESZ = 1;
  static constexpr bool kScaled = true;
  static __device__ __forceinline__ floatx16 mma_s(Frag a, Frag b, floatx16 c, int sa, int sb) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
  }
  static __device__ __forceinline__ floatx16 mma(Frag a, Frag b, floatx16 c) {
    return mma_s(a, b, c, 127, 127);
  }
};


constexpr int kSliceAlign = 128;       // reference slices start on this multiple
constexpr float kOpen = FLT_MAX;       // threshold of a pool that is not full yet

// QG: 32-query groups per wave; MINB: resident workgroups per CU the kernel is compiled
// for (its register budget).  Two groups at three workgroups per CU measured faster than
// four groups at two (BF16 K=64, 1M x 1M: 215 ms against 231 ms) and than one group at
// four (176 ms against 184 ms).
template <class Op, int K>
struct ScreenShape {
  static constexpr int QG = 2;
  static constexpr int QPB = 4 * 32 * QG;  // queries per workgroup
  static constexpr int MINB = 3;
};
// Register budget of one instantiation: three workgroups per CU (168 VGPRs) unless the
// query fragments or the compaction sort need more -- the FP8 screen at K >= 128 and the
// 1024-key buffers -- which then run two per CU instead of spilling to scratch.  The BF16
// screen at K <= 64 with 256-key buffers runs four per CU (128 VGPRs; its spills fall
// outside the tile loop): 166.6 against 174.6 ms at 1M x 1M, alternated runs on one GPU.
// The FP8 screen at K = 128 with 512-key buffers needs 165 VGPRs with the late index add
// (append_quad) and so runs three per CU instead of two.
template <class Op, int K, int CAP>
constexpr bool screen_late() {
  return (std::is_same<Op, OpBF16>::value && K <= 64 && CAP <= 256) ||
         (std::is_same<Op, OpFP8>::value && K > 64 && K <= 128 && CAP == 512);
}
template <class Op, int K, int CAP>
constexpr int screen_minb() {
  if constexpr (std::is_same<Op, OpBF16>::value && K <= 64 && CAP <= 256) return 4;
  if constexpr (std::is_same<Op, OpFP8>::value && K > 64 && K <= 128 && CAP == 512) return 3;
  return (Op::kScaled && K >= 128) || CAP >= 1024 ? 2 : ScreenShape<Op, K>::MINB;
}

// Candidate buffer of one (query, slice): CAP keys in two halves of CAP / 2, one per lane
// half (lanes n and n + 32 both serve query n); each lane appends to its own half with a
// register counter, so inserting needs no atomics.  Compaction (the whole wave, for one
// query) sorts both halves' c0 + c1 keys, keeps the best `pool` split between the two
// halves (or writes them to `out`), and returns the new threshold: the pool-th key once
// the query has seen `pool` candidates, else kOpen.
template <int CAP>
__device__ __forceinline__ int compact_pair(uint64_t *base, int c0, int c1, int pool,
                                            uint64_t *out, int32_t *out_cnt, float &tau_new) {
  constexpr int E = CAP / 64;
  constexpr int HALF = CAP / 2;
  const int lane = threadIdx.x & 63;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  uint64_t v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int e = lane * E + i;
    v[i] = e < c0 ? base[e] : (e - c0 < c1 ? base[HALF + (e - c0)] : ~0ull);
  }
  wave_bitonic_sort<E>(v);
  const int total = c0 + c1;
  const int keep = total < pool ? total : pool;
  const int k0 = (keep + 1) / 2;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int e = lane * E + i;
    if (e < keep) {
      if (out)
        out[e] = v[i];
      else
        base[e < k0 ? e : HALF + (e - k0)] = v[i];
    }
  }
  const uint64_t kp = wave_elem_u64<E>(v, pool - 1);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  if (out_cnt && lane == 0) *out_cnt = keep;
  tau_new = total >= pool ? ord_float((uint32_t)(kp >> 32)) : kOpen;
  return keep;
}

// True when the predicate holds on any lane of the wave (one ballot, no lane shuffles).
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Appends this lane's values of one 4-reference quad that pass tau to its buffer half.
// Queries past the batch end carry tau = -inf, so nothing of theirs passes.
// LATE: the reference index is added inside the store (an asm statement the compiler
// cannot hoist).  Hoisted, a tile that takes this path computes all 16 indices and keeps
// them in registers, which the four-workgroups-per-CU budget cannot spare; at three per CU
// the hoisted form is faster.
template <int V, bool LATE>
__device__ __forceinline__ void append_one(const floatx16 &acc, float tau, uint32_t ref0,
                                           uint64_t *half, int &cnt) {
  constexpr int kOff = 8 * (V >> 2) + (V & 3);  // the value's reference within the tile
  if (acc[V] <= tau) {
    uint32_t ref;
    if constexpr (LATE)
      asm volatile("v_add_u32 %0, %1, %2" : "=v"(ref) : "i"(kOff), "v"(ref0));
    else
      ref = ref0 + (uint32_t)kOff;
    half[cnt] = make_key(acc[V], ref);
    ++cnt;
  }
}

template <int V0, bool LATE>
__device__ __forceinline__ void append_quad(const floatx16 &acc, float tau, uint32_t ref0,
                                            uint64_t *half, int &cnt) {
  append_one<V0, LATE>(acc, tau, ref0, half, cnt);
  append_one<V0 + 1, LATE>(acc, tau, ref0, half, cnt);
  append_one<V0 + 2, LATE>(acc, tau, ref0, half, cnt);
  append_one<V0 + 3, LATE>(acc, tau, ref0, half, cnt);
}

template <class Op, int K, int CAP>
__global__ void __launch_bounds__(256, (screen_minb<Op, K, CAP>())) k_screen(
    const typename Op::Frag *__restrict__ Qp_, int64_t nq, const unsigned char *__restrict__ Rp,
    int64_t nr, int64_t slice_len, int S, int pool, const int64_t *__restrict__ seeds,
    uint64_t *__restrict__ buf, uint64_t *__restrict__ out, int32_t *__restrict__ out_cnt) {
  typedef typename Op::Frag Frag;
  constexpr int QG = ScreenShape<Op, K>::QG;
  constexpr bool kLate = screen_late<Op, K, CAP>();  // see append_quad
  constexpr int QPW = 32 * QG;                // queries per wave
  constexpr int QPB = ScreenShape<Op, K>::QPB;
  constexpr int HALFC = CAP / 2;              // keys per buffer half
  constexpr int HALF = K / 2;                 // elements per lane half
  constexpr int NS = HALF / Op::EPL;          // MFMA steps per tile
  constexpr int GROWB = K * Op::ESZ + (Op::kScaled ? 16 : 0);  // packed row bytes in HBM
  constexpr int ROWB = K * Op::ESZ + 16;      // LDS row bytes (pad or FP8 trailer)
  constexpr int CHUNKS = GROWB / 16;          // 16-byte chunks per packed row
  // references per LDS stage: as many as two stage buffers fit in 48 KiB
  constexpr int RT = (2 * 128 * ROWB <= 49152) ? 128 : (2 * 64 * ROWB <= 49152) ? 64 : 32;
  // stage buffers are filled by global_load_lds_dwordx4: each wave instruction writes
  // 64 consecutive 16-byte chunks (1 KiB), so a buffer is a whole number of KiB
  constexpr int CPR = ROWB / 16;                      // chunks per padded LDS row
  constexpr int NINSTR = (RT * CPR + 63) / 64;        // wave instructions per stage
  constexpr int BUFB = NINSTR * 1024;
  __shared__ __attribute__((aligned(1024))) unsigned char lds_r[2][BUFB];

  const unsigned char *Qp = reinterpret_cast<const unsigned char *>(Qp_);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, n = lane & 31;
  const int slice = blockIdx.y;
  const int64_t r_begin = (int64_t)slice * slice_len;
  const int64_t r_end = min(nr, r_begin + slice_len);
  const int qw0 = w * QPW;  // wave's first local query
  const int64_t qb0 = (int64_t)blockIdx.x * QPB;

  Frag bq[QG][NS];
  float tau[QG];
  bool qvalid[QG];
  int qscale[QG];
  int cnt[QG];
#pragma unroll
  for (int g = 0; g < QG; ++g) {
    const int64_t q = qb0 + qw0 + g * 32 + n;
    qvalid[g] = q < nq;
    tau[g] = qvalid[g] ? kOpen : -INFINITY;  // a query past the batch end never passes
    cnt[g] = 0;
    qscale[g] = 127;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      bq[g][s] = qvalid[g] ? *reinterpret_cast<const Frag *>(
                                 Qp + q * GROWB + (h * HALF + s * Op::EPL) * Op::ESZ)
                           : Frag{};
    if (Op::kScaled && qvalid[g])
      qscale[g] = *reinterpret_cast<const int32_t *>(Qp + q * GROWB + K * Op::ESZ + 4);
  }

  // double buffering: stage t+1 streams into the other buffer (LDS DMA) while stage t
  // computes.  Chunk ci of a stage lands at byte 16*ci: row ci / CPR, column chunk
  // ci % CPR; the padding chunk of each row and rows past the slice re-read a valid
  // address (their values are never used).
  // Byte offset of this lane's chunk within a stage for each of its wave's instructions
  // (instruction j = w + 4 i fills LDS bytes [1024 j, 1024 j + 1024)); the padding chunk
  // of a row re-reads the row's first chunk (its value is never used).
  constexpr int JPW = (NINSTR + 3) / 4;  // instructions per wave per stage (at most)
  // rows the chunk layout touches per stage (can exceed RT by one when RT * CPR is not a
  // multiple of 64)
  constexpr int RROWS = (64 * NINSTR + CPR - 1) / CPR;
  uint32_t choff[JPW];
  int chrow[JPW];
#pragma unroll
  for (int i = 0; i < JPW; ++i) {
    const int ci = 64 * (w + 4 * i) + lane;
    const int rr = ci / CPR, c = ci % CPR;
    chrow[i] = rr;
    choff[i] = (uint32_t)(rr * GROWB + (c < CHUNKS ? c : 0) * 16);
  }
#define SH_SCREEN_FETCH(T0, B)                                                           \
  {                                                                                      \
    const unsigned char *stage = Rp + (T0) * GROWB;                                     \
    const bool whole = (T0) + RROWS <= r_end; /* uniform: no row past the slice */       \
    _Pragma("unroll") for (int i = 0; i < JPW; ++i) {                                    \
      const int j = w + 4 * i;                                                           \
      if (j < NINSTR) {                                                                  \
        const unsigned char *src;                                                        \
        if (whole) {                                                                     \
          src = stage + choff[i];                                                        \
        } else { /* rows past the slice re-read its last row */                          \
          const int64_t last = r_end - 1 - (T0);                                         \
          const int64_t over = chrow[i] > last ? (int64_t)chrow[i] - last : 0;           \
          src = stage + choff[i] - over * GROWB;                                         \
        }                                                                                \
        __builtin_amdgcn_global_load_lds(                                                \
            (const void *)src,                                                           \
            (__attribute__((address_space(3))) void *)(&lds_r[(B)][1024 * j]), 16, 0, 0); \
      }                                                                                  \
    }                                                                                    \
  }

  // Seeding: before the scan, every query of the workgroup screens the workgroup's QPB
  // seed references (seeds[qb0 .. qb0 + QPB), distinct rows of the reference set, e.g.
  // the block's own rows when queries are the references in locality order) and starts
  // from the pool-th smallest of those keys instead of an open threshold.  The seeds are
  // references of the slice, screened by the same MFMA arithmetic, so at least `pool`
  // references have a key at or below that threshold and the kept pool is the same; the
  // scan only skips candidates that could not have stayed.  (A 2^-16 relative margin
  // keeps this true even if a key were rounded differently in another tile position.)
  static_assert(QPB % 64 == 0 && QPB <= 256, "at most one seed slot per thread");
  static_assert(QPB % RT == 0 && (QPB / 32) * 16 <= HALFC, "seed stages must fit the buffer");
  if (seeds != nullptr && r_begin < r_end) {
    __shared__ uint32_t seed_row[QPB];
    __shared__ uint32_t seed_bad[QPB / 32];  // bit s of word i: slot 32 i + s has no seed
    if (threadIdx.x < QPB) {  // whole waves
      const int64_t q = qb0 + threadIdx.x;
      const int64_t r = q < nq ? seeds[q] : -1;
      const bool ok = r >= r_begin && r < r_end;
      seed_row[threadIdx.x] = (uint32_t)(ok ? r : r_begin);
      const uint64_t bad = __ballot(!ok);
      if (lane == 0) {
        seed_bad[2 * w] = (uint32_t)bad;
        seed_bad[2 * w + 1] = (uint32_t)(bad >> 32);
      }
    }
    __syncthreads();
    for (int s0 = 0; s0 < QPB; s0 += RT) {
      // gather the stage's seed rows: LDS row rr holds seed slot s0 + rr (rows the chunk
      // layout touches past the stage re-read a valid row; they are never used)
#pragma unroll
      for (int i = 0; i < JPW; ++i) {
        const int j = w + 4 * i;
        if (j < NINSTR) {
          const int slot = s0 + chrow[i];
          const uint32_t row = seed_row[slot < QPB ? slot : 0];
          const unsigned char *src =
              Rp + (int64_t)row * GROWB + (choff[i] - (uint32_t)(chrow[i] * GROWB));
          __builtin_amdgcn_global_load_lds(
              (const void *)src, (__attribute__((address_space(3))) void *)(&lds_r[0][1024 * j]),
              16, 0, 0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
#pragma unroll
      for (int sub = 0; sub < RT / 32; ++sub) {
        floatx16 acc[QG];
        if constexpr (Op::kScaled) {
          floatx16 nrm;
#pragma unroll
          for (int v = 0; v < 16; ++v)
            nrm[v] = *reinterpret_cast<const float *>(
                &lds_r[0][(sub * 32 + 8 * (v >> 2) + 4 * h + (v & 3)) * ROWB + K * Op::ESZ]);
#pragma unroll
          for (int g = 0; g < QG; ++g) acc[g] = nrm;
        } else {
#pragma unroll
          for (int g = 0; g < QG; ++g)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[g][v] = 0.f;
        }
        const unsigned char *arow = &lds_r[0][(sub * 32 + n) * ROWB + h * HALF * Op::ESZ];
        int ascale = 127;
        if constexpr (Op::kScaled)
          ascale = *reinterpret_cast<const int32_t *>(
              &lds_r[0][(sub * 32 + n) * ROWB + K * Op::ESZ + 4]);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          Frag a = *reinterpret_cast<const Frag *>(arow + s * Op::EPL * Op::ESZ);
#pragma unroll
          for (int g = 0; g < QG; ++g) {
            if constexpr (Op::kScaled)
              acc[g] = Op::mma_s(a, bq[g][s], acc[g], ascale, qscale[g]);
            else
              acc[g] = Op::mma(a, bq[g][s], acc[g]);
          }
        }
        const uint32_t badw = seed_bad[(s0 >> 5) + sub];
#pragma unroll
        for (int g = 0; g < QG; ++g) {
#pragma unroll
          for (int v = 0; v < 16; ++v)
            if ((badw >> (8 * (v >> 2) + 4 * h + (v & 3))) & 1u) acc[g][v] = INFINITY;
          const int64_t q = qb0 + qw0 + g * 32 + n;
          uint64_t *half = buf + ((qvalid[g] ? q : 0) * S + slice) * CAP + h * HALFC;
          append_quad<0, kLate>(acc[g], tau[g], 0u, half, cnt[g]);
          append_quad<4, kLate>(acc[g], tau[g], 0u, half, cnt[g]);
          append_quad<8, kLate>(acc[g], tau[g], 0u, half, cnt[g]);
          append_quad<12, kLate>(acc[g], tau[g], 0u, half, cnt[g]);
        }
      }
      __syncthreads();  // the next stage overwrites the buffer
    }
    // each query's threshold: the pool-th smallest seed key (open below `pool` seeds);
    // the seed keys themselves are dropped, the scan meets those references again
#pragma unroll
    for (int g = 0; g < QG; ++g) {
      for (int src = 0; src < 32; ++src) {
        const int64_t q = qb0 + qw0 + g * 32 + src;
        if (q >= nq) break;
        const int c0 = __builtin_amdgcn_readlane(cnt[g], src);
        const int c1 = __builtin_amdgcn_readlane(cnt[g], src + 32);
        float tnew;
        compact_pair<CAP>(buf + (q * S + slice) * CAP, c0, c1, pool, nullptr, nullptr, tnew);
        if (n == src)
          tau[g] = tnew < kOpen ? fminf(tnew + fabsf(tnew) * 0x1p-16f, kOpen) : kOpen;
      }
      cnt[g] = 0;
    }
  }

  int b = 0;
  if (r_begin < r_end) {
    SH_SCREEN_FETCH(r_begin, 0)
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int64_t t0 = r_begin; t0 < r_end; t0 += RT) {
    const int64_t tn = t0 + RT;
    if (tn < r_end) {
      SH_SCREEN_FETCH(tn, b ^ 1)
    }
#pragma unroll
    for (int sub = 0; sub < RT / 32; ++sub) {
      const int64_t rb = t0 + sub * 32;
      if (rb >= r_end) continue;
      floatx16 acc[QG];
      if constexpr (Op::kScaled) {
        // accumulator starts at ||r||^2 of each of the lane's 16 references
        floatx16 nrm;
#pragma unroll
        for (int v = 0; v < 16; ++v)
          nrm[v] = *reinterpret_cast<const float *>(
              &lds_r[b][(sub * 32 + 8 * (v >> 2) + 4 * h + (v & 3)) * ROWB + K * Op::ESZ]);
#pragma unroll
        for (int g = 0; g < QG; ++g) acc[g] = nrm;
      } else {
#pragma unroll
        for (int g = 0; g < QG; ++g)
#pragma unroll
          for (int v = 0; v < 16; ++v) acc[g][v] = 0.f;
      }
      const unsigned char *arow = &lds_r[b][(sub * 32 + n) * ROWB + h * HALF * Op::ESZ];
      // LLVM's small-GEMM interleave of the fragment LDS reads with the MFMAs: 175.1 ->
      // 172.3 ms for the 1M BF16 screen, alternated runs on one GPU
      __builtin_amdgcn_iglp_opt(0);
      int ascale = 127;
      if constexpr (Op::kScaled)
        ascale = *reinterpret_cast<const int32_t *>(&lds_r[b][(sub * 32 + n) * ROWB + K * Op::ESZ + 4]);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        Frag a = *reinterpret_cast<const Frag *>(arow + s * Op::EPL * Op::ESZ);
#pragma unroll
        for (int g = 0; g < QG; ++g) {
          if constexpr (Op::kScaled)
            acc[g] = Op::mma_s(a, bq[g][s], acc[g], ascale, qscale[g]);
          else
            acc[g] = Op::mma(a, bq[g][s], acc[g]);
        }
      }
      if (rb + 32 > r_end) {  // references past the slice end never pass
#pragma unroll
        for (int g = 0; g < QG; ++g)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int64_t ref = rb + 8 * (v >> 2) + 4 * h + (v & 3);
            if (ref >= r_end) acc[g][v] = INFINITY;
          }
      }
      // quad minima, then the tile minimum, per query group
      float m4[QG][4];
      bool hit[QG];
      bool any_hit = false;
#pragma unroll
      for (int g = 0; g < QG; ++g) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          m4[g][j] = fminf(fminf(acc[g][4 * j], acc[g][4 * j + 1]),
                           fminf(acc[g][4 * j + 2], acc[g][4 * j + 3]));
        const float m = fminf(fminf(m4[g][0], m4[g][1]), fminf(m4[g][2], m4[g][3]));
        hit[g] = m <= tau[g];
        any_hit |= hit[g];
      }
      if (wave_any(any_hit)) {
        // lane's first reference of the tile (quads add 8 * (v / 4) + v % 4)
        const uint32_t ref0 = (uint32_t)(rb + 4 * h);
#pragma unroll
        for (int g = 0; g < QG; ++g) {
          if (!wave_any(hit[g])) continue;
          const int64_t q = qb0 + qw0 + g * 32 + n;
          uint64_t *half = buf + ((qvalid[g] ? q : 0) * S + slice) * CAP + h * HALFC;
          if (wave_any(m4[g][0] <= tau[g]))
            append_quad<0, kLate>(acc[g], tau[g], ref0, half, cnt[g]);
          if (wave_any(m4[g][1] <= tau[g]))
            append_quad<4, kLate>(acc[g], tau[g], ref0, half, cnt[g]);
          if (wave_any(m4[g][2] <= tau[g]))
            append_quad<8, kLate>(acc[g], tau[g], ref0, half, cnt[g]);
          if (wave_any(m4[g][3] <= tau[g]))
            append_quad<12, kLate>(acc[g], tau[g], ref0, half, cnt[g]);
          // compact every query of the group one of whose halves cannot take a tile
          const uint64_t mk = __ballot(cnt[g] > HALFC - 16);
          uint32_t mq = (uint32_t)(mk | (mk >> 32));
          while (mq) {
            const int src = __ffs(mq) - 1;
            mq &= mq - 1;
            const int c0 = __builtin_amdgcn_readlane(cnt[g], src);
            const int c1 = __builtin_amdgcn_readlane(cnt[g], src + 32);
            const int64_t qs = qb0 + qw0 + g * 32 + src;
            float tnew;
            const int keep = compact_pair<CAP>(buf + (qs * S + slice) * CAP, c0, c1, pool,
                                               nullptr, nullptr, tnew);
            if (n == src) {
              cnt[g] = h == 0 ? (keep + 1) / 2 : keep - (keep + 1) / 2;
              tau[g] = tnew;
            }
          }
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);  // the next stage's DMA has landed
    __syncthreads();
    b ^= 1;
  }
#undef SH_SCREEN_FETCH
  // final: sort each query's buffer and emit its best `pool`
#pragma unroll
  for (int g = 0; g < QG; ++g) {
    for (int src = 0; src < 32; ++src) {
      const int64_t q = qb0 + qw0 + g * 32 + src;
      if (q >= nq) break;
      const int c0 = __builtin_amdgcn_readlane(cnt[g], src);
      const int c1 = __builtin_amdgcn_readlane(cnt[g], src + 32);
      float tnew;
      compact_pair<CAP>(buf + (q * S + slice) * CAP, c0, c1, pool, out + (q * S + slice) * pool,
                        out_cnt + q * S + slice, tnew);
    }
  }
}

// Exact FP64 squared distance, terms added in feature order, no contraction: the
// same evaluation as the CPU oracle, so results agree bit for bit.
__device__ __forceinline__ double sqdist_exact(const float *__restrict__ q_lds,
                                               const float *__restrict__ r, int D) {
#pragma clang fp contract(off)
  double acc = 0.0;
  int t = 0;
  if ((reinterpret_cast<uintptr_t>(r) & 7) == 0) {
    // 8-byte rows: half the load instructions (each lane reads a different row), the
    // same terms in the same order
    const float2 *r2 = reinterpret_cast<const float2 *>(r);
    for (; t + 1 < D; t += 2) {
      const float2 rv = r2[t >> 1];
      double df = (double)q_lds[t] - (double)rv.x;
      double sq = df * df;
      acc = acc + sq;
      df = (double)q_lds[t + 1] - (double)rv.y;
      sq = df * df;
      acc = acc + sq;
    }
  }
  for (; t < D; ++t) {
    double df = (double)q_lds[t] - (double)r[t];
    double sq = df * df;
    acc = acc + sq;
  }
  return acc;
}

// One wave per query: merge the slices' pools to the global best `pool` by screened
// key, re-rank them exactly, take the best k (+1 with self repair).
template <int PE>  // pool elements per lane (pool <= 64 * PE)
__global__ void __launch_bounds__(256) k_rerank(
    const float *__restrict__ Q, int64_t nq, int64_t ldq, const float *__restrict__ R,
    int64_t nr, int64_t ldr, int32_t D, const uint64_t *__restrict__ cand,
    const int32_t *__restrict__ cand_cnt, int pool, int S, int k, int64_t self_offset,
    const int64_t *__restrict__ self_idx,
    double err_coef, float r_norm_max, double margin, int64_t *__restrict__ idx_out,
    float *__restrict__ dist_out, int32_t *__restrict__ flags) {
  __shared__ float qv[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t q = (int64_t)blockIdx.x * 4 + w;
  if (q >= nq) return;
  float *ql = qv[w];
  for (int t = lane; t < D; t += 64) ql[t] = Q[q * ldq + t];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();

  // 1. merged pool: best `pool` keys over the S slices
  uint64_t v[PE];
  int total = 0;
  {
    const int c0 = cand_cnt[q * S];
    const uint64_t *p0 = cand + (q * S) * pool;
#pragma unroll
    for (int i = 0; i < PE; ++i) {
      const int e = lane * PE + i;
      v[i] = e < c0 ? p0[e] : ~0ull;
    }
    total = c0;
  }
  for (int s = 1; s < S; ++s) {
    const int cs = cand_cnt[q * S + s];
    const uint64_t *ps = cand + (q * S + s) * pool;
    uint64_t m[2 * PE];
#pragma unroll
    for (int i = 0; i < PE; ++i) m[i] = v[i];
    // second half of the 2*PE-per-lane array: the new slice
#pragma unroll
    for (int i = 0; i < PE; ++i) {
      const int e = lane * PE + i;
      m[PE + i] = e < cs ? ps[e] : ~0ull;
    }
    wave_bitonic_sort<2 * PE>(m);
    // the best `pool` now sit at elements 0..pool-1 = lanes 0..pool/(2PE)-1; move
    // them back to the PE-per-lane layout: element e lives at lane e/(2PE), slot e%(2PE)
#pragma unroll
    for (int i = 0; i < PE; ++i) {
      const int e = lane * PE + i;
      const int src = e / (2 * PE), slot = e % (2 * PE);
      uint64_t mine = 0;
      // every lane offers its slot `slot`... gather with a shuffle per slot value
#pragma unroll
      for (int sl = 0; sl < 2 * PE; ++sl) {
        uint32_t lo = __shfl((uint32_t)m[sl], src, kWave);
        uint32_t hi = __shfl((uint32_t)(m[sl] >> 32), src, kWave);
        if (sl == slot) mine = ((uint64_t)hi << 32) | lo;
      }
      v[i] = mine;
    }
    total += cs;
  }
  const int npool = total < pool ? total : pool;
  // screened key of the pool's last element: the bound on everything left out
  const uint64_t lastkey = wave_elem_u64<PE>(v, npool > 0 ? npool - 1 : 0);

  // 2. exact distances of the pool
  double qn = 0.0;  // ||q||^2 in FP64 (the certificate and the margin test)
  if (err_coef >= 0.0 || margin > 0.0)
    for (int t = 0; t < D; ++t) qn += (double)ql[t] * (double)ql[t];
  KeyIdx c[PE];
  double emax = 0.0;  // largest screening error |key + ||q||^2 - exact| over the pool
#pragma unroll
  for (int i = 0; i < PE; ++i) {
    const int e = lane * PE + i;
    if (e < npool) {
      const uint32_t j = (uint32_t)(v[i] & 0xffffffffu);
      c[i].d = sqdist_exact(ql, R + (int64_t)j * ldr, D);
      c[i].i = j;
      if (margin > 0.0)
        emax = fmax(emax, fabs((double)ord_float((uint32_t)(v[i] >> 32)) + qn - c[i].d));
    } else {
      c[i].d = INFINITY;
      c[i].i = 0xffffffffu;
    }
  }
  if (margin > 0.0) emax = wave_max_f64(emax);
  wave_bitonic_sort<PE>(c);

  const bool self = self_idx != nullptr || self_offset >= 0;
  const int kk = self ? k + 1 : k;
  // broadcast the first kk (<= 64*PE) sorted elements through LDS-free shuffles
  // element e is at lane e/PE, slot e%PE
  int32_t flag = 0;
  // certificate: every reference outside the pool has exact distance
  // >= lastkey + ||q||^2 - err; the k-th chosen one must be strictly below that.
  // References were left out whenever the index holds more rows than the merged pool
  // (each slice stores at most `pool` keys, so `total` cannot tell with one slice).
  if (err_coef >= 0.0) {
    bool certified = true;
    if (nr > (int64_t)npool && kk > 0) {
      const double rq = sqrt(qn);
      const double err =
          err_coef * ((double)r_norm_max * r_norm_max + 2.0 * rq * (double)r_norm_max + qn);
      const double bound = (double)ord_float((uint32_t)(lastkey >> 32)) + qn - err;
      // kk-th exact distance (element kk-1)
      const int e = (kk <= npool ? kk : npool) - 1;
      double dk = 0.0;
      {
        double mine = 0.0;
#pragma unroll
        for (int i = 0; i < PE; ++i)
          if (lane * PE + i == e) mine = c[i].d;
        dk = __shfl(mine, e / PE, kWave);
      }
      certified = dk < bound;
    }
    if (!certified) flag |= SH_FLAG_UNCERTIFIED;
  }
  // margin test of a low-precision screen: a reference left out of the pool has a
  // screened key >= lastkey, so a true neighbour was lost only if its screening error
  // reached gap = lastkey + ||q||^2 - d_kk.  The pool shows the errors the screen made
  // near this query; rows whose gap is under margin times the largest of them are
  // flagged for a wider second screen.
  if (margin > 0.0 && nr > (int64_t)npool && kk > 0) {
    const int e = (kk <= npool ? kk : npool) - 1;
    double mine = 0.0;
#pragma unroll
    for (int i = 0; i < PE; ++i)
      if (lane * PE + i == e) mine = c[i].d;
    const double dk = __shfl(mine, e / PE, kWave);
    const double gap = (double)ord_float((uint32_t)(lastkey >> 32)) + qn - dk;
    if (gap <= margin * emax) flag |= SH_FLAG_THIN;
  }
  // gather the top kk (index, distance) into lane-indexed registers: lane t <- elem t
  double dd[2];
  uint32_t ii[2];
#pragma unroll
  for (int rep = 0; rep < 2; ++rep) {
    const int e = lane + 64 * rep;
    const int src = (e < 64 * PE ? e : 0) / PE, slot = e % PE;
    double dm = INFINITY;
    uint32_t im = 0xffffffffu;
#pragma unroll
    for (int sl = 0; sl < PE; ++sl) {
      double t_d = __shfl(c[sl].d, src, kWave);
      uint32_t t_i = (uint32_t)__shfl((int)c[sl].i, src, kWave);
      if (sl == slot) {
        dm = t_d;
        im = t_i;
      }
    }
    if (e >= npool) {
      dm = INFINITY;
      im = 0xffffffffu;
    }
    dd[rep] = dm;
    ii[rep] = im;
  }
  // self repair (ann.py:31-52) on the kk = k+1 list
  int drop = -1;
  if (self) {
    const int64_t me = self_idx ? self_idx[q] : self_offset + q;
    int pos = -1;
    for (int e = 0; e < kk; ++e) {
      const uint32_t ie = (uint32_t)__shfl((int)ii[e / 64], e % 64, kWave);
      if ((int64_t)ie == me) {
        pos = e;
        break;
      }
    }
    if (pos < 0) {
      drop = kk - 1;
      flag |= SH_FLAG_SELF_MISSING;
    } else {
      drop = pos;
    }
    if (pos != 0) flag |= SH_FLAG_SELF_NOT_FIRST;
  }
  // output column col takes element col (col < drop) or col + 1 (col >= drop); the
  // element may sit in another lane: shuffle
#pragma unroll
  for (int rep = 0; rep < 2; ++rep) {
    const int col = lane + 64 * rep;
    const int e = (drop >= 0 && col >= drop) ? col + 1 : col;
    const int esrc = e < 128 ? e : 127;
    double d0 = __shfl(dd[0], esrc % 64, kWave), d1 = __shfl(dd[1], esrc % 64, kWave);
    uint32_t i0 = (uint32_t)__shfl((int)ii[0], esrc % 64, kWave);
    uint32_t i1 = (uint32_t)__shfl((int)ii[1], esrc % 64, kWave);
    const double de = esrc < 64 ? d0 : d1;
    const uint32_t ie = esrc < 64 ? i0 : i1;
    if (col < k) {
      idx_out[q * k + col] = ie == 0xffffffffu ? -1 : (int64_t)ie;
      dist_out[q * k + col] = (float)de;
    }
  }
  if (flags && lane == 0) flags[q] = flag;
}

// FP64 brute force over all references for selected query rows (rows the
// certificate rejected, or every row when used as the FP64 reference).  One wave per
// query; each lane keeps a sorted top-KK of the references it visits
// (lane, lane + 64, ...), then the wave sorts the 64 * KK survivors.
template <int KK>
__global__ void __launch_bounds__(256) k_exact(
    const float *__restrict__ Q, int64_t ldq, const int64_t *__restrict__ q_rows, int64_t n_sel,
    const float *__restrict__ R, int64_t nr, int64_t ldr, int32_t D, int k, int64_t self_offset,
    const int64_t *__restrict__ self_idx,
    int64_t *__restrict__ idx_out, float *__restrict__ dist_out, double *__restrict__ dist64_out,
    int32_t *__restrict__ flags) {
  __shared__ float qv[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t sel = (int64_t)blockIdx.x * 4 + w;
  if (sel >= n_sel) return;
  const int64_t q = q_rows ? q_rows[sel] : sel;
  float *ql = qv[w];
  for (int t = lane; t < D; t += 64) ql[t] = Q[q * ldq + t];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  KeyIdx best[KK];
#pragma unroll
  for (int i = 0; i < KK; ++i) {
    best[i].d = INFINITY;
    best[i].i = 0xffffffffu;
  }
  for (int64_t j = lane; j < nr; j += 64) {
    KeyIdx x;
    x.d = sqdist_exact(ql, R + j * ldr, D);
    x.i = (uint32_t)j;
    if (less_ki(x, best[KK - 1])) {
#pragma unroll
      for (int p = KK - 1; p >= 0; --p) {
        const bool here = less_ki(x, best[p]);
        const bool prev = p > 0 && less_ki(x, best[p - 1]);
        if (here) best[p] = prev ? best[p - 1] : x;
      }
    }
  }
  wave_bitonic_sort<KK>(best);
  const bool self = self_idx != nullptr || self_offset >= 0;
  const int kk = self ? k + 1 : k;
  int drop = -1;
  int32_t flag = 0;
  if (self) {
    const int64_t me = self_idx ? self_idx[q] : self_offset + q;
    int pos = -1;
    for (int e = 0; e < kk; ++e) {
      uint32_t mine = 0;
#pragma unroll
      for (int i = 0; i < KK; ++i)
        if (lane * KK + i == e) mine = best[i].i;
      const uint32_t ie = (uint32_t)__shfl((int)mine, e / KK, kWave);
      if ((int64_t)ie == me) {
        pos = e;
        break;
      }
    }
    drop = pos < 0 ? kk - 1 : pos;
    if (pos < 0) flag |= SH_FLAG_SELF_MISSING;
    if (pos != 0) flag |= SH_FLAG_SELF_NOT_FIRST;
  }
  for (int col = 0; col < k; ++col) {
    const int e = (drop >= 0 && col >= drop) ? col + 1 : col;
    double md = 0.0;
    uint32_t mi = 0;
#pragma unroll
    for (int i = 0; i < KK; ++i)
      if (lane * KK + i == e) {
        md = best[i].d;
        mi = best[i].i;
      }
    const double de = __shfl(md, e / KK, kWave);
    const uint32_t ie = (uint32_t)__shfl((int)mi, e / KK, kWave);
    if (lane == 0) {
      idx_out[q * k + col] = ie == 0xffffffffu ? -1 : (int64_t)ie;
      dist_out[q * k + col] = (float)de;
      if (dist64_out) dist64_out[q * k + col] = de;
    }
  }
  // a row recomputed here after the re-rank's certificate failed keeps that flag
  if (flags && lane == 0) flags[q] = (flags[q] & SH_FLAG_UNCERTIFIED) | flag;
}


// fix_knn_query (ann.py:31-52) on a (k+1)-column table, one thread per row: drop the
// query's own column (column 0 in the common case), else the last column.
__global__ void k_fix_knn(const int64_t *__restrict__ idx_in, const float *__restrict__ dist_in,
                          const int64_t *__restrict__ ref_idx, int64_t n, int k1,
                          int64_t *__restrict__ idx_out, float *__restrict__ dist_out,
                          int32_t *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t me = ref_idx[i];
  int pos = -1;
  for (int e = 0; e < k1; ++e)
    if (idx_in[i * k1 + e] == me) {
      pos = e;
      break;
    }
  int32_t flag = 0;
  if (pos != 0) flag |= SH_FLAG_SELF_NOT_FIRST;
  if (pos < 0) flag |= SH_FLAG_SELF_MISSING;
  const int drop = pos < 0 ? k1 - 1 : pos;
  for (int c = 0; c < k1 - 1; ++c) {
    const int e = c < drop ? c : c + 1;
    idx_out[i * (k1 - 1) + c] = idx_in[i * k1 + e];
    dist_out[i * (k1 - 1) + c] = dist_in[i * k1 + e];
  }
  if (flags) flags[i] = flag;
}

}  // namespace sh

using namespace sh;

extern "C" int sh_fix_knn_query(const int64_t *idx_in, const float *dist_in,
                                const int64_t *ref_idx, int64_t n, int32_t k1, int64_t *idx_out,
                                float *dist_out, int32_t *flags_out, sh_stream_t stream) {
  SH_ARG(n >= 0 && k1 >= 2, "sh_fix_knn_query: need n >= 0 and at least 2 columns");
  SH_ARG(idx_in && dist_in && ref_idx && idx_out && dist_out, "sh_fix_knn_query: NULL pointer");
  if (n == 0) return SH_OK;
  hipLaunchKernelGGL(k_fix_knn, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, idx_in,
                     dist_in, ref_idx, n, k1, idx_out, dist_out, flags_out);
  SH_LAUNCH_CHECK("k_fix_knn");
  return SH_OK;
}

extern "C" int32_t sh_knn_pack_k(int dtype, int32_t D) { return pack_k(dtype, D); }
extern "C" int64_t sh_knn_pack_row_bytes(int dtype, int32_t D) { return pack_row_bytes(dtype, D); }

extern "C" int sh_knn_pack(const float *E, int64_t n, int32_t D, int64_t lde, int dtype, int role,
                           void *out, float *aux_out, sh_stream_t stream) {
  SH_ARG(n >= 0 && D > 0 && D <= 128 && lde >= D, "sh_knn_pack: bad sizes D=%d lde=%lld", D,
         (long long)lde);
  SH_ARG(role == 0 || role == 1, "sh_knn_pack: role must be 0 (reference) or 1 (query)");
  SH_ARG(E && out, "sh_knn_pack: NULL pointer");
  const int32_t K = pack_k(dtype, D);
  if (K < 0) {
    set_error("sh_knn_pack: unsupported dtype %d", dtype);
    return SH_EUNSUPPORTED;
  }
  if (n == 0) return SH_OK;
  dim3 grid(cdiv(n, 256)), block(256);
  if (dtype == SH_F32)
    hipLaunchKernelGGL(k_pack_f32, grid, block, 0, (hipStream_t)stream, E, n, D, lde, role, K,
                       (float *)out, aux_out);
  else if (dtype == SH_BF16) {
    const size_t per_row = D * sizeof(float) + K * sizeof(uint16_t);
    int rb = kPackRows;
    while (rb > 1 && (size_t)rb * per_row > 65536) rb /= 2;
    SH_ARG((size_t)rb * per_row <= 65536, "sh_knn_pack: D=%d too large for the BF16 pack", D);
    hipLaunchKernelGGL(k_pack_bf16, dim3((uint32_t)cdiv(n, rb)), dim3(rb), (size_t)rb * per_row,
                       (hipStream_t)stream, E, n, D, lde, role, K, (uint16_t *)out, aux_out);
  }
  else
    hipLaunchKernelGGL(k_pack_fp8, grid, block, 0, (hipStream_t)stream, E, n, D, lde, role, K,
                       (uint8_t *)out, aux_out);
  SH_LAUNCH_CHECK("k_pack");
  return SH_OK;
}

static int screen_cap(int pool) { return pool <= 64 ? 256 : pool <= 128 ? 512 : 1024; }

extern "C" size_t sh_knn_screen_workspace(int64_t nq, int32_t pool, int32_t n_slices) {
  return (size_t)nq * n_slices * screen_cap(pool) * sizeof(uint64_t);
}

template <class Op, int K, int CAP>
static void launch_screen(hipStream_t st, const void *qpack, int64_t nq,
                          const void *rpack, int64_t nr, int64_t slice_len, int S, int pool,
                          const int64_t *seeds, uint64_t *buf, uint64_t *out, int32_t *cnt) {
  dim3 grid((uint32_t)cdiv(nq, ScreenShape<Op, K>::QPB), S);
  hipLaunchKernelGGL((k_screen<Op, K, CAP>), grid, dim3(256), 0, st,
                     (const typename Op::Frag *)qpack, nq, (const unsigned char *)rpack, nr,
                     slice_len, S, pool, seeds, buf, out, cnt);
}

extern "C" int sh_knn_screen(const void *qpack, int64_t nq, const void *rpack, int64_t nr,
                             int32_t D, int dtype, const float *q_aux, const float *r_aux,
                             int32_t pool, int32_t n_slices, const int64_t *seed_refs,
                             uint64_t *cand_keys, int32_t *cand_cnt, void *workspace,
                             size_t workspace_bytes, sh_stream_t stream) {
  (void)q_aux;
  (void)r_aux;
  SH_ARG(nq >= 0 && nr > 0 && D > 0, "sh_knn_screen: bad sizes nq=%lld nr=%lld D=%d",
         (long long)nq, (long long)nr, D);
  SH_ARG(nr < 0xffffffffll, "sh_knn_screen: more than 2^32-1 references");
  SH_ARG(pool >= 1 && pool <= 256, "sh_knn_screen: pool=%d outside [1, 256]", pool);
  SH_ARG(pool <= 128 || dtype == SH_FP8, "sh_knn_screen: pools above 128 are built for the FP8 "
         "screen only (pool=%d)", pool);
  SH_ARG(n_slices >= 1 && n_slices <= 4096, "sh_knn_screen: n_slices=%d", n_slices);
  SH_ARG(qpack && rpack && cand_keys && cand_cnt, "sh_knn_screen: NULL pointer");
  const size_t need = sh_knn_screen_workspace(nq, pool, n_slices);
  if (!workspace || workspace_bytes < need) {
    set_error("sh_knn_screen: workspace %zu < %zu bytes", workspace_bytes, need);
    return SH_EWORKSPACE;
  }
  if (nq == 0) return SH_OK;
  const int32_t K = pack_k(dtype, D);
  int64_t slice_len = align_up((size_t)cdiv(nr, n_slices), kSliceAlign);
  const int S = (int)cdiv(nr, slice_len);
  SH_ARG(S == n_slices, "sh_knn_screen: %d slices requested but %lld references give %d; use "
         "a slice count whose slices hold >= 128 references", n_slices, (long long)nr, S);
  SH_ARG(cdiv(nq, 128) <= 0x7fffffffll, "sh_knn_screen: too many queries");
  hipStream_t st = (hipStream_t)stream;
  uint64_t *buf = (uint64_t *)workspace;
  const bool big = pool > 64;
  const bool huge = pool > 128;
#define SH_SCREEN_CASE_1024(OP, KV)                                                          \
  if (K == KV && huge) {                                                                     \
    launch_screen<OP, KV, 1024>(st, qpack, nq, rpack, nr, slice_len, S, pool, seed_refs,     \
                                buf, cand_keys, cand_cnt);                                   \
    SH_LAUNCH_CHECK("k_screen");                                                             \
    return SH_OK;                                                                            \
  }
#define SH_SCREEN_CASE(OP, KV)                                                               \
  if (K == KV) {                                                                             \
    if (big)                                                                                 \
      launch_screen<OP, KV, 512>(st, qpack, nq, rpack, nr, slice_len, S, pool, seed_refs,    \
                                 buf, cand_keys, cand_cnt);                                  \
    else                                                                                     \
      launch_screen<OP, KV, 256>(st, qpack, nq, rpack, nr, slice_len, S, pool, seed_refs,    \
                                 buf, cand_keys, cand_cnt);                                  \
    SH_LAUNCH_CHECK("k_screen");                                                             \
    return SH_OK;                                                                            \
  }
  if (dtype == SH_BF16) {
    SH_SCREEN_CASE(OpBF16, 16)
    SH_SCREEN_CASE(OpBF16, 32)
    SH_SCREEN_CASE(OpBF16, 48)
    SH_SCREEN_CASE(OpBF16, 64)
    SH_SCREEN_CASE(OpBF16, 80)
    SH_SCREEN_CASE(OpBF16, 96)
    SH_SCREEN_CASE(OpBF16, 112)
    SH_SCREEN_CASE(OpBF16, 128)
  } else if (dtype == SH_FP8) {
    SH_SCREEN_CASE_1024(OpFP8, 64)
    SH_SCREEN_CASE_1024(OpFP8, 128)
    SH_SCREEN_CASE(OpFP8, 64)
    SH_SCREEN_CASE(OpFP8, 128)
  } else if (dtype == SH_F32) {
    SH_SCREEN_CASE(OpF32, 8)
    SH_SCREEN_CASE(OpF32, 16)
    SH_SCREEN_CASE(OpF32, 24)
    SH_SCREEN_CASE(OpF32, 32)
    SH_SCREEN_CASE(OpF32, 40)
    SH_SCREEN_CASE(OpF32, 48)
    SH_SCREEN_CASE(OpF32, 56)
    SH_SCREEN_CASE(OpF32, 64)
    SH_SCREEN_CASE(OpF32, 72)
    SH_SCREEN_CASE(OpF32, 80)
    SH_SCREEN_CASE(OpF32, 88)
    SH_SCREEN_CASE(OpF32, 96)
    SH_SCREEN_CASE(OpF32, 104)
    SH_SCREEN_CASE(OpF32, 112)
    SH_SCREEN_CASE(OpF32, 120)
    SH_SCREEN_CASE(OpF32, 128)
    SH_SCREEN_CASE(OpF32, 136)
  }
#undef SH_SCREEN_CASE
#undef SH_SCREEN_CASE_1024
  set_error("sh_knn_screen: dtype %d with packed K=%d not supported", dtype, K);
  return SH_EUNSUPPORTED;
}

extern "C" int sh_knn_rerank(const float *Q, int64_t nq, int64_t ldq, const float *R, int64_t nr,
                             int64_t ldr, int32_t D, const uint64_t *cand_keys,
                             const int32_t *cand_cnt, int32_t pool, int32_t n_slices, int32_t k,
                             int64_t self_offset, const int64_t *self_idx, double err_coef,
                             float r_norm_max, double margin,
                             int64_t *idx_out, float *dist_out, int32_t *flags_out,
                             sh_stream_t stream) {
  SH_ARG(nq >= 0 && nr > 0 && D > 0 && D <= 256, "sh_knn_rerank: bad sizes");
  SH_ARG(pool >= 1 && pool <= 256 && n_slices >= 1, "sh_knn_rerank: bad pool/slices");
  const int kk = k + ((self_offset >= 0 || self_idx) ? 1 : 0);
  SH_ARG(k >= 1 && kk <= pool && kk <= 127, "sh_knn_rerank: k=%d (+self) must be <= pool=%d", k,
         pool);
  SH_ARG(Q && R && cand_keys && cand_cnt && idx_out && dist_out, "sh_knn_rerank: NULL pointer");
  if (nq == 0) return SH_OK;
  dim3 grid(cdiv(nq, 4)), block(256);
  if (pool <= 64)
    hipLaunchKernelGGL(k_rerank<1>, grid, block, 0, (hipStream_t)stream, Q, nq, ldq, R, nr, ldr,
                       D, cand_keys, cand_cnt, pool, n_slices, k, self_offset, self_idx, err_coef,
                       r_norm_max, margin, idx_out, dist_out, flags_out);
  else if (pool <= 128)
    hipLaunchKernelGGL(k_rerank<2>, grid, block, 0, (hipStream_t)stream, Q, nq, ldq, R, nr, ldr,
                       D, cand_keys, cand_cnt, pool, n_slices, k, self_offset, self_idx, err_coef,
                       r_norm_max, margin, idx_out, dist_out, flags_out);
  else
    hipLaunchKernelGGL(k_rerank<4>, grid, block, 0, (hipStream_t)stream, Q, nq, ldq, R, nr, ldr,
                       D, cand_keys, cand_cnt, pool, n_slices, k, self_offset, self_idx, err_coef,
                       r_norm_max, margin, idx_out, dist_out, flags_out);
  SH_LAUNCH_CHECK("k_rerank");
  return SH_OK;
}

extern "C" int sh_knn_exact(const float *Q, int64_t nq, int64_t ldq, const int64_t *q_rows,
                            int64_t n_sel, const float *R, int64_t nr, int64_t ldr, int32_t D,
                            int32_t k, int64_t self_offset, const int64_t *self_idx,
                            int64_t *idx_out, float *dist_out,
                            double *dist64_out, int32_t *flags_out, sh_stream_t stream) {
  (void)nq;
  SH_ARG(n_sel >= 0 && nr > 0 && D > 0 && D <= 256, "sh_knn_exact: bad sizes");
  SH_ARG(nr < 0xffffffffll, "sh_knn_exact: more than 2^32-1 references");
  const int kk = k + ((self_offset >= 0 || self_idx) ? 1 : 0);
  SH_ARG(k >= 1 && kk <= 64, "sh_knn_exact: k=%d (+self) must be <= 64", k);
  SH_ARG(Q && R && idx_out && dist_out, "sh_knn_exact: NULL pointer");
  if (n_sel == 0) return SH_OK;
  dim3 grid(cdiv(n_sel, 4)), block(256);
  hipStream_t st = (hipStream_t)stream;
  if (kk <= 16)
    hipLaunchKernelGGL(k_exact<16>, grid, block, 0, st, Q, ldq, q_rows, n_sel, R, nr, ldr, D, k,
                       self_offset, self_idx, idx_out, dist_out, dist64_out, flags_out);
  else if (kk <= 32)
    hipLaunchKernelGGL(k_exact<32>, grid, block, 0, st, Q, ldq, q_rows, n_sel, R, nr, ldr, D, k,
                       self_offset, self_idx, idx_out, dist_out, dist64_out, flags_out);
  else
    hipLaunchKernelGGL(k_exact<64>, grid, block, 0, st, Q, ldq, q_rows, n_sel, R, nr, ldr, D, k,
                       self_offset, self_idx, idx_out, dist_out, dist64_out, flags_out);
  SH_LAUNCH_CHECK("k_exact");
  return SH_OK;
}
