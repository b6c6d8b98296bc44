// Wave-wide bitonic sort (wave64): 64*E elements held E per lane, element index
// e = lane * E + i.  Partners closer than E are swapped in registers, farther ones
// are exchanged with __shfl_xor.  Fully unrolled for a compile-time E.
#pragma once
#include "sh_common.h"

namespace sh {

struct KeyIdx {  // (FP64 distance, reference index), ordered lexicographically
  double d;
  uint32_t i;
};

__device__ __forceinline__ bool less_u64(uint64_t a, uint64_t b) { return a < b; }
__device__ __forceinline__ bool less_ki(const KeyIdx &a, const KeyIdx &b) {
  return a.d < b.d || (a.d == b.d && a.i < b.i);
}
__device__ __forceinline__ uint64_t shfl_elem(uint64_t v, int m) { return shfl_xor_u64(v, m); }
__device__ __forceinline__ KeyIdx shfl_elem(const KeyIdx &v, int m) {
  KeyIdx o;
  o.d = __shfl_xor(v.d, m, kWave);
  o.i = (uint32_t)__shfl_xor((int)v.i, m, kWave);
  return o;
}
__device__ __forceinline__ bool elem_less(uint64_t a, uint64_t b) { return a < b; }
__device__ __forceinline__ bool elem_less(const KeyIdx &a, const KeyIdx &b) { return less_ki(a, b); }

template <int E, typename T>
__device__ __forceinline__ void wave_bitonic_sort(T (&v)[E]) {
  const int lane = threadIdx.x & 63;
  constexpr int N = 64 * E;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= E) {
        const int lm = j / E;
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int e = lane * E + i;
          T o = shfl_elem(v[i], lm);
          const bool up = (e & k) == 0;
          const bool lower = (e & j) == 0;
          const bool o_less = elem_less(o, v[i]);
          const bool take_o = (lower == up) ? o_less : !o_less;
          if (take_o) v[i] = o;
        }
      } else {
#pragma unroll
        for (int i = 0; i < E; ++i) {
          if ((i & j) == 0) {
            const int e = lane * E + i;
            const bool up = (e & k) == 0;
            T a = v[i], b = v[i + j];
            const bool sw = up ? elem_less(b, a) : elem_less(a, b);
            if (sw) {
              v[i] = b;
              v[i + j] = a;
            }
          }
        }
      }
    }
  }
}

// Broadcast element e (0 <= e < 64*E) of a wave-distributed array to all lanes.
template <int E>
__device__ __forceinline__ uint64_t wave_elem_u64(const uint64_t (&v)[E], int e) {
  const int src = e / E, ii = e % E;
  uint64_t mine = 0;
#pragma unroll
  for (int i = 0; i < E; ++i)
    if (i == ii) mine = v[i];
  uint32_t lo = __shfl((uint32_t)mine, src, kWave);
  uint32_t hi = __shfl((uint32_t)(mine >> 32), src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

}  // namespace sh
