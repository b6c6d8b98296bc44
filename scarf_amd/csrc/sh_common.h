// Shared helpers for the scarf_amd HIP library (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "scarf_hip.h"

namespace sh {

// Records a printf-style message for sh_last_error() (thread-local).
void set_error(const char *fmt, ...);

constexpr int kWave = 64;

// Orderable 32-bit image of a float: a < b  <=>  ord(a) < ord(b) (for non-NaN).
__device__ __forceinline__ uint32_t float_ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_float(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(u);
}
// 64-bit sort key: distance first, reference index second (lower index wins ties).
__device__ __forceinline__ uint64_t make_key(float d, uint32_t idx) {
  return ((uint64_t)float_ord(d) << 32) | (uint64_t)idx;
}
__device__ __forceinline__ uint64_t double_ord(double d) {
  uint64_t u = (uint64_t)__double_as_longlong(d);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
  return __shfl_xor(v, mask, kWave);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl_xor(lo, mask, kWave);
  hi = __shfl_xor(hi, mask, kWave);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}
// FP64 sum over the wave with DPP lane moves (no LDS round trips): two quad swaps, two
// rotations within each 16-lane row, then the four row sums read from lanes 0, 16, 32, 48
// and added in that fixed order.  Every lane gets the same value; the association differs
// from wave_sum_f64's butterfly.
template <int CTRL>
__device__ __forceinline__ double dpp_mov_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_f64_dpp(double v) {
  v += dpp_mov_f64<0xB1>(v);   // quad_perm [1, 0, 3, 2]: lane ^ 1
  v += dpp_mov_f64<0x4E>(v);   // quad_perm [2, 3, 0, 1]: lane ^ 2
  v += dpp_mov_f64<0x124>(v);  // row_ror:4
  v += dpp_mov_f64<0x128>(v);  // row_ror:8: each lane holds its row's sum
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
  return v;
}
__device__ __forceinline__ double wave_max_f64(double v) {
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, kWave));
  return v;
}
__device__ __forceinline__ float wave_max_f32(float v) {
  for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, kWave));
  return v;
}

// Normalised count x = log1p(sf * c / l) in FP64 with an IEEE division, as numpy evaluates
// it (assay.py:54-64).
__device__ __forceinline__ double norm_value(uint32_t c, double sf, double l, int log_t) {
  double v = (sf * (double)c) / l;
  return log_t ? log1p(v) : v;
}

// Feature of input column c (identity without a map); -1 when c is out of range or unmapped.
__device__ __forceinline__ int map_col(const int32_t *col_map, int32_t c, int32_t n_in) {
  if (c < 0 || c >= n_in) return -1;
  return col_map ? col_map[c] : c;
}

// Tile t of the row-major enumeration of an nb x nb upper block triangle (row bi holds
// nb - bi tiles) -> block coordinates (bi, bj), bi <= bj.
__device__ __forceinline__ void tile_coords(int t, int nb, int &bi, int &bj) {
  int r = 0, base = 0;
  while (t >= base + (nb - r)) {
    base += nb - r;
    ++r;
  }
  bi = r;
  bj = r + (t - base);
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace sh

#define SH_ARG(cond, ...)                 \
  do {                                    \
    if (!(cond)) {                        \
      ::sh::set_error(__VA_ARGS__);       \
      return SH_EINVAL;                   \
    }                                     \
  } while (0)

#define SH_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::sh::set_error("%s failed: %s", #call, hipGetErrorString(e_));        \
      return SH_EHIP;                                                        \
    }                                                                        \
  } while (0)

#define SH_LAUNCH_CHECK(name)                                                \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess) {                                                  \
      ::sh::set_error("%s launch failed: %s", name, hipGetErrorString(e_));   \
      return SH_EHIP;                                                        \
    }                                                                        \
  } while (0)
