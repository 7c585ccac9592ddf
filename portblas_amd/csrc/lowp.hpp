// Reduced-precision conversions shared by the split/quantise pre-pass and
// the tests' device-side checks.  All conversions round to nearest even.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pblowp {

// FP32 -> BF16 bits, round to nearest even; NaN stays a quiet NaN.
__host__ __device__ inline uint16_t f32_to_bf16(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  const uint32_t u = uint32_t(h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// FP32 -> OCP FP8 E4M3 ("e4m3fn": bias 7, no infinities, 0x7f/0xff = NaN,
// largest finite 448).  Round to nearest even, saturate finite overflow to
// +-448.  Scaling by powers of two is exact, so rintf does the rounding.
__host__ __device__ inline uint8_t f32_to_e4m3(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  const uint8_t sign = uint8_t((u >> 24) & 0x80u);
  const float ax = fabsf(x);
  if (!(ax == ax)) return uint8_t(sign | 0x7f);
  if (ax >= 448.0f) return uint8_t(sign | 0x7e);
  if (ax < 0.015625f) {  // below 2^-6: subnormal grid of 2^-9
    const float r = rintf(ax * 512.0f);  // 0..8 (8 = smallest normal)
    return uint8_t(sign | uint8_t(r));
  }
  int e;
  frexpf(ax, &e);  // ax = f * 2^e, f in [0.5, 1)
  e -= 1;          // ax in [2^e, 2^(e+1))
  float r = rintf(ldexpf(ax, 3 - e));  // in [8, 16]
  if (r >= 16.0f) {
    r = 8.0f;
    e += 1;
  }
  const int code = ((e + 7) << 3) | (int(r) - 8);
  if (code >= 0x7f) return uint8_t(sign | 0x7e);
  return uint8_t(sign | code);
}

__host__ __device__ inline float e4m3_to_f32(uint8_t v) {
  const int s = v >> 7, e = (v >> 3) & 0xf, m = v & 7;
  float r;
  if (e == 15 && m == 7) r = __builtin_nanf("");
  else if (e == 0) r = ldexpf(float(m), -9);
  else r = ldexpf(float(8 + m), e - 10);
  return s ? -r : r;
}

}  // namespace pblowp
