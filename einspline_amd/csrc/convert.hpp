// convert.hpp -- device-side conversion of coefficients to the table's
// storage type (round to nearest even), shared by the upload and the GPU
// table build so that both store identical bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace esk {

// double -> float rounded to odd: exact when representable, otherwise the
// truncated value with its last bit set.  Rounding that result to a format
// with at least two fewer mantissa bits (bf16, f16) equals rounding the
// double directly.
__device__ __forceinline__ float f64_to_f32_odd(double d) {
  float f = (float)d;
  if ((double)f == d || f != f) return f;
  unsigned u = __float_as_uint(f);
  if (fabs((double)f) > fabs(d)) u -= 1;        // step the magnitude back toward zero
  return __uint_as_float(u | 1u);
}

__device__ __forceinline__ uint16_t f32_to_bf16(float f, int &ovf) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);   // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  if ((u & 0x7f800000u) == 0x7f800000u) ovf = 1;                             // rounded to inf
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint16_t f32_to_f16(float f, int &ovf) {
  if (fabsf(f) >= 65520.0f) ovf = 1;            // beyond the largest half, rounds to inf
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}

}  // namespace esk
