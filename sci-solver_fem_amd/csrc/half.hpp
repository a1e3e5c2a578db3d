// FP16 storage of the level-0 smoother matrix (precision 2); device sources only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fem {

// A value is rounded to FP32 first, then to FP16; every producer and checker
// of FP16 values rounds this same way.
__host__ __device__ inline uint16_t half_bits(double v) {
  const _Float16 hv = (_Float16)(float)v;
  return __builtin_bit_cast(uint16_t, hv);
}
__host__ __device__ inline double half_value(double v) { return (double)(_Float16)(float)v; }

}  // namespace fem
