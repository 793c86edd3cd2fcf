// Division of 31-bit unsigned integers by a runtime-invariant divisor via one
// multiply and one shift (Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, Theorem 4.2 with N = 31):
//   l = ceil(log2 d),  m = ceil(2^(31+l) / d) < 2^32,
//   floor(n / d) = (n * m) >> (31 + l)   for 0 <= n < 2^31.
#pragma once

#include <cstdint>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#endif

namespace lbmh {

struct FastDiv {
  uint32_t d;
  uint32_t m;
  uint32_t shift;  // 31 + l

  FastDiv() = default;
  explicit FastDiv(uint32_t divisor) : d(divisor)
  {
    uint32_t l = 0;
    while ((uint64_t(1) << l) < divisor) ++l;
    const uint64_t num = uint64_t(1) << (31 + l);
    m = uint32_t((num + divisor - 1) / divisor);
    shift = 31 + l;
  }

  __host__ __device__ inline uint32_t div(uint32_t n) const
  {
    return uint32_t((uint64_t(n) * m) >> shift);
  }
};

}  // namespace lbmh
