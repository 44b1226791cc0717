// Host construction of the multiply-high divisors used by the kernels
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994): for 1 < d < 2^31 and x < 2^31,
//   x / d == umulhi(x, m) >> s,  l = ceil(log2 d), m = ceil(2^(31+l) / d), s = l - 1.
#pragma once

#include <cstdint>

#include "../kernels/pass_desc.h"

namespace sa {

inline sa_fastdiv_t make_fastdiv(uint32_t d) {
  sa_fastdiv_t f{};
  f.d = d == 0 ? 1 : d;
  if (f.d == 1) {
    f.m = 0;
    f.s = 0;
    return f;
  }
  uint32_t l = 0;
  while ((uint64_t(1) << l) < f.d) ++l;
  const uint64_t p = uint64_t(1) << (31 + l);
  f.m = (uint32_t)((p + f.d - 1) / f.d);
  f.s = l - 1;
  return f;
}

inline uint32_t fastdiv_host(uint32_t x, sa_fastdiv_t f) {
  if (f.d == 1) return x;
  return (uint32_t)(((uint64_t)x * f.m) >> 32) >> f.s;
}

}  // namespace sa
