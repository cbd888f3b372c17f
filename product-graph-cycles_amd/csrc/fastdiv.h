// Unsigned 32-bit division by a run-time invariant divisor with one
// multiply-high, an add and two shifts: Granlund and Montgomery, "Division by
// invariant integers using multiplication" (PLDI 1994), Figure 4.1. The GPU
// has no integer divide instruction; the solver divides expanded node ids by
// the same two divisors (nodes per layer, nodes per expanded layer) all the
// time.
#pragma once
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
#define __host__
#define __device__
#endif

namespace pgc {

struct FastDiv {
  uint32_t m;
  int s1, s2;

  __host__ __device__ void init(uint32_t d) {
    int l = 0;  // ceil(log2 d)
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;
    m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    s1 = l < 1 ? l : 1;
    s2 = l - 1 > 0 ? l - 1 : 0;
  }
  __host__ __device__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t t = __umulhi(m, n);
#else
    const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
#endif
    return (t + ((n - t) >> s1)) >> s2;
  }
};

}  // namespace pgc
