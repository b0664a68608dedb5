/* Division of a non-negative count below 2^31 by a fixed divisor d as a multiply-high, an
 * add and a shift (Granlund & Montgomery, "Division by invariant integers using
 * multiplication", PLDI 1994; Warren, "Hacker's Delight", 10-8): with l = ceil(log2 d) and
 * m = floor(2^32 (2^l - d) / d) + 1, floor(n / d) = (mulhi(m, n) + n) >> l.  The kernels
 * use it for the ray of a sample (sample / samples per ray), which the compiler would
 * otherwise expand into a full integer division.  Plain C, so the host test compiles the
 * same header with gcc. */
#ifndef VXN_FASTDIV_H
#define VXN_FASTDIV_H

#include <stdint.h>

#ifndef VXN_HD
#define VXN_HD
#endif

typedef struct {
  uint32_t mul;
  uint32_t shift;
} vxn_fastdiv;

/* d in [1, 2^31) */
static inline vxn_fastdiv vxn_fastdiv_make(uint32_t d) {
  uint32_t l = 0;
  while (((uint64_t)1 << l) < d) ++l;
  vxn_fastdiv f;
  f.mul = (uint32_t)(((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d) + 1);
  f.shift = l;
  return f;
}

/* floor(n / d) for n in [0, 2^31) */
VXN_HD static inline uint32_t vxn_fastdiv_apply(vxn_fastdiv f, uint32_t n) {
  return (uint32_t)(((((uint64_t)n * f.mul) >> 32) + n) >> f.shift);
}

#endif
