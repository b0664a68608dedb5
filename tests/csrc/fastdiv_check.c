/* Checks vxn_fastdiv_apply against integer division: every divisor in [1, 4096] and a
 * sample of larger ones, numerators at the multiples of d (and one either side), near
 * 2^31 and random; prints "mismatches N". */
#include <stdio.h>
#include <stdlib.h>
#include "vxn_fastdiv.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint32_t rnd31(void) {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return (uint32_t)(st >> 33);
}

static long check(uint32_t d, long nrand) {
  long bad = 0;
  vxn_fastdiv f = vxn_fastdiv_make(d);
  uint32_t edge[] = {0u, 1u, d - 1u, d, d + 1u, 2u * d - 1u, 0x7FFFFFFFu, 0x7FFFFFFFu - d, 0x40000000u};
  for (unsigned k = 0; k < sizeof(edge) / sizeof(edge[0]); ++k) {
    uint32_t n = edge[k] & 0x7FFFFFFFu;
    if (vxn_fastdiv_apply(f, n) != n / d) ++bad;
  }
  for (uint64_t q = 1; q * d < 0x80000000ull; q = q * 3 + 1) {
    uint32_t n = (uint32_t)(q * d);
    if (vxn_fastdiv_apply(f, n) != n / d) ++bad;
    if (vxn_fastdiv_apply(f, n - 1) != (n - 1) / d) ++bad;
    if (n + 1 < 0x80000000u && vxn_fastdiv_apply(f, n + 1) != (n + 1) / d) ++bad;
  }
  for (long i = 0; i < nrand; ++i) {
    uint32_t n = rnd31();
    if (vxn_fastdiv_apply(f, n) != n / d) {
      if (bad < 5) printf("mismatch d=%u n=%u\n", d, n);
      ++bad;
    }
  }
  return bad;
}

int main(int argc, char** argv) {
  long nrand = argc > 1 ? atol(argv[1]) : 20000;
  long bad = 0;
  for (uint32_t d = 1; d <= 4096; ++d) bad += check(d, nrand);
  for (int k = 0; k < 2000; ++k) bad += check(1u + rnd31(), nrand / 10);
  printf("mismatches %ld\n", bad);
  return bad != 0;
}
