/* CPU check of the stencil division used by the grid kernels: a / b from y = RN(1 / b)
 * with two fma residual corrections (Markstein), against the correctly rounded float
 * division, on random quotients.  Prints the number of quotients compared and of
 * mismatches with a normal (non-subnormal) result; exits 1 on any mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float div_rn(float a, float b, float y) {
  const float q0 = a * y;
#ifdef NO_CORRECTION
  return q0; /* planted weakening: the test expects this form to fail */
#endif
  const float e0 = fmaf(-q0, b, a);
  const float q1 = fmaf(e0, y, q0);
  const float e1 = fmaf(-q1, b, a);
  const float q2 = copysignf(fmaf(e1, y, q1), q0);
  return isinf(q0) ? q0 : q2;
}

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(void) {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return (uint32_t)(st >> 11);
}
static float u01(void) { return (float)(rnd() & 0xFFFFFF) / 16777216.0f; }
static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static long n_cmp = 0, n_bad = 0;
static void check(float a, float b) {
  const float y = 1.0f / b;
  const float want = a / b, got = div_rn(a, b, y);
  ++n_cmp;
  if (bits(want) != bits(got) && !(want != want && got != got) && fabsf(want) >= 1.17549435e-38f) {
    if (n_bad < 5) printf("mismatch a=%a b=%a want=%a got=%a\n", a, b, want, got);
    ++n_bad;
  }
}

int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : (1L << 24);
  /* general positive normal divisors, dividends of every exponent in [-60, 60] */
  for (long i = 0; i < n; ++i) {
    const float b = from_bits(((uint32_t)(127 - 40 + rnd() % 80) << 23) | (rnd() & 0x7FFFFF));
    const float a = from_bits(((rnd() & 1u) << 31) | ((uint32_t)(127 - 60 + rnd() % 120) << 23) | (rnd() & 0x7FFFFF));
    check(a, b);
  }
  /* the stencil's divisors: eps = (hi - lo) / (res - 1) and 2 eps, dividends are
     differences of SDF values and products of weights and gradients */
  for (long i = 0; i < n; ++i) {
    const float lo = -1.0f - 3.0f * u01(), hi = 1.0f + 3.0f * u01();
    const int res = 2 + (int)(rnd() % 4095);
    const float eps = (hi - lo) / ((float)res - 1.0f);
    const float b = (rnd() & 1) ? 2.0f * eps : eps;
    const float f1 = (u01() - 0.5f) * 4.0f, f2 = f1 + (u01() - 0.5f) * eps * 8.0f;
    const float a = (rnd() & 1) ? f2 - f1 : u01() * (u01() - 0.5f) * 1e-3f;
    check(a, b);
  }
  /* special dividends */
  const float sp[] = {0.0f, -0.0f, INFINITY, -INFINITY, NAN, 1.0f, -1.0f};
  for (int i = 0; i < 7; ++i) check(sp[i], 0.0078125f * 3.0f);
  if (bits(div_rn(-0.0f, 3.0f, 1.0f / 3.0f)) != bits(-0.0f)) ++n_bad;
  if (div_rn(INFINITY, 3.0f, 1.0f / 3.0f) != INFINITY) ++n_bad;
  printf("compared %ld, mismatches %ld\n", n_cmp, n_bad);
  return n_bad != 0;
}
