/* Check of the gather's quotients over a shared denominator b,
 *   r = RN(1/b), y = RN(x * r), e = fma(-y, b, x), z = fma(e, r, y),
 * against the IEEE quotient RN(x / b).  Scaling x or b by a power of two
 * scales y, e and z exactly (no intermediate under- or overflows for the
 * magnitudes the gather divides), and a sign flip of x or b flips them, so
 * the significands decide: every significand of b in [1, 2) is checked
 * against fixed numerators (1, the largest significand, b's neighbours) and
 * `per_b` pseudo-random significands of x in [1, 2), then x = +-0 (whose
 * quotient keeps its sign: the gather returns x itself).
 * Usage: check_common_div [per_b]; prints "ok N" or the first mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float bits2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float fast_div(float x, float b, float r) {
  float y = x * r;
  float e = fmaf(-y, b, x);
  float z = fmaf(e, r, y);
  return x == 0.0f ? x : z;
}

static int check(float x, float b, float r, unsigned long long *n) {
  (*n)++;
  float ref = x / b, z = fast_div(x, b, r);
  if (f2bits(z) != f2bits(ref)) {
    printf("mismatch b=%a x=%a ref=%a fast=%a\n", b, x, ref, z);
    return 1;
  }
  return 0;
}

int main(int argc, char **argv) {
  int per_b = argc > 1 ? atoi(argv[1]) : 32;
  uint64_t s = 0x9E3779B97F4A7C15ull;
  unsigned long long n = 0;
  const uint32_t one = f2bits(1.0f);
  for (uint32_t m = 0; m < (1u << 23); m++) {
    float b = bits2f(one | m);
    float r = 1.0f / b;
    float fixed[5] = {1.0f, bits2f(one | 0x7fffffu), b, bits2f(one | (m ? m - 1 : 0)),
                      bits2f(one | (m < 0x7fffffu ? m + 1 : m))};
    for (int i = 0; i < 5; i++)
      if (check(fixed[i], b, r, &n) || check(-fixed[i], b, r, &n) || check(fixed[i], -b, -r, &n)) return 1;
    for (int i = 0; i < per_b; i++) {
      s ^= s << 13;
      s ^= s >> 7;
      s ^= s << 17;
      float x = bits2f(one | (uint32_t)(s & 0x7fffffu));
      if (check(x, b, r, &n)) return 1;
    }
  }
  /* signed zeros and exponent spread on a few denominators */
  const float bs[4] = {1.0f, 3.0f, 0.7f, 1234.567f};
  for (int i = 0; i < 4; i++) {
    float b = bs[i], r = 1.0f / b;
    if (check(0.0f, b, r, &n) || check(-0.0f, b, r, &n)) return 1;
    for (int k = -60; k <= 60; k += 7)
      if (check(ldexpf(1.2345678f, k), b, r, &n) || check(ldexpf(-1.9999999f, k), b, r, &n)) return 1;
  }
  printf("ok %llu\n", n);
  return 0;
}
