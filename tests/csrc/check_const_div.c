/* Exhaustive check that the gather's division by the viewport extent,
 *   y = RN(q * r), e = fma(-y, b, q), z = fma(e, r, y),  r = RN(1/b),
 * equals the IEEE quotient RN(q / b) for every float q in [2^-12, 4b).
 * Below 2^-12 both quotients are < 2^-12 / b, so the texel index
 * floor(s * extent_l) is 0 and the sign tests agree either way.
 * Usage: check_const_div b1 b2 ...; prints "ok" or the first mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float bits2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char **argv) {
  for (int a = 1; a < argc; a++) {
    float b = (float)atof(argv[a]);
    float r = 1.0f / b;
    uint32_t lo = f2bits(ldexpf(1.0f, -12)), hi = f2bits(4.0f * b);
    for (uint32_t u = lo; u < hi; u++) {
      float q = bits2f(u);
      float ref = q / b;
      float y = q * r;
      float e = fmaf(-y, b, q);
      float z = fmaf(e, r, y);
      if (f2bits(z) != f2bits(ref)) {
        printf("mismatch b=%g q=%a ref=%a fast=%a\n", b, q, ref, z);
        return 1;
      }
    }
  }
  printf("ok\n");
  return 0;
}
