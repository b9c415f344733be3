/* Exhaustive check of the kernels' division by the lattice constants: C^2,
 * 2C^4, 2C^2 with C = 1/sqrt(3) in FP32 (D2Q9), and C^2, 2C^4, 2C^2, 2C^6
 * with C = 1/sqrt(2) (D2Q13; C^4 = C^2 C^2, C^6 = C^2 C^4):
 *   q0 = x*r, rem = fma(-q0, y, x), q = fma(rem, r, q0),  r = RN(1/y)
 * against the IEEE quotient x / y, for x = +0 and every positive FP32 x whose
 * biased exponent lies in [lo, hi] (the method and the quotient are both odd in x).
 * Prints the number of mismatches per constant. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char **argv) {
  const uint32_t lo = argc > 1 ? (uint32_t)atoi(argv[1]) : 25, hi = argc > 2 ? (uint32_t)atoi(argv[2]) : 251;
  const float C = 1.0f / sqrtf(3.0f);
  const float C2 = C * C, C4x2 = 2.0f * (C2 * C2), C2x2 = 2.0f * C2;
  const float D = 1.0f / sqrtf(2.0f);
  const float D2 = D * D, D4 = D2 * D2, D6 = D2 * D4;
  const float ys[7] = {C2, C4x2, C2x2, D2, 2.0f * D4, 2.0f * D2, 2.0f * D6};
  for (int k = 0; k < 7; ++k) {
    const float y = ys[k], r = 1.0f / y;
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long u = (long long)lo << 23; u < ((long long)hi + 1) << 23; ++u) {
      const float x = bits((uint32_t)u);
      const float q0 = x * r;
      const float q = fmaf(fmaf(-q0, y, x), r, q0);
      if (ubits(q) != ubits(x / y)) ++bad;
    }
    {
      const float z = 0.0f, q0 = z * r;
      if (ubits(fmaf(fmaf(-q0, y, z), r, q0)) != ubits(z / y)) ++bad;
    }
    printf("%.9g %lld\n", y, bad);
  }
  return 0;
}
