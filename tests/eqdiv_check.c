/* The FP32 equilibrium of the step kernels divides by the lattice constants
 * with fast_div (q0 = x*r, rem = fma(-q0, y, x), q = fma(rem, r, q0),
 * r = RN(1/y)) when u.u < 1/2, and with the IEEE division otherwise. This
 * program evaluates the equilibrium both ways, in the kernels' operation
 * order, for many velocity pairs with u.u < 1/2 (random magnitudes from
 * subnormal to 1, exact zeros, ux = +-uy, u.u just under 1/2) and counts
 * outputs whose bits differ. Lattices: D2Q9 in the any-lattice form
 * (0-initialised ksi.u sums), D2Q9 in the tuned form (ux, uy, ux + uy,
 * uy - ux directly) and D2Q13. Prints "<form> <samples> <mismatches>". */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t ubits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float fbits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static uint64_t rs = 0x9E3779B97F4A7C15ull;
static uint64_t rnd(void) {  /* xorshift64* */
  rs ^= rs >> 12; rs ^= rs << 25; rs ^= rs >> 27;
  return rs * 2685821657736338717ull;
}
static float uni(void) { return (float)((rnd() >> 40) * (1.0 / 16777216.0)); }

/* a velocity component: sign, then zero, a subnormal, or 2^e * m */
static float comp(void) {
  const uint64_t k = rnd() % 16;
  const float s = (rnd() & 1) ? -1.0f : 1.0f;
  if (k == 0) return 0.0f;
  if (k == 1) return s * fbits((uint32_t)(1 + rnd() % 0x7fffff));  /* subnormal */
  const int e = -(int)(rnd() % 140);                              /* 2^0 .. 2^-139 */
  return s * ldexpf(1.0f + uni(), e - 1);
}

typedef struct { float W[13], C2, C4x2, C2x2, C6x2, r2, r4, r2x2, r6; } Lat;

static Lat lattice(int q13) {
  static const double w9[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                               1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  static const double w13[13] = {3.0 / 8, 1.0 / 12, 1.0 / 12, 1.0 / 12, 1.0 / 12, 1.0 / 16, 1.0 / 16,
                                 1.0 / 16, 1.0 / 16, 1.0 / 96, 1.0 / 96, 1.0 / 96, 1.0 / 96};
  Lat k;
  for (int q = 0; q < 13; ++q) k.W[q] = (float)(q13 ? w13[q] : (q < 9 ? w9[q] : 0.0));
  const float C = 1.0f / sqrtf(q13 ? 2.0f : 3.0f);
  k.C2 = C * C;
  const float c4 = k.C2 * k.C2, c6 = k.C2 * c4;
  k.C4x2 = 2.0f * c4; k.C2x2 = 2.0f * k.C2; k.C6x2 = 2.0f * c6;
  k.r2 = 1.0f / k.C2; k.r4 = 1.0f / k.C4x2; k.r2x2 = 1.0f / k.C2x2; k.r6 = 1.0f / k.C6x2;
  return k;
}

static float dv(int fast, float x, float y, float r) {
  if (!fast) return x / y;
  const float q0 = x * r;
  return fmaf(fmaf(-q0, y, x), r, q0);
}

static const int KX[13] = {0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0};
static const int KY[13] = {0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 2, 0, -2};
static const int POS[6] = {1, 2, 5, 6, 9, 10}, NEG[6] = {3, 4, 7, 8, 11, 12};

/* any-lattice kernels (fvdbm_lattice.hip: equilibrium_div) */
static void eq_lattice(const Lat *k, int q13, int fast, float rho, float ux, float uy, float *out) {
  float uu = 0.0f; uu += ux * ux; uu += uy * uy;
  const float D = dv(fast, uu, k->C2x2, k->r2x2);
  float c0 = 0.0f; c0 += 0.0f * ux; c0 += 0.0f * uy;
  float br = ((1.0f + c0) + c0 * c0) - D;
  if (q13) br = (br + c0 * (c0 * c0)) - (3.0f * c0) * uu;
  out[0] = (k->W[0] * rho) * br;
  for (int i = 0; i < (q13 ? 6 : 4); ++i) {
    const int qp = POS[i], qn = NEG[i];
    float cu = 0.0f; cu += (float)KX[qp] * ux; cu += (float)KY[qp] * uy;
    const float A = dv(fast, cu, k->C2, k->r2), B = dv(fast, cu * cu, k->C4x2, k->r4);
    float bp = ((1.0f + A) + B) - D, bn = ((1.0f + (-A)) + B) - D;
    if (q13) {
      const float E = dv(fast, cu * (cu * cu), k->C6x2, k->r6);
      const float G = dv(fast, (3.0f * cu) * uu, k->C4x2, k->r4);
      bp = (bp + E) - G;
      bn = (bn + (-E)) - (-G);
    }
    out[qp] = (k->W[qp] * rho) * bp;
    out[qn] = (k->W[qn] * rho) * bn;
  }
}

/* tuned D2Q9 kernels (fvdbm_step.hip: eq_terms + equilibrium) */
static void eq_tuned(const Lat *k, int fast, float rho, float ux, float uy, float *out) {
  float uu = 0.0f; uu += ux * ux; uu += uy * uy;
  const float c5 = ux + uy, c6 = (-ux) + uy;
  const float D = dv(fast, uu, k->C2x2, k->r2x2);
  const float A1 = dv(fast, ux, k->C2, k->r2), A2 = dv(fast, uy, k->C2, k->r2);
  const float A5 = dv(fast, c5, k->C2, k->r2), A6 = dv(fast, c6, k->C2, k->r2);
  const float B1 = dv(fast, ux * ux, k->C4x2, k->r4), B2 = dv(fast, uy * uy, k->C4x2, k->r4);
  const float B5 = dv(fast, c5 * c5, k->C4x2, k->r4), B6 = dv(fast, c6 * c6, k->C4x2, k->r4);
  const float w0 = k->W[0] * rho, w1 = k->W[1] * rho, w5 = k->W[5] * rho;
  out[0] = w0 * (((1.0f + 0.0f) + 0.0f) - D);
  out[1] = w1 * (((1.0f + A1) + B1) - D);
  out[2] = w1 * (((1.0f + A2) + B2) - D);
  out[3] = w1 * (((1.0f - A1) + B1) - D);
  out[4] = w1 * (((1.0f - A2) + B2) - D);
  out[5] = w5 * (((1.0f + A5) + B5) - D);
  out[6] = w5 * (((1.0f + A6) + B6) - D);
  out[7] = w5 * (((1.0f - A5) + B5) - D);
  out[8] = w5 * (((1.0f - A6) + B6) - D);
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? atol(argv[1]) : 4000000;
  const char *names[3] = {"lattice_d2q9", "tuned_d2q9", "lattice_d2q13"};
  for (int form = 0; form < 3; ++form) {
    const int q13 = form == 2, nq = q13 ? 13 : 9;
    const Lat k = lattice(q13);
    long tested = 0, bad = 0;
    for (long i = 0; i < n; ++i) {
      float ux = comp(), uy = comp();
      const uint64_t kind = rnd() % 8;
      if (kind == 0) uy = -ux;                       /* ksi.u of q5 = 0 */
      if (kind == 1) uy = ux;                        /* ksi.u of q6 = 0 */
      if (kind == 2) {                               /* u.u just under 1/2 */
        const float a = uni() * 6.2831853f;
        const float m = 0.70710677f * (1.0f - ldexpf(uni(), -20));
        ux = m * cosf(a); uy = m * sinf(a);
      }
      float uu = 0.0f; uu += ux * ux; uu += uy * uy;
      if (!(uu < 0.5f)) continue;
      const float rho = (rnd() & 1) ? 1.0f : ldexpf(1.0f + uni(), (int)(rnd() % 40) - 20);
      float a[13], b[13];
      if (form == 1) { eq_tuned(&k, 1, rho, ux, uy, a); eq_tuned(&k, 0, rho, ux, uy, b); }
      else { eq_lattice(&k, q13, 1, rho, ux, uy, a); eq_lattice(&k, q13, 0, rho, ux, uy, b); }
      ++tested;
      for (int q = 0; q < nq; ++q)
        if (ubits(a[q]) != ubits(b[q])) { ++bad; break; }
    }
    printf("%s %ld %ld\n", names[form], tested, bad);
  }
  return 0;
}
