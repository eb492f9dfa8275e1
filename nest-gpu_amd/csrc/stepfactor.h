// Elementary functions of the RK5 path evaluated from IEEE basic operations
// only, so every IEEE machine returns the same float (vendor libm expf/powf
// differ by an ulp between host and device): the step-size factors
// err^(-1/4) and err^(-1/5) of the RK5 controller (src/rk5.h:135-153,
// src/rk5.cu:68-70), the AdEx exponential, and division by a uniform
// parameter through its reciprocal.
#pragma once
#include <cmath>

#ifndef NGPU_HD
#ifdef __HIPCC__
#define NGPU_HD __host__ __device__
#else
#define NGPU_HD
#endif
#endif

namespace ngpu {

// x = m 2^e with m in [1, 2)
NGPU_HD inline double split_binade(float x, int* e) {
  int ex;
  const double m = frexp((double)x, &ex);
  *e = ex - 1;
  return 2.0 * m;
}

// x^(-1/4) for x >= 0 (inf -> 0, 0 -> inf, NaN -> NaN, as powf).  With
// x = m 2^(4q+r), x^(-1/4) = 2^(-q) 2^(-r/4) m^(-1/4); m^(-1/4) on [1, 2)
// from a quadratic start (relative error 2.2e-3) and three Newton steps
// z <- z (5 - m z^4) / 4 for the inverse root (error 2.5 e^2 per step),
// all in double with +,-,* only.  Equal to the correctly rounded float power
// for every positive float (exhaustive check against long-double powl).
NGPU_HD inline float pow_m025(float x) {
  if (!(x == x)) return x;
  if (x == INFINITY) return 0.0f;
  if (x == 0.0f) return INFINITY;
  int e;
  const double m = split_binade(x, &e);
  int q = e / 4, r = e % 4;
  if (r < 0) { r += 4; q -= 1; }
  // 2^(-r/4) by selects, not a table: a table read would be a memory load
  // (and wait) inside the RK5 loop
  const double c4 = r == 0 ? 1.0 : r == 1 ? 0.8408964152537145 : r == 2 ? 0.7071067811865476
                                                                          : 0.5946035575013605;
  double z = 1.2852 + m * (-0.35326 + m * 0.06590);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double z2 = z * z;
    const double z4 = z2 * z2;
    z = z * ((5.0 - m * z4) * 0.25);
  }
  return (float)ldexp(z * c4, -q);
}

// x^(-1/5) for x >= 0, the same construction: x = m 2^(5q+r), a quadratic
// start for m^(-1/5) (relative error 1.7e-3) and three Newton steps
// z <- z (6 - m z^5) / 5.  Correctly rounded for every positive float.
NGPU_HD inline float pow_m02(float x) {
  if (!(x == x)) return x;
  if (x == INFINITY) return 0.0f;
  if (x == 0.0f) return INFINITY;
  int e;
  const double m = split_binade(x, &e);
  int q = e / 5, r = e % 5;
  if (r < 0) { r += 5; q -= 1; }
  const double c5 = r == 0   ? 1.0
                    : r == 1 ? 0.8705505632961241
                    : r == 2 ? 0.757858283255199
                    : r == 3 ? 0.6597539553864471
                             : 0.5743491774985174;
  double z = 1.2284 + m * (-0.28173 + m * 0.05166);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double z2 = z * z;
    const double z4 = z2 * z2;
    const double z5 = z4 * z;
    z = z * ((6.0 - m * z5) * 0.2);
  }
  return (float)ldexp(z * c5, -q);
}

// RN(1/d) for a divisor usable by div_rn (normal, |d| in [2^-32, 2^32]),
// else 0.  1/d rounded to double and then to float is RN(1/d): the exact
// reciprocal of a float is never within 2^-49 (relative) of a float midpoint.
// err^-1/5 (five) or err^-1/4 (not five), lane by lane the operations of
// pow_m02 / pow_m025 with the constants selected: one evaluation serves a
// wave whose lanes took both the accept and the reject branch.
NGPU_HD inline float pow_step(float x, bool five) {
  if (!(x == x)) return x;
  if (x == INFINITY) return 0.0f;
  if (x == 0.0f) return INFINITY;
  int e;
  const double m = split_binade(x, &e);
  int q4 = e / 4, r4 = e % 4;
  if (r4 < 0) { r4 += 4; q4 -= 1; }
  int q5 = e / 5, r5 = e % 5;
  if (r5 < 0) { r5 += 5; q5 -= 1; }
  const int q = five ? q5 : q4, r = five ? r5 : r4;
  // both branches' constants stay literals (selected results, not selected
  // operands), so the evaluation holds no more registers than one power
  const double z5 = 1.2284 + m * (-0.28173 + m * 0.05166);
  const double z4i = 1.2852 + m * (-0.35326 + m * 0.06590);
  double z = five ? z5 : z4i;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double z2 = z * z;
    const double z4 = z2 * z2;
    const double mz = m * (five ? z4 * z : z4);
    z = z * (five ? (6.0 - mz) * 0.2 : (5.0 - mz) * 0.25);
  }
  const double c = five ? (r == 0   ? 1.0
                           : r == 1 ? 0.8705505632961241
                           : r == 2 ? 0.757858283255199
                           : r == 3 ? 0.6597539553864471
                                    : 0.5743491774985174)
                        : (r == 0   ? 1.0
                           : r == 1 ? 0.8408964152537145
                           : r == 2 ? 0.7071067811865476
                                    : 0.5946035575013605);
  return (float)ldexp(z * c, -q);
}

NGPU_HD inline float recip_rn(float d) {
  const float a = fabsf(d);
  if (!(a >= 0x1p-32f && a <= 0x1p32f)) return 0.0f;
  return (float)(1.0 / (double)d);
}

// a / d correctly rounded, given r = recip_rn(d) != 0: q = RN(a r) is within
// one ulp of a/d, the remainder a - d q is exact with an fma, and
// RN(q + rem r) is RN(a/d) (Markstein's theorem).  Exhaustive checks over all
// 2^32 dividends show the identity for |a| in [2^-64, 2^64) and failures only
// where the quotient or remainder reaches the subnormal range or a is
// infinite, so dividends outside that range (zeros, subnormals, inf, NaN)
// take the IEEE division: one integer range test on the magnitude bits.
NGPU_HD inline float div_rn(float a, float d, float r) {
  unsigned int u;
  __builtin_memcpy(&u, &a, sizeof(u));
  if ((u & 0x7fffffffu) - 0x1f800000u >= 0x40000000u) return a / d;
  const float q = a * r;
  const float rem = fmaf(-q, d, a);
  return fmaf(rem, r, q);
}

// e^x for the exponential term of the AdEx models, in float: Cody-Waite
// reduction x = k ln2 + r with ln2 split in two floats (the high part has 15
// significant bits, so k * ln2_hi is exact for the |k| <= 160 reached here),
// a degree-7 Taylor polynomial of e^r (|r| <= ln2/2) in Horner form with
// fma, and the scaling by 2^k (rounded once where the result is
// subnormal, overflowing to +inf above ~88.72).  Error < 1 ulp over the
// whole float range (0.93 ulp at most; 99.5% of results correctly rounded).
NGPU_HD inline float exp_fma(float x) {
  // branch-free: the polynomial runs on the clamped argument and the special
  // cases are selected afterwards (NaN -> x, above 100 -> +inf, below -110 -> 0)
  const float xc = fminf(fmaxf(x, -110.0f), 100.0f);
  const float k = rintf(xc * 1.44269504f);
  float r = fmaf(-k, 0.693145752f, xc);
  r = fmaf(-k, 1.42860677e-06f, r);
  float p = 1.0f / 5040.0f;
  p = fmaf(p, r, 1.0f / 720.0f);
  p = fmaf(p, r, 1.0f / 120.0f);
  p = fmaf(p, r, 1.0f / 24.0f);
  p = fmaf(p, r, 1.0f / 6.0f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  float e = ldexpf(p, (int)k);
  e = x > 100.0f ? (float)INFINITY : e;
  e = x < -110.0f ? 0.0f : e;
  return x == x ? e : x;
}

}  // namespace ngpu
