// Elementary functions of the RK5 path evaluated from IEEE basic operations
// only, so every IEEE machine returns the same float (vendor libm expf/powf
// differ by an ulp between host and device):
//
// Step-size factors err^(-1/4) and err^(-1/5) of the RK5 controller
// (src/rk5.h:135-153, src/rk5.cu:68-70), evaluated in double with IEEE
// basic operations only (sqrt, division, multiply, add, exact scaling), then
// rounded to float.  The result is the correctly rounded float power except
// in ties closer than double rounding, and it is the same on every IEEE
// machine, so host and device integrate identical step sequences (a libm
// powf differs between vendors by an ulp).
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

// x^(-1/4) for x >= 0 (inf -> 0, NaN -> NaN, as powf)
NGPU_HD inline float pow_m025(float x) {
  if (!(x == x)) return x;
  if (x == INFINITY) return 0.0f;
  const double d = (double)x;
  return (float)(1.0 / sqrt(sqrt(d)));
}

// x^(-1/5) for x > 0: x = m*2^e (m in [0.5,1)), e = 5q + r, so
// x^(1/5) = 2^q * t^(1/5) with t = m*2^r in [0.5, 16); Newton's iteration
// y <- (4y + t/y^4)/5 from a linear start converges in 6 steps.
NGPU_HD inline float pow_m02(float x) {
  if (!(x == x)) return x;
  if (x == INFINITY) return 0.0f;
  if (x == 0.0f) return INFINITY;
  int e;
  const double m = frexp((double)x, &e);
  int q = e / 5, r = e % 5;
  if (r < 0) { r += 5; q -= 1; }
  const double t = ldexp(m, r);
  double y = 0.75 + 0.12 * t;
  for (int k = 0; k < 6; ++k) {
    const double y2 = y * y;
    const double y4 = y2 * y2;
    y = (4.0 * y + t / y4) / 5.0;
  }
  return (float)(1.0 / ldexp(y, q));
}

// e^x for the exponential term of the AdEx models: Cody-Waite reduction
// x = k ln2 + r (|r| <= ln2/2) and a degree-11 Taylor polynomial, all in
// double (relative error < 1e-14 before the final rounding to float).
NGPU_HD inline float exp_ieee(float x) {
  if (!(x == x)) return x;
  if (x > 89.0f) return (float)INFINITY;
  if (x < -104.0f) return 0.0f;
  const double xd = (double)x;
  const double k = rint(xd * 1.4426950408889634);
  const double r = xd - k * 0.6931471805599453;
  double p = 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return (float)ldexp(p, (int)k);
}

}  // namespace ngpu
