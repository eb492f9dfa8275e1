// Host check of the device elementary functions (nest-gpu_amd/csrc/stepfactor.h,
// compiled for the CPU): div_rn against IEEE division, pow_m025/pow_m02
// against the long-double power rounded to float.  Prints one line per
// check and exits non-zero on any mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "stepfactor.h"

static float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

int main() {
  long long bad = 0;
  // divisors: model time constants and capacitances, eps, and significands
  // that stress the correction (all ones, 1 + ulp, pi)
  const float ds[] = {281.0f, 250.0f, 144.0f, 1.0f, 2.0f, 20.0f, 0.5f, 10.0f, 3.0f, 7.0f,
                      0.1f, 1.0e-6f, 30.0f, -3.0f, bits_f(0x3fffffffu), bits_f(0x3f800001u),
                      bits_f(0x40490fdbu), 1.0e-9f, 4.0e9f};
  const int exps[] = {-64, -40, -7, 0, 1, 9, 30, 63};
  long long n = 0;
  for (float d : ds) {
    const float r = ngpu::recip_rn(d);
    if (r == 0.0f) { std::printf("recip_rn(%a) unusable\n", d); ++bad; continue; }
#pragma omp parallel for reduction(+ : bad, n) schedule(static)
    for (int m = 0; m < (1 << 23); ++m) {
      for (int e : exps)
        for (int sg = 0; sg < 2; ++sg) {
          float a = std::ldexp(bits_f(0x3f800000u | (uint32_t)m), e);
          if (sg) a = -a;
          ++n;
          if (f_bits(ngpu::div_rn(a, d, r)) != f_bits(a / d)) ++bad;
        }
    }
    // special and out-of-range dividends take the IEEE division
    const float sp[] = {0.0f, -0.0f, INFINITY, -INFINITY, 1e-40f, -1e-45f, 1e-30f, 3e30f, 3e38f};
    for (float a : sp) {
      ++n;
      if (f_bits(ngpu::div_rn(a, d, r)) != f_bits(a / d)) { std::printf("special %a / %a\n", a, d); ++bad; }
    }
  }
  std::printf("div_rn: %lld dividends, %lld mismatches\n", n, bad);
  // inverse roots: every 257th positive float
  long long bad_r = 0, nr = 0;
#pragma omp parallel for reduction(+ : bad_r, nr) schedule(static)
  for (long long u = 1; u < 0x7f800000LL; u += 257) {
    const float x = bits_f((uint32_t)u);
    const float w4 = (float)std::pow((long double)x, -0.25L);
    const float w5 = (float)std::pow((long double)x, -0.2L);
    ++nr;
    if (f_bits(ngpu::pow_m025(x)) != f_bits(w4)) ++bad_r;
    if (f_bits(ngpu::pow_m02(x)) != f_bits(w5)) ++bad_r;
    // the one-evaluation form used by the persistent kernel
    if (f_bits(ngpu::pow_step(x, false)) != f_bits(w4)) ++bad_r;
    if (f_bits(ngpu::pow_step(x, true)) != f_bits(w5)) ++bad_r;
  }
  if (!(ngpu::pow_m025(0.0f) == INFINITY && ngpu::pow_m02(0.0f) == INFINITY &&
        ngpu::pow_m025(INFINITY) == 0.0f && ngpu::pow_m02(INFINITY) == 0.0f &&
        std::isnan(ngpu::pow_m02(NAN)) && ngpu::pow_step(0.0f, true) == INFINITY &&
        ngpu::pow_step(0.0f, false) == INFINITY && ngpu::pow_step(INFINITY, true) == 0.0f &&
        ngpu::pow_step(INFINITY, false) == 0.0f && std::isnan(ngpu::pow_step(NAN, true))))
    ++bad_r;
  std::printf("pow_m025/pow_m02: %lld inputs, %lld mismatches\n", nr, bad_r);
  return (bad || bad_r) ? 1 : 0;
}
