// Single-precision sine and cosine of one argument for the field formation.
//
// Argument reduction after Cody and Waite (W. J. Cody, W. Waite, "Software
// Manual for the Elementary Functions", 1980): x = k * pi/2 + r with
// k = rint(x * 2/pi) and |r| <= pi/4 (up to rounding), where pi/2 is split
// into a float C1 and the float C2 nearest to pi/2 - C1; with fused
// multiply-adds the two subtractions carry the product k * C1 exactly, so r
// is within half an ulp of pi/4 plus |k| * 2^-48 of the true remainder. sin
// and cos of r are their Taylor polynomials (to r^9 and r^10; the first
// omitted terms are below 2e-9 and 1.5e-10 on |r| <= pi/4), evaluated by
// Horner's rule in r^2, and k mod 4 picks the quadrant. The absolute error
// is about 1e-7 for |x| up to 1e4 (the reduction's k * 2^-48 term grows
// with |x| beyond that); there is no large-argument path.
#pragma once
#include <hip/hip_runtime.h>

namespace mtd {

__host__ __device__ __forceinline__ void sincos_cw(float x, float* s, float* c) {
  constexpr double kHalfPi = 1.57079632679489661923;
  constexpr float C1 = (float)kHalfPi;
  constexpr float C2 = (float)(kHalfPi - (double)C1);
  constexpr float kTwoOverPi = (float)(2.0 / 3.14159265358979323846);
  const float k = rintf(x * kTwoOverPi);
  const float r = fmaf(-k, C2, fmaf(-k, C1, x));
  const float z = r * r;
  // Taylor coefficients (-1)^n / (2n+1)! and (-1)^n / (2n)!
  constexpr float s3 = (float)(-1.0 / 6.0), s5 = (float)(1.0 / 120.0);
  constexpr float s7 = (float)(-1.0 / 5040.0), s9 = (float)(1.0 / 362880.0);
  constexpr float c2 = -0.5f, c4 = (float)(1.0 / 24.0), c6 = (float)(-1.0 / 720.0);
  constexpr float c8 = (float)(1.0 / 40320.0), c10 = (float)(-1.0 / 3628800.0);
  const float ps = fmaf(r * z, fmaf(z, fmaf(z, fmaf(z, s9, s7), s5), s3), r);
  const float pc = fmaf(z, fmaf(z, fmaf(z, fmaf(z, fmaf(z, c10, c8), c6), c4), c2), 1.0f);
  const int q = (int)k & 3;  // two's complement: k mod 4 also for k < 0
  const float sa = (q & 1) ? pc : ps;
  const float ca = (q & 1) ? ps : pc;
  *s = (q & 2) ? -sa : sa;
  *c = ((q + 1) & 2) ? -ca : ca;
}

}  // namespace mtd
