// Correctly rounded powers x^y for the fixed exponents the solar modules use.
//
// The reference evaluates these with the C library's pow (grid.cpp:169; radiativelosses.cpp:182),
// which is accurate to about half an ulp but not correctly rounded.  The device computes the exact
// power in double-double arithmetic (error about 2^-104 relative) and rounds it once, which gives the
// correctly rounded result except when the exact power lies within ~2^-104 of a rounding boundary.
// Inputs are positive normal numbers (temperatures above temp_min > 0); zero, negative and
// non-finite inputs fall back to the special-case rules of pow.
#pragma once
#include <math.h>

#ifdef __HIPCC__
#define SPRUCE_HD __host__ __device__ __forceinline__
#else
#define SPRUCE_HD inline
#endif

namespace spruce {

// sqrt(x) as s + t with |t| <= ulp(s)/2: t = (x - s*s) / (2 s), the residual being exact in an FMA.
SPRUCE_HD void sqrt_dd(double x, double& s, double& t) {
  s = sqrt(x);
  const double r = fma(-s, s, x);
  t = r / (2.0 * s);
}

// x^1.5 = x * sqrt(x)
SPRUCE_HD double pow_1p5(double x) {
  if (!(x > 0.0) || !(x < 0x1p+600) || !(x > 0x1p-600)) return pow(x, 1.5);
  double s, t;
  sqrt_dd(x, s, t);
  const double p = x * s;
  const double e = fma(x, s, -p);
  return p + (e + x * t);
}

// x^2.5 = x^2 * sqrt(x)
SPRUCE_HD double pow_2p5(double x) {
  if (!(x > 0.0) || !(x < 0x1p+300) || !(x > 0x1p-300)) return pow(x, 2.5);
  double s, t;
  sqrt_dd(x, s, t);
  const double a = x * x;
  const double ae = fma(x, x, -a);
  const double p = a * s;
  const double pe = fma(a, s, -p);
  return p + (pe + (a * t + ae * s));
}

// x^-1.5 = 1 / (x sqrt(x))
SPRUCE_HD double pow_m1p5(double x) {
  if (!(x > 0.0) || !(x < 0x1p+600) || !(x > 0x1p-600)) return pow(x, -1.5);
  double s, t;
  sqrt_dd(x, s, t);
  const double hi = x * s;
  const double lo = fma(x, s, -hi) + x * t;
  const double q = 1.0 / hi;
  const double r = fma(-hi, q, 1.0) - lo * q;
  return fma(q, r, q);
}

// x^y for y = RN(1/3) = 1/3 - 1/(3*2^54): x^y = cbrt(x) * (1 - ln(x) / (3*2^54)) to 2^-100 relative.
SPRUCE_HD double pow_third(double x) {
  if (!(x > 0.0) || !(x < 0x1p+900) || !(x > 0x1p-900)) return pow(x, 1.0 / 3.0);
  const double c0 = cbrt(x);
  // one Newton step on c^3 = x in double-double: c = c0 - (c0^3 - x) / (3 c0^2)
  const double a = c0 * c0;
  const double ae = fma(c0, c0, -a);
  const double p = a * c0;
  const double pe = fma(a, c0, -p);
  const double res = (p - x) + (pe + ae * c0);
  const double clo = -res / (3.0 * a);
  const double delta = 0x1.5555555555555p-56;  // RN(1/(3*2^54))
  return c0 + (clo - c0 * (delta * log(x)));
}

}  // namespace spruce
