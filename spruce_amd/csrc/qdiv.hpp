// Correctly rounded division by a prepared reciprocal, for device and host code.
//
// qdiv(a, b, rb) with rb = RN(1/b) returns RN(a/b), bit for bit the IEEE quotient: q = RN(a*rb) is
// within one ulp of a/b, r = a - b*q is exact in an FMA, and RN(q + r*rb) = RN(a/b) (Markstein's
// theorem: P. Markstein, "Computation of elementary functions on the IBM RISC System/6000
// processor", IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point Arithmetic,
// thm 4.10), provided nothing underflows or overflows: |a| and |q| within 2^-900 .. 2^900.  Zeros,
// infinities, NaN and extreme exponents take the IEEE division, so the result always equals a/b.
// b must be a finite positive normal number and rb its correctly rounded reciprocal (both are
// geometry or physical constants prepared on the host).
#pragma once
#include <math.h>

#ifdef __HIPCC__
#define SPRUCE_QD __host__ __device__ __forceinline__
#else
#define SPRUCE_QD inline
#endif

namespace spruce {

SPRUCE_QD double qdiv(double a, double b, double rb) {
  const double q = a * rb;
  const double aq = fabs(q), aa = fabs(a);
  if (__builtin_expect(aq > 0x1p-900 && aq < 0x1p+900 && aa > 0x1p-900, 1)) {
    const double r = fma(-b, q, a);
    return fma(r, rb, q);
  }
  if (a == 0.0) return q;  // +-0 / b keeps the sign of a
  return a / b;
}

// whether qdiv takes the corrected-reciprocal path for (a, q = a*rb)
SPRUCE_QD bool qdiv_safe(double a, double q) {
  return fabs(q) > 0x1p-900 && fabs(q) < 0x1p+900 && fabs(a) > 0x1p-900;
}

}  // namespace spruce
