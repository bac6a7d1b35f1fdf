// Device arithmetic shared by all kernels.
//
// Parity with the reference CPU build requires every value to be produced by the same sequence of
// IEEE-754 binary64 operations, each rounded to nearest: the code is compiled with
// -ffp-contract=off so no multiply-add is fused implicitly, and min/max reproduce std::min/std::max
// (grid.cpp:110-113, 161-163), whose NaN behaviour differs from fmin/fmax.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spruce {

// std::min(a, b): b if b < a, else a.  std::max(a, b): b if a < b, else a.
__device__ __host__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __host__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// Correctly rounded a/b from rb = RN(1/b), using one Markstein correction step
// (P. Markstein, "Computation of elementary functions on the IBM RISC System/6000 processor",
// IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point Arithmetic, thm 4.10):
// with q = RN(a*rb) within one ulp of a/b, r = a - b*q is exact in an FMA and RN(q + r*rb) = RN(a/b),
// provided nothing underflows or overflows.  Outside the normal range, and for zeros, infinities and
// NaN, the IEEE division is taken instead, so the result always equals a/b bit for bit.
// b must be a finite positive normal number and rb its correctly rounded reciprocal (both are
// geometry or physical constants prepared on the host).
__device__ __host__ __forceinline__ double qdiv(double a, double b, double rb) {
  double q = a * rb;
  double aq = __builtin_fabs(q), aa = __builtin_fabs(a);
  if (__builtin_expect(aq > 0x1p-900 && aq < 0x1p+900 && aa > 0x1p-900, 1)) {
    double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, rb, q);
  }
  if (a == 0.0) return q;  // +-0 / b keeps the sign of a
  return a / b;
}

// Physical constants of the reference (source/constants.hpp:8-19).
constexpr double K_B = 1.3807e-16;
constexpr double M_ELECTRON = 9.1094e-28;
constexpr double KAPPA_0 = 1.0e-6;
constexpr double PI = 3.14159265358979323846;
constexpr double C_LIGHT = 29979245800.0;
constexpr double E_CHARGE = 4.80320425e-10;
constexpr int N_GHOST = 2;

}  // namespace spruce
