// Device arithmetic shared by all kernels.
//
// Parity with the reference CPU build requires every value to be produced by the same sequence of
// IEEE-754 binary64 operations, each rounded to nearest: the code is compiled with
// -ffp-contract=off so no multiply-add is fused implicitly, and min/max reproduce std::min/std::max
// (grid.cpp:110-113, 161-163), whose NaN behaviour differs from fmin/fmax.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qdiv.hpp"

namespace spruce {

// std::min(a, b): b if b < a, else a.  std::max(a, b): b if a < b, else a.
__device__ __host__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __host__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// Physical constants of the reference (source/constants.hpp:8-19).
constexpr double K_B = 1.3807e-16;
constexpr double M_ELECTRON = 9.1094e-28;
constexpr double KAPPA_0 = 1.0e-6;
constexpr double PI = 3.14159265358979323846;
constexpr double C_LIGHT = 29979245800.0;
constexpr double E_CHARGE = 4.80320425e-10;
constexpr int N_GHOST = 2;

}  // namespace spruce
