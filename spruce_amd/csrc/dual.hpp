// Forward-mode numbers with K derivative components, for the local Jacobians of the module kernels'
// per-cell functions in the adjoint (the value follows the forward's arithmetic; the derivative is
// that of the exact function, taking the branch the value takes at kinks).
#pragma once
#include <hip/hip_runtime.h>

#include "arith.hpp"
#include "crpow.hpp"

namespace spruce {

template <int K>
struct Dn {
  double v;
  double d[K];
};

template <int K> __host__ __device__ inline Dn<K> dn_const(double x) {
  Dn<K> r;
  r.v = x;
  for (int k = 0; k < K; ++k) r.d[k] = 0.0;
  return r;
}
template <int K> __host__ __device__ inline Dn<K> dn_var(double x, int k) {
  Dn<K> r = dn_const<K>(x);
  r.d[k] = 1.0;
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator+(const Dn<K>& a, const Dn<K>& b) {
  Dn<K> r; r.v = a.v + b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator-(const Dn<K>& a, const Dn<K>& b) {
  Dn<K> r; r.v = a.v - b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator-(const Dn<K>& a) {
  Dn<K> r; r.v = -a.v;
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator*(const Dn<K>& a, const Dn<K>& b) {
  Dn<K> r; r.v = a.v * b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator/(const Dn<K>& a, const Dn<K>& b) {
  Dn<K> r; r.v = a.v / b.v;
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator+(const Dn<K>& a, double b) {
  Dn<K> r = a; r.v = a.v + b; return r;
}
template <int K> __host__ __device__ inline Dn<K> operator+(double a, const Dn<K>& b) { return b + a; }
template <int K> __host__ __device__ inline Dn<K> operator-(const Dn<K>& a, double b) {
  Dn<K> r = a; r.v = a.v - b; return r;
}
template <int K> __host__ __device__ inline Dn<K> operator-(double a, const Dn<K>& b) {
  Dn<K> r; r.v = a - b.v;
  for (int k = 0; k < K; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator*(const Dn<K>& a, double b) {
  Dn<K> r; r.v = a.v * b;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator*(double a, const Dn<K>& b) { return b * a; }
template <int K> __host__ __device__ inline Dn<K> operator/(const Dn<K>& a, double b) {
  Dn<K> r; r.v = a.v / b;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int K> __host__ __device__ inline Dn<K> operator/(double a, const Dn<K>& b) {
  Dn<K> r; r.v = a / b.v;
  for (int k = 0; k < K; ++k) r.d[k] = -r.v * b.d[k] / b.v;
  return r;
}
template <int K> __host__ __device__ inline Dn<K> dn_sqrt(const Dn<K>& a) {
  Dn<K> r; r.v = sqrt(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
  return r;
}
// x^1.5 and x^2.5 with the forward's correctly rounded values
template <int K> __host__ __device__ inline Dn<K> dn_pow_1p5(const Dn<K>& a) {
  Dn<K> r; r.v = pow_1p5(a.v);
  const double g = 1.5 * sqrt(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = g * a.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> dn_pow_2p5(const Dn<K>& a) {
  Dn<K> r; r.v = pow_2p5(a.v);
  const double g = 2.5 * pow_1p5(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = g * a.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> dn_pow_m1p5(const Dn<K>& a) {
  Dn<K> r; r.v = pow_m1p5(a.v);
  const double g = -1.5 * r.v / a.v;
  for (int k = 0; k < K; ++k) r.d[k] = g * a.d[k];
  return r;
}
template <int K> __host__ __device__ inline Dn<K> dn_pow_third(const Dn<K>& a) {
  Dn<K> r; r.v = pow_third(a.v);
  const double g = (1.0 / 3.0) * r.v / a.v;
  for (int k = 0; k < K; ++k) r.d[k] = g * a.d[k];
  return r;
}
// std::max / std::min semantics: (a < b) ? b : a and (b < a) ? b : a
template <int K> __host__ __device__ inline Dn<K> dn_max(const Dn<K>& a, const Dn<K>& b) { return (a.v < b.v) ? b : a; }
template <int K> __host__ __device__ inline Dn<K> dn_min(const Dn<K>& a, const Dn<K>& b) { return (b.v < a.v) ? b : a; }
template <int K> __host__ __device__ inline Dn<K> dn_max(const Dn<K>& a, double b) { return (a.v < b) ? dn_const<K>(b) : a; }
template <int K> __host__ __device__ inline Dn<K> dn_min(const Dn<K>& a, double b) { return (b < a.v) ? dn_const<K>(b) : a; }

}  // namespace spruce
