// ThermalConduction::thermalEnergyDerivative without flux saturation at one interior cell
// (thermalconduction.cpp:158-176), written once for the scalar type S: double in the forward
// conduction kernel (solar.hip), Dn<K> in the adjoint for its local Jacobian with respect to the
// 3x3 temperatures, b_hat and the four derivatives of b_hat at the cell, a few components per
// pass.
#pragma once
#include <hip/hip_runtime.h>

#include "arith.hpp"
#include "crpow.hpp"
#include "dual.hpp"
#include "mhd.hpp"

namespace spruce {
namespace tcc {

__host__ __device__ __forceinline__ int wrp(int k, int n, int per) {
  if (per) {
    if (k < 0) k += n;
    else if (k >= n) k -= n;
  }
  return k;
}
__host__ __device__ __forceinline__ bool inside(const Mesh& M, int i, int j) {
  return i >= M.x.lo && i <= M.x.hi && j >= M.y.lo && j <= M.y.hi;
}
// boundaryInterpolate across the face whose upper cell is c on axis ax (derivs.cpp:477-487)
__host__ __device__ __forceinline__ double fint(const Axis& ax, int c, int cm1, double qlo, double qup) {
  return qdiv(qlo * ax.h[c] + qup * ax.h[cm1], ax.fden[c], ax.rfden[c]);
}
template <int K>
__host__ __device__ __forceinline__ Dn<K> fint(const Axis& ax, int c, int cm1, const Dn<K>& qlo, const Dn<K>& qup) {
  return (qlo * ax.h[c] + qup * ax.h[cm1]) / ax.fden[c];
}
__host__ __device__ __forceinline__ double dv(double a, double b, double rb) { return qdiv(a, b, rb); }
template <int K> __host__ __device__ __forceinline__ Dn<K> dv(const Dn<K>& a, double b, double) { return a / b; }
__host__ __device__ __forceinline__ double p15(double x) { return pow_1p5(x); }
__host__ __device__ __forceinline__ double p25(double x) { return pow_2p5(x); }
template <int K> __host__ __device__ __forceinline__ Dn<K> p15(const Dn<K>& x) { return dn_pow_1p5(x); }
template <int K> __host__ __device__ __forceinline__ Dn<K> p25(const Dn<K>& x) { return dn_pow_2p5(x); }
template <class S> struct Zero;
template <> struct Zero<double> { __host__ __device__ static double get() { return 0.0; } };
template <int K> struct Zero<Dn<K>> { __host__ __device__ static Dn<K> get() { return dn_const<K>(0.0); } };
template <class S> __host__ __device__ __forceinline__ S zero() { return Zero<S>::get(); }

// T[a][b] is the temperature of cell (i-1+a, j-1+b) (periodic neighbours wrapped).
template <class S>
__host__ __device__ inline S tc_core(const Mesh& M, int i, int j, const S T[3][3], const S& bx, const S& by,
                                     const S& dbx_x, const S& dbx_y, const S& dby_x, const S& dby_y, double k0) {
  const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
  const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
  const S Tc = T[1][1];
  // face interpolants of T on the four faces of (i,j)
  const S Txu = fint(M.x, ip, i, Tc, T[2][1]), Txl = fint(M.x, i, im, T[0][1], Tc);
  const S Tyu = fint(M.y, jp, j, Tc, T[1][2]), Tyl = fint(M.y, j, jm, T[1][0], Tc);
  const double dx = M.x.d[i], rdx = M.x.rd[i], dy = M.y.d[j], rdy = M.y.rd[j];
  const S gx = dv(Txu - Txl, dx, rdx), gy = dv(Tyu - Tyl, dy, rdy);
  // derivative1D(dtemp_dx, 1): dtemp_dx is zero outside the interior (derivs.cpp:229-233)
  auto gx_at = [&](int jj, int b1) -> S {
    if (!inside(M, i, jj)) return zero<S>();
    const S u = fint(M.x, ip, i, T[1][b1], T[2][b1]), l = fint(M.x, i, im, T[0][b1], T[1][b1]);
    return dv(u - l, dx, rdx);
  };
  auto gy_at = [&](int ii, int a1) -> S {
    if (!inside(M, ii, j)) return zero<S>();
    const S u = fint(M.y, jp, j, T[a1][1], T[a1][2]), l = fint(M.y, j, jm, T[a1][0], T[a1][1]);
    return dv(u - l, dy, rdy);
  };
  const S gxm = gx_at(jm, 0), gxp = gx_at(jp, 2);
  const S gym = gy_at(im, 0), gyp = gy_at(ip, 2);
  const S dgx_y = dv(fint(M.y, jp, j, gx, gxp) - fint(M.y, j, jm, gxm, gx), dy, rdy);
  const S dgy_x = dv(fint(M.x, ip, i, gy, gyp) - fint(M.x, i, im, gym, gy), dx, rdx);
  // secondDerivative1D (derivs.cpp:450)
  const S s0 = dv(Txu - 2.0 * Tc + Txl, M.x.hsq[i], M.x.rhsq[i]);
  const S s1 = dv(Tyu - 2.0 * Tc + Tyl, M.y.hsq[j], M.y.rhsq[j]);

  const S bdg = bx * gx + by * gy;
  const S t1x = bx * s0 + by * dgx_y, t1y = bx * dgy_x + by * s1;
  const S t2x = gx * dbx_x + gy * dbx_y, t2y = gx * dby_x + gy * dby_y;
  const S curl = dby_x - dbx_y;
  const S t3x = (-(-1.0 * curl)) * gy, t3y = (-1.0 * curl) * gx;
  const S q15 = p15(Tc), q25 = p25(Tc);
  const S ttx = 5.0 / 2.0 * q15 * gx * bdg + q25 * (t1x + t2x + t3x);
  const S tty = 5.0 / 2.0 * q15 * gy * bdg + q25 * (t1y + t2y + t3y);
  const S divb = dbx_x + dby_y;
  return ((q25 * bdg * divb) + (bx * ttx + by * tty)) * k0;
}

}  // namespace tcc
}  // namespace spruce
