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

// The energy derivative from the cell's temperature Tc, its gradient (gx, gy), second derivatives
// (s0 = d2T/dx2, s1 = d2T/dy2), the cross derivatives of the gradient (dgx_y = d(dT/dx)/dy,
// dgy_x = d(dT/dy)/dx), b_hat and the four derivatives of b_hat (thermalconduction.cpp:158-176,
// operation for operation).
template <class S>
__host__ __device__ __forceinline__ S tc_combine(const S& Tc, const S& gx, const S& gy, const S& s0, const S& s1,
                                                 const S& dgx_y, const S& dgy_x, const S& bx, const S& by,
                                                 const S& dbx_x, const S& dbx_y, const S& dby_x, const S& dby_y,
                                                 double k0) {
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

  return tc_combine<S>(Tc, gx, gy, s0, s1, dgx_y, dgy_x, bx, by, dbx_x, dbx_y, dby_x, dby_y, k0);
}

// Adjoints of tc_combine's arguments for the seed ld of its value.
struct TcCombAdj {
  double T, gx, gy, s0, s1, dgx_y, dgy_x, bx, by, dbx_x, dbx_y, dby_x, dby_y;
};

// Reverse mode of tc_combine: returns its value and sets g to ld times the partial derivatives with
// respect to its thirteen arguments.  CR: the powers of Tc correctly rounded, as tc_combine takes
// them (the value is then tc_combine's); otherwise Tc sqrt(Tc) and Tc^2 sqrt(Tc) (within a few ulps).
template <bool CR = true>
__host__ __device__ inline double tc_combine_reverse(double Tc, double gx, double gy, double s0, double s1,
                                                     double dgx_y, double dgy_x, double bx, double by, double dbx_x,
                                                     double dbx_y, double dby_x, double dby_y, double k0, double ld,
                                                     TcCombAdj& g) {
  const double bdg = bx * gx + by * gy;
  const double t1x = bx * s0 + by * dgx_y, t1y = bx * dgy_x + by * s1;
  const double t2x = gx * dbx_x + gy * dbx_y, t2y = gx * dby_x + gy * dby_y;
  const double curl = dby_x - dbx_y;
  const double t3x = (-(-1.0 * curl)) * gy, t3y = (-1.0 * curl) * gx;
  const double sq = sqrt(Tc);
  const double q15 = CR ? pow_1p5(Tc) : Tc * sq;
  const double q25 = CR ? pow_2p5(Tc) : Tc * q15;
  const double ttx = 5.0 / 2.0 * q15 * gx * bdg + q25 * (t1x + t2x + t3x);
  const double tty = 5.0 / 2.0 * q15 * gy * bdg + q25 * (t1y + t2y + t3y);
  const double divb = dbx_x + dby_y;
  const double R = ((q25 * bdg * divb) + (bx * ttx + by * tty)) * k0;
  const double gR = ld * k0;
  double g_q25 = gR * bdg * divb, g_bdg = gR * q25 * divb;
  const double g_divb = gR * q25 * bdg;
  double g_bx = gR * ttx, g_by = gR * tty;
  const double g_ttx = gR * bx, g_tty = gR * by;
  const double g_q15 = g_ttx * 2.5 * gx * bdg + g_tty * 2.5 * gy * bdg;
  double g_gx = g_ttx * 2.5 * q15 * bdg, g_gy = g_tty * 2.5 * q15 * bdg;
  g_bdg += g_ttx * 2.5 * q15 * gx + g_tty * 2.5 * q15 * gy;
  g_q25 += g_ttx * (t1x + t2x + t3x) + g_tty * (t1y + t2y + t3y);
  const double gSx = g_ttx * q25, gSy = g_tty * q25;
  double g_dbx_x = g_divb, g_dby_y = g_divb, g_dbx_y = 0.0, g_dby_x = 0.0;
  // t3x = curl gy, t3y = -curl gx; curl = dby_x - dbx_y
  const double g_curl = gSx * gy - gSy * gx;
  g_gy += gSx * curl;
  g_gx -= gSy * curl;
  g_dby_x += g_curl;
  g_dbx_y -= g_curl;
  // t2
  g_gx += gSx * dbx_x + gSy * dby_x;
  g_gy += gSx * dbx_y + gSy * dby_y;
  g_dbx_x += gSx * gx;
  g_dbx_y += gSx * gy;
  g_dby_x += gSy * gx;
  g_dby_y += gSy * gy;
  // t1
  g_bx += gSx * s0 + gSy * dgy_x;
  g_by += gSx * dgx_y + gSy * s1;
  g.s0 = gSx * bx;
  g.dgx_y = gSx * by;
  g.dgy_x = gSy * bx;
  g.s1 = gSy * by;
  // bdg
  g_bx += g_bdg * gx;
  g_by += g_bdg * gy;
  g_gx += g_bdg * bx;
  g_gy += g_bdg * by;
  // powers of Tc
  g.T = g_q15 * 1.5 * sq + g_q25 * 2.5 * q15;
  g.gx = g_gx;
  g.gy = g_gy;
  g.bx = g_bx;
  g.by = g_by;
  g.dbx_x = g_dbx_x;
  g.dbx_y = g_dbx_y;
  g.dby_x = g_dby_x;
  g.dby_y = g_dby_y;
  return R;
}

// Reverse mode of tc_core<double> at interior cell (i, j): returns the value and adds ld * dR/dT to
// gT[3][3] and ld * dR/d(b_hat_x, b_hat_y, d b_hat_x/dx, /dy, d b_hat_y/dx, /dy) to gfz[6].  The
// partial derivatives are those of the exact function, with the geometric divisions taken as
// products with the prepared reciprocals (a relative error of an ulp in a derivative); the value is
// tc_core's.
__host__ __device__ inline double tc_core_reverse(const Mesh& M, int i, int j, const double T[3][3], double bx,
                                                  double by, double dbx_x, double dbx_y, double dby_x, double dby_y,
                                                  double k0, double ld, double gT[3][3], double gfz[6]) {
  const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
  const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
  const Axis& X = M.x;
  const Axis& Y = M.y;
  const double dx = X.d[i], rdx = X.rd[i], dy = Y.d[j], rdy = Y.rd[j];
  // interpolation weights: upper face (c, c+1) and lower face (c-1, c) on each axis
  const double xu_c = X.h[ip] * X.rfden[ip], xu_p = X.h[i] * X.rfden[ip];
  const double xl_m = X.h[i] * X.rfden[i], xl_c = X.h[im] * X.rfden[i];
  const double yu_c = Y.h[jp] * Y.rfden[jp], yu_p = Y.h[j] * Y.rfden[jp];
  const double yl_m = Y.h[j] * Y.rfden[j], yl_c = Y.h[jm] * Y.rfden[j];
  // forward, as tc_core
  const double Tc = T[1][1];
  const double Txu = fint(X, ip, i, Tc, T[2][1]), Txl = fint(X, i, im, T[0][1], Tc);
  const double Tyu = fint(Y, jp, j, Tc, T[1][2]), Tyl = fint(Y, j, jm, T[1][0], Tc);
  const double gx = qdiv(Txu - Txl, dx, rdx), gy = qdiv(Tyu - Tyl, dy, rdy);
  const bool in_jm = inside(M, i, jm), in_jp = inside(M, i, jp), in_im = inside(M, im, j), in_ip = inside(M, ip, j);
  auto gxc = [&](int b1) {
    return qdiv(fint(X, ip, i, T[1][b1], T[2][b1]) - fint(X, i, im, T[0][b1], T[1][b1]), dx, rdx);
  };
  auto gyr = [&](int a1) {
    return qdiv(fint(Y, jp, j, T[a1][1], T[a1][2]) - fint(Y, j, jm, T[a1][0], T[a1][1]), dy, rdy);
  };
  const double gxm = in_jm ? gxc(0) : 0.0, gxp = in_jp ? gxc(2) : 0.0;
  const double gym = in_im ? gyr(0) : 0.0, gyp = in_ip ? gyr(2) : 0.0;
  const double dgx_y = qdiv(fint(Y, jp, j, gx, gxp) - fint(Y, j, jm, gxm, gx), dy, rdy);
  const double dgy_x = qdiv(fint(X, ip, i, gy, gyp) - fint(X, i, im, gym, gy), dx, rdx);
  const double s0 = qdiv(Txu - 2.0 * Tc + Txl, X.hsq[i], X.rhsq[i]);
  const double s1 = qdiv(Tyu - 2.0 * Tc + Tyl, Y.hsq[j], Y.rhsq[j]);
  TcCombAdj G;
  const double R = tc_combine_reverse(Tc, gx, gy, s0, s1, dgx_y, dgy_x, bx, by, dbx_x, dbx_y, dby_x, dby_y, k0, ld, G);
  double g_Tc = G.T, g_gx = G.gx, g_gy = G.gy;
  const double g_s0 = G.s0, g_s1 = G.s1, g_dgx_y = G.dgx_y, g_dgy_x = G.dgy_x;
  // second derivatives
  const double sx = g_s0 * X.rhsq[i], sy = g_s1 * Y.rhsq[j];
  double g_Txu = sx, g_Txl = sx;
  g_Tc -= 2.0 * sx;
  double g_Tyu = sy, g_Tyl = sy;
  g_Tc -= 2.0 * sy;
  // nested derivatives
  const double ry = g_dgx_y * rdy;
  g_gx += ry * (yu_c - yl_c);
  const double g_gxp = ry * yu_p, g_gxm = -ry * yl_m;
  const double rx = g_dgy_x * rdx;
  g_gy += rx * (xu_c - xl_c);
  const double g_gyp = rx * xu_p, g_gym = -rx * xl_m;
  auto rev_gxc = [&](int b1, double g) {   // gx of column b1 (cells (i-1..i+1, j-1+b1))
    const double r = g * rdx;
    gT[1][b1] += r * (xu_c - xl_c);
    gT[2][b1] += r * xu_p;
    gT[0][b1] -= r * xl_m;
  };
  auto rev_gyr = [&](int a1, double g) {   // gy of row a1 (cells (i-1+a1, j-1..j+1))
    const double r = g * rdy;
    gT[a1][1] += r * (yu_c - yl_c);
    gT[a1][2] += r * yu_p;
    gT[a1][0] -= r * yl_m;
  };
  if (in_jm) rev_gxc(0, g_gxm);
  if (in_jp) rev_gxc(2, g_gxp);
  if (in_im) rev_gyr(0, g_gym);
  if (in_ip) rev_gyr(2, g_gyp);
  // first derivatives at the cell
  g_Txu += g_gx * rdx;
  g_Txl -= g_gx * rdx;
  g_Tyu += g_gy * rdy;
  g_Tyl -= g_gy * rdy;
  // face interpolants
  g_Tc += g_Txu * xu_c + g_Txl * xl_c + g_Tyu * yu_c + g_Tyl * yl_c;
  gT[2][1] += g_Txu * xu_p;
  gT[0][1] += g_Txl * xl_m;
  gT[1][2] += g_Tyu * yu_p;
  gT[1][0] += g_Tyl * yl_m;
  gT[1][1] += g_Tc;
  gfz[0] += G.bx;
  gfz[1] += G.by;
  gfz[2] += G.dbx_x;
  gfz[3] += G.dbx_y;
  gfz[4] += G.dby_x;
  gfz[5] += G.dby_y;
  return R;
}

}  // namespace tcc
}  // namespace spruce
