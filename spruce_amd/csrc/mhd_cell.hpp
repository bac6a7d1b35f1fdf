// Per-cell IdealMHD time derivative (idealmhd.cpp:42-105 with derivs.cpp:10-264), shared by the
// stage kernels and the adjoint, callable on the host for the adjoint's algebra checks.
#pragma once
#include <hip/hip_runtime.h>

#include "arith.hpp"
#include "mhd.hpp"

namespace spruce {

// Transported quantities in the order idealmhd.cpp:52-86 transports them.
enum TQ { T_RHO = 0, T_MX, T_MY, T_MZ, T_TE, T_BIX, T_BEX, T_BIY, T_BEY, T_BIZ, T_BEZ, T_N };

__host__ __device__ __forceinline__ int wrapi(int k, int n, int per) {
  if (per) {
    if (k < 0) k += n;
    else if (k >= n) k -= n;
  }
  return k;
}

// Everything a cell needs from one of its faces.
struct Face {
  double vf;       // face velocity along the axis (boundaryInterpolate of v, derivs.cpp:50,155)
  double s[T_N];   // Barton upwind surface values (derivs.cpp:47-68)
  double iv[3];    // interpolated v_x, v_y, v_z (derivative1D, derivs.cpp:259)
  double ip;       // interpolated pressure
  double ib[3];    // interpolated bi_x, bi_y, bi_z (== the upwind d2 of those quantities)
};

// Face between cells c-1 and c along axis AX; (i, j) is the cell c.
template <int AX>
__host__ __device__ __forceinline__ void compute_face(const MhdSet& S, const MhdStatic& st, const MhdConst& K,
                                             const Mesh& M, int i, int j, Face& F) {
  const Axis& ax = AX == 0 ? M.x : M.y;
  const int c = AX == 0 ? i : j;
  const int cm2 = wrapi(c - 2, ax.n, ax.per), cm1 = wrapi(c - 1, ax.n, ax.per), cp1 = wrapi(c + 1, ax.n, ax.per);
  auto at = [&](int k) -> size_t { return AX == 0 ? (size_t)k * M.ny + j : (size_t)i * M.ny + k; };
  const size_t i0 = at(cm2), i1 = at(cm1), i2 = at(c), i3 = at(cp1);
  const double hlo = ax.h[cm1], hup = ax.h[c];
  const double den = ax.fden[c], rden = ax.rfden[c];
  // boundaryInterpolate(q, lower, upper) = (q_lo*h_up + q_up*h_lo)/(h_up + h_lo)
  auto interp = [&](double qlo, double qup) { return qdiv(qlo * hup + qup * hlo, den, rden); };

  const double r1 = S.rho[i1], r2 = S.rho[i2];
  const double v1[3] = {S.mx[i1] / r1, S.my[i1] / r1, S.mz[i1] / r1};
  const double v2[3] = {S.mx[i2] / r2, S.my[i2] / r2, S.mz[i2] / r2};
#pragma unroll
  for (int k = 0; k < 3; ++k) F.iv[k] = interp(v1[k], v2[k]);
  F.vf = F.iv[AX];
  F.ip = interp(S.te[i1] * K.gm1, S.te[i2] * K.gm1);

  const double vf = F.vf;
  const bool pos = vf > 0.0, neg = vf < 0.0;
  // boundaryExtrapolate through (c-2, c-1) when the flow is positive, through (c+1, c) otherwise
  // (derivs.cpp:52,60,490-499).
  const double g = pos ? ax.gup[c] : ax.gdn[c];
  const int ce = pos ? cm1 : cp1;
  const double eden = ax.fden[ce], reden = ax.rfden[ce];
  const double* src[T_N] = {S.rho, S.mx, S.my, S.mz, S.te, S.bx, st.bex, S.by, st.bey, S.bz, st.bez};
#pragma unroll
  for (int k = 0; k < T_N; ++k) {
    const double q0 = src[k][i0], q1 = src[k][i1], q2 = src[k][i2], q3 = src[k][i3];
    const double d2 = interp(q1, q2);
    const double a = pos ? q0 : q3, b = pos ? q1 : q2;
    const double d1 = a + qdiv((b - a) * g, eden, reden);
    const double d3 = b;
    // derivs.cpp:53-65: positive flow takes min(d3, max(d1,d2)) when q_c <= q_{c-1}, else
    // max(d3, min(d1,d2)); negative flow the mirror image.
    const bool le = q2 <= q1;
    const bool outer_min = pos ? le : !le;
    const double inner = outer_min ? smax(d1, d2) : smin(d1, d2);
    const double lim = outer_min ? smin(d3, inner) : smax(d3, inner);
    F.s[k] = (pos || neg) ? lim : d2;
    if (k == T_BIX) F.ib[0] = d2;
    if (k == T_BIY) F.ib[1] = d2;
    if (k == T_BIZ) F.ib[2] = d2;
  }
}

// IdealMHD::computeTimeDerivativesDerived at one interior cell (idealmhd.cpp:42-105) from its four
// faces: xl/yl are the faces computed at (i, j), xu/yu those at (i+1, j) and (i, j+1), wrapped.
__host__ __device__ __forceinline__ void mhd_rhs_faces(const MhdSet& S, const MhdStatic& st, const MhdConst& K,
                                              const Mesh& M, int i, int j, const Face& xl, const Face& xu,
                                              const Face& yl, const Face& yu, double out[8]) {
  const size_t c = (size_t)i * M.ny + j;
  const double dx = M.x.d[i], rdx = M.x.rd[i], dy = M.y.d[j], rdy = M.y.rd[j];

  // transportDerivative1D (derivs.cpp:155-156) summed over both axes (derivs.cpp:219)
  double tdiv[T_N];
#pragma unroll
  for (int k = 0; k < T_N; ++k) {
    const double tx = qdiv(xu.s[k] * xu.vf - xl.s[k] * xl.vf, dx, rdx);
    const double ty = qdiv(yu.s[k] * yu.vf - yl.s[k] * yl.vf, dy, rdy);
    tdiv[k] = tx + ty;
  }
  // derivative1D (derivs.cpp:259)
  auto ddx = [&](double up, double lo) { return qdiv(up - lo, dx, rdx); };
  auto ddy = [&](double up, double lo) { return qdiv(up - lo, dy, rdy); };
  const double dpx = ddx(xu.ip, xl.ip), dpy = ddy(yu.ip, yl.ip);
  const double dvx_x = ddx(xu.iv[0], xl.iv[0]), dvx_y = ddy(yu.iv[0], yl.iv[0]);
  const double dvy_x = ddx(xu.iv[1], xl.iv[1]), dvy_y = ddy(yu.iv[1], yl.iv[1]);
  const double dvz_x = ddx(xu.iv[2], xl.iv[2]), dvz_y = ddy(yu.iv[2], yl.iv[2]);
  const double dbiy_x = ddx(xu.ib[1], xl.ib[1]), dbix_y = ddy(yu.ib[0], yl.ib[0]);
  const double dbiz_x = ddx(xu.ib[2], xl.ib[2]), dbiz_y = ddy(yu.ib[2], yl.ib[2]);

  const double rho = S.rho[c], mx = S.mx[c], my = S.my[c], mz = S.mz[c];
  const double press = S.te[c] * K.gm1;
  const double vx = mx / rho, vy = my / rho, vz = mz / rho;
  const double bix = S.bx[c], biy = S.by[c], biz = S.bz[c];
  const double bex = st.bex[c], bey = st.bey[c], bez = st.bez[c];

  const double curl_db = qdiv(dbiy_x - dbix_y, 4.0 * PI, K.r4pi);     // idealmhd.cpp:54
  const double ext0 = (-curl_db) * bey, ext1 = curl_db * bex;         // :55 CrossProductZ2D
  const double in0 = (-curl_db) * biy, in1 = curl_db * bix;           // :56
  const double cz0 = dbiz_y, cz1 = -dbiz_x;                           // curlZ(bi_z)
  const double bzi = qdiv(biz, 4.0 * PI, K.r4pi), bze = qdiv(bez, 4.0 * PI, K.r4pi);
  const double inz0 = bzi * cz1, inz1 = (-bzi) * cz0;                 // :57 CrossProduct2DZ
  const double extz0 = bze * cz1, extz1 = (-bze) * cz0;               // :58
  const double fz_ext = qdiv(cz0 * bey - cz1 * bex, 4.0 * PI, K.r4pi);  // :59
  const double fz_in = qdiv(cz0 * biy - cz1 * bix, 4.0 * PI, K.r4pi);   // :60

  double k[8];
  k[0] = -tdiv[T_RHO];
  k[1] = -tdiv[T_MX] - dpx + rho * st.gx[c] + ext0 + in0 + inz0 + extz0;
  k[2] = -tdiv[T_MY] - dpy + rho * st.gy[c] + ext1 + in1 + inz1 + extz1;
  k[3] = -tdiv[T_MZ] + fz_ext + fz_in;
  k[4] = -tdiv[T_TE] - press * (dvx_x + dvy_y);
  k[5] = -tdiv[T_BIX] - tdiv[T_BEX] + (bix + bex) * dvx_x + (biy + bey) * dvx_y;
  k[6] = -tdiv[T_BIY] - tdiv[T_BEY] + (bix + bex) * dvy_x + (biy + bey) * dvy_y;
  k[7] = -tdiv[T_BIZ] - tdiv[T_BEZ] + (bix + bex) * dvz_x + (biy + bey) * dvz_y;
  // Ghost mask and the characteristic-boundary terms, which are zero grids without open_moc
  // boundaries; the momentum entries are rho*0 + v*0 (idealmhd.cpp:99-103, 324-329).
  out[0] = k[0] + 0.0;
  out[1] = k[1] + (rho * 0.0 + vx * 0.0);
  out[2] = k[2] + (rho * 0.0 + vy * 0.0);
  out[3] = k[3] + (rho * 0.0 + vz * 0.0);
  out[4] = k[4] + 0.0;
  out[5] = k[5] + 0.0;
  out[6] = k[6] + 0.0;
  out[7] = k[7] + 0.0;
}

__host__ __device__ __forceinline__ void mhd_rhs_cell(const MhdSet& S, const MhdStatic& st, const MhdConst& K,
                                             const Mesh& M, int i, int j, double out[8]) {
  Face xl, xu, yl, yu;
  compute_face<0>(S, st, K, M, i, j, xl);
  compute_face<0>(S, st, K, M, wrapi(i + 1, M.x.n, M.x.per), j, xu);
  compute_face<1>(S, st, K, M, i, j, yl);
  compute_face<1>(S, st, K, M, i, wrapi(j + 1, M.y.n, M.y.per), yu);
  mhd_rhs_faces(S, st, K, M, i, j, xl, xu, yl, yu, out);
}

__host__ __device__ __forceinline__ bool interior(const Mesh& M, int i, int j) {
  return i >= M.x.lo && i <= M.x.hi && j >= M.y.lo && j <= M.y.hi;
}

}  // namespace spruce
