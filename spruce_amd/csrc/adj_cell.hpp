// Hand-written reverse (adjoint) of the IdealMHD per-cell time derivative of mhd_cell.hpp.
//
// The forward k(c) reads its four faces; each face reads four cells along its axis.  The reverse is
// split the same way, so that no cell accumulates from more than one thread except through the
// face scatter:
//   adj_cell_coeffs:  lambda_k(c) -> direct adjoint of c's own values + per-axis face coefficients
//                     cx(c), cy(c) (the adjoint of c's flux differences, already divided by d);
//   adj_face:         face f receives cx(f-1) - cx(f) (upper face of f-1, lower face of f) and is
//                     differentiated back onto its four cells f-2 .. f+1.
// Limiter and flow-direction branches are taken as in the forward (the derivative of the selected
// expression); min/max keep the derivative of the selected argument.
#pragma once
#include <hip/hip_runtime.h>

#include "arith.hpp"
#include "mhd.hpp"
#include "mhd_cell.hpp"

namespace spruce {

// Adjoint state layout: 8 arrays of N doubles in evolved order (rho, mom_x, mom_y, mom_z,
// thermal_energy, bi_x, bi_y, bi_z).
enum AdjVar { AV_RHO = 0, AV_MX, AV_MY, AV_MZ, AV_TE, AV_BX, AV_BY, AV_BZ };
constexpr int ADJ_NCOEF = 17;   // 11 fluxes, pressure, 3 velocities, 2 magnetic components per axis

// Reverse of mhd_rhs_cell at interior cell (i, j): lam = dJ/dk(c).  Writes the direct adjoint of
// the cell's own variables to own[8] and the face coefficients to cx, cy.
__host__ __device__ inline void adj_cell_coeffs(const MhdSet& S, const MhdStatic& st, const MhdConst& K,
                                                const Mesh& M, int i, int j, const double lam[8], double own[8],
                                                double cx[ADJ_NCOEF], double cy[ADJ_NCOEF]) {
  Face xl, xu, yl, yu;
  compute_face<0>(S, st, K, M, i, j, xl);
  compute_face<0>(S, st, K, M, wrapi(i + 1, M.x.n, M.x.per), j, xu);
  compute_face<1>(S, st, K, M, i, j, yl);
  compute_face<1>(S, st, K, M, i, wrapi(j + 1, M.y.n, M.y.per), yu);
  const size_t c = (size_t)i * M.ny + j;
  const double rdx = M.x.rd[i], rdy = M.y.rd[j];
  const double dx = M.x.d[i], dy = M.y.d[j];
  const double dvx_x = (xu.iv[0] - xl.iv[0]) / dx, dvx_y = (yu.iv[0] - yl.iv[0]) / dy;
  const double dvy_x = (xu.iv[1] - xl.iv[1]) / dx, dvy_y = (yu.iv[1] - yl.iv[1]) / dy;
  const double dvz_x = (xu.iv[2] - xl.iv[2]) / dx, dvz_y = (yu.iv[2] - yl.iv[2]) / dy;
  const double dbiy_x = (xu.ib[1] - xl.ib[1]) / dx, dbix_y = (yu.ib[0] - yl.ib[0]) / dy;
  const double dbiz_x = (xu.ib[2] - xl.ib[2]) / dx, dbiz_y = (yu.ib[2] - yl.ib[2]) / dy;
  const double press = S.te[c] * K.gm1;
  const double bix = S.bx[c], biy = S.by[c], biz = S.bz[c];
  const double bex = st.bex[c], bey = st.bey[c], bez = st.bez[c];
  const double r4pi = K.r4pi;
  const double curl = (dbiy_x - dbix_y) * r4pi;
  const double cz0 = dbiz_y, cz1 = -dbiz_x;
  const double bzi = biz * r4pi, bze = bez * r4pi;
  const double l0 = lam[0], l1 = lam[1], l2 = lam[2], l3 = lam[3], l4 = lam[4], l5 = lam[5], l6 = lam[6], l7 = lam[7];

  for (int v = 0; v < 8; ++v) own[v] = 0.0;
  // transport divergences: k0 = -tdiv(rho), k1 = -tdiv(mx)..., k5 = -tdiv(bix) - tdiv(bex) ...
  const double gt[T_N] = {-l0, -l1, -l2, -l3, -l4, -l5, -l5, -l6, -l6, -l7, -l7};
  for (int q = 0; q < T_N; ++q) { cx[q] = gt[q] * rdx; cy[q] = gt[q] * rdy; }
  // pressure gradients
  cx[11] = -l1 * rdx;
  cy[11] = -l2 * rdy;
  // velocity derivatives: k4 -= press (dvx_x + dvy_y); k5..k7 += (b_x) dv?_x + (b_y) dv?_y
  const double bxs = bix + bex, bys = biy + bey;
  const double g_dvx_x = -l4 * press + l5 * bxs, g_dvy_y = -l4 * press + l6 * bys;
  const double g_dvx_y = l5 * bys, g_dvy_x = l6 * bxs, g_dvz_x = l7 * bxs, g_dvz_y = l7 * bys;
  cx[12] = g_dvx_x * rdx; cx[13] = g_dvy_x * rdx; cx[14] = g_dvz_x * rdx;
  cy[12] = g_dvx_y * rdy; cy[13] = g_dvy_y * rdy; cy[14] = g_dvz_y * rdy;
  own[AV_TE] += -l4 * (dvx_x + dvy_y) * K.gm1;
  own[AV_BX] += l5 * dvx_x + l6 * dvy_x + l7 * dvz_x;
  own[AV_BY] += l5 * dvx_y + l6 * dvy_y + l7 * dvz_y;
  // gravity
  own[AV_RHO] += l1 * st.gx[c] + l2 * st.gy[c];
  // Lorentz forces: ext0 = -curl bey, ext1 = curl bex, in0 = -curl biy, in1 = curl bix
  const double g_curl = -l1 * bey + l2 * bex - l1 * biy + l2 * bix;
  own[AV_BY] += -l1 * curl;
  own[AV_BX] += l2 * curl;
  // inz0 = bzi cz1, inz1 = -bzi cz0, extz0 = bze cz1, extz1 = -bze cz0 (momentum x, y)
  // fz_ext = (cz0 bey - cz1 bex)/4pi, fz_in = (cz0 biy - cz1 bix)/4pi (momentum z)
  const double g_bzi = l1 * cz1 - l2 * cz0;
  const double g_cz1 = l1 * (bzi + bze) - l3 * (bex + bix) * r4pi;
  const double g_cz0 = -l2 * (bzi + bze) + l3 * (bey + biy) * r4pi;
  own[AV_BY] += l3 * cz0 * r4pi;
  own[AV_BX] += -l3 * cz1 * r4pi;
  own[AV_BZ] += g_bzi * r4pi;
  // cz0 = d(bi_z)/dy, cz1 = -d(bi_z)/dx, curl = (d(bi_y)/dx - d(bi_x)/dy)/4pi
  const double g_dbiz_y = g_cz0, g_dbiz_x = -g_cz1;
  const double g_dbiy_x = g_curl * r4pi, g_dbix_y = -g_curl * r4pi;
  cx[15] = g_dbiy_x * rdx; cx[16] = g_dbiz_x * rdx;   // x faces: interpolated bi_y, bi_z
  cy[15] = g_dbix_y * rdy; cy[16] = g_dbiz_y * rdy;   // y faces: interpolated bi_x, bi_z
}

// Reverse of compute_face<AX> for the face between cells c-1 and c ((i, j) is cell c), given the
// face adjoint fa[ADJ_NCOEF].  add(slot, var, value) accumulates into the adjoint of cell
// c-2+slot along the axis (slot 0..3: c-2, c-1, c, c+1).
template <int AX, class ADD>
__host__ __device__ inline void adj_face_slots(const MhdSet& S, const MhdStatic& st, const MhdConst& K, const Mesh& M,
                                               int i, int j, const double fa[ADJ_NCOEF], ADD&& add) {
  const Axis& ax = AX == 0 ? M.x : M.y;
  const int c = AX == 0 ? i : j;
  const int cm2 = wrapi(c - 2, ax.n, ax.per), cm1 = wrapi(c - 1, ax.n, ax.per), cp1 = wrapi(c + 1, ax.n, ax.per);
  auto at = [&](int k) -> size_t { return AX == 0 ? (size_t)k * M.ny + j : (size_t)i * M.ny + k; };
  const size_t i0 = at(cm2), i1 = at(cm1), i2 = at(c), i3 = at(cp1);
  const double hlo = ax.h[cm1], hup = ax.h[c];
  const double den = ax.fden[c], rden = ax.rfden[c];
  const double wlo = hup * rden, wup = hlo * rden;   // d interp / d q_lower, d interp / d q_upper
  auto interp = [&](double a, double b) { return qdiv(a * hup + b * hlo, den, rden); };
  const double r1 = S.rho[i1], r2 = S.rho[i2];
  const double m1[3] = {S.mx[i1], S.my[i1], S.mz[i1]}, m2[3] = {S.mx[i2], S.my[i2], S.mz[i2]};
  const double v1[3] = {m1[0] / r1, m1[1] / r1, m1[2] / r1};
  const double v2[3] = {m2[0] / r2, m2[1] / r2, m2[2] / r2};
  const double vf = interp(v1[AX], v2[AX]);
  const bool pos = vf > 0.0, neg = vf < 0.0;
  const double g = pos ? ax.gup[c] : ax.gdn[c];
  const int ce = pos ? cm1 : cp1;
  const double eden = ax.fden[ce], reden = ax.rfden[ce];
  const double t = g * reden;                        // d d1 / d b; d d1 / d a = 1 - t
  const double* src[T_N] = {S.rho, S.mx, S.my, S.mz, S.te, S.bx, st.bex, S.by, st.bey, S.bz, st.bez};
  const int var_of[T_N] = {AV_RHO, AV_MX, AV_MY, AV_MZ, AV_TE, AV_BX, -1, AV_BY, -1, AV_BZ, -1};
  double g_vf = 0.0;
  for (int q = 0; q < T_N; ++q) {
    const double q0 = src[q][i0], q1 = src[q][i1], q2 = src[q][i2], q3 = src[q][i3];
    const double d2 = interp(q1, q2);
    const double a = pos ? q0 : q3, b = pos ? q1 : q2;
    const double d1 = a + qdiv((b - a) * g, eden, reden);
    const bool le = q2 <= q1;
    const bool outer_min = pos ? le : !le;
    // selection: 0 -> d2, 1 -> d1, 2 -> b
    int sel;
    double inner;
    int inner_sel;
    if (outer_min) { inner_sel = (d1 < d2) ? 0 : 1; inner = (d1 < d2) ? d2 : d1; }
    else { inner_sel = (d2 < d1) ? 0 : 1; inner = (d2 < d1) ? d2 : d1; }
    if (outer_min) sel = (inner < b) ? inner_sel : 2;
    else sel = (b < inner) ? inner_sel : 2;
    if (!(pos || neg)) sel = 0;
    const double s = sel == 0 ? d2 : (sel == 1 ? d1 : b);
    g_vf += fa[q] * s;
    const int v = var_of[q];
    if (v < 0) continue;
    const double gs = fa[q] * vf;
    const int ia = pos ? 0 : 3, ib = pos ? 1 : 2;
    if (sel == 0) { add(1, v, gs * wlo); add(2, v, gs * wup); }
    else if (sel == 1) { add(ia, v, gs * (1.0 - t)); add(ib, v, gs * t); }
    else add(ib, v, gs);
  }
  // interpolated velocities (iv[AX] is also the face velocity), pressure and fields
  double giv[3] = {fa[12], fa[13], fa[14]};
  giv[AX] += g_vf;
  for (int k = 0; k < 3; ++k) {
    const double g1 = giv[k] * wlo, g2 = giv[k] * wup;
    add(1, AV_MX + k, g1 / r1);
    add(1, AV_RHO, -g1 * v1[k] / r1);
    add(2, AV_MX + k, g2 / r2);
    add(2, AV_RHO, -g2 * v2[k] / r2);
  }
  add(1, AV_TE, fa[11] * wlo * K.gm1);
  add(2, AV_TE, fa[11] * wup * K.gm1);
  const int bfirst = AX == 0 ? AV_BY : AV_BX;        // x faces carry bi_y, y faces bi_x
  add(1, bfirst, fa[15] * wlo);
  add(2, bfirst, fa[15] * wup);
  add(1, AV_BZ, fa[16] * wlo);
  add(2, AV_BZ, fa[16] * wup);
}

// adj_face_slots with add(cell_index, var, value) on the grid's cell indices.
template <int AX, class ADD>
__host__ __device__ inline void adj_face(const MhdSet& S, const MhdStatic& st, const MhdConst& K, const Mesh& M,
                                         int i, int j, const double fa[ADJ_NCOEF], ADD&& add) {
  const Axis& ax = AX == 0 ? M.x : M.y;
  const int c = AX == 0 ? i : j;
  auto at = [&](int k) -> size_t { return AX == 0 ? (size_t)k * M.ny + j : (size_t)i * M.ny + k; };
  const size_t cell[4] = {at(wrapi(c - 2, ax.n, ax.per)), at(wrapi(c - 1, ax.n, ax.per)), at(c),
                          at(wrapi(c + 1, ax.n, ax.per))};
  adj_face_slots<AX>(S, st, K, M, i, j, fa, [&](int slot, int v, double val) { add(cell[slot], v, val); });
}

// Forward-mode number for the gradient of the per-cell step size with respect to the cell's 8
// evolved variables (used once per step, at the cell that sets the step).
struct Dual8 {
  double v;
  double d[8];
};
__host__ __device__ inline Dual8 dconst(double x) { Dual8 r; r.v = x; for (int k = 0; k < 8; ++k) r.d[k] = 0.0; return r; }
__host__ __device__ inline Dual8 dvar(double x, int k) { Dual8 r = dconst(x); r.d[k] = 1.0; return r; }
__host__ __device__ inline Dual8 operator+(const Dual8& a, const Dual8& b) { Dual8 r; r.v = a.v + b.v; for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] + b.d[k]; return r; }
__host__ __device__ inline Dual8 operator-(const Dual8& a, const Dual8& b) { Dual8 r; r.v = a.v - b.v; for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] - b.d[k]; return r; }
__host__ __device__ inline Dual8 operator*(const Dual8& a, const Dual8& b) { Dual8 r; r.v = a.v * b.v; for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
__host__ __device__ inline Dual8 operator/(const Dual8& a, const Dual8& b) {
  Dual8 r; r.v = a.v / b.v;
  for (int k = 0; k < 8; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
__host__ __device__ inline Dual8 dsqrt(const Dual8& a) {
  Dual8 r; r.v = sqrt(a.v);
  // at 0 the one-sided derivative is taken as 0 (|v| = sqrt(v*v) at v = 0)
  for (int k = 0; k < 8; ++k) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
  return r;
}
// std::max semantics: (a < b) ? b : a
__host__ __device__ inline Dual8 dmax(const Dual8& a, const Dual8& b) { return (a.v < b.v) ? b : a; }

// IdealMHD::recomputeDT at one cell (idealmhd.cpp:279-304) with its gradient.
__host__ __device__ inline Dual8 dt_cell_dual(const MhdSet& L, const MhdStatic& st, const MhdConst& K, const Mesh& M,
                                              int i, int j) {
  const size_t c = (size_t)i * M.ny + j;
  const Dual8 rho = dvar(L.rho[c], AV_RHO), mx = dvar(L.mx[c], AV_MX), my = dvar(L.my[c], AV_MY);
  const Dual8 te = dvar(L.te[c], AV_TE);
  const Dual8 bx = dconst(st.bex[c]) + dvar(L.bx[c], AV_BX), by = dconst(st.bey[c]) + dvar(L.by[c], AV_BY);
  const Dual8 bz = dconst(st.bez[c]) + dvar(L.bz[c], AV_BZ);
  const Dual8 press = te * dconst(K.gm1);
  const Dual8 bmag = dsqrt(bx * bx + by * by + bz * bz);
  const Dual8 vx = mx / rho, vy = my / rho;
  const Dual8 cs = dsqrt(press * dconst(K.gamma) / rho);
  const Dual8 cs2 = cs * cs;
  const Dual8 va = bmag / dsqrt(rho * dconst(4.0 * PI));
  const Dual8 va2 = va * va;
  const Dual8 sum = cs2 + va2;
  const Dual8 one = dconst(1.0);
  const Dual8 delta = dsqrt(one - cs2 * dconst(4.0) * va2 / (sum * sum));
  const Dual8 vfast = dsqrt(sum * dconst(0.5) * (one + delta));
  const Dual8 vslow = dsqrt(sum * dconst(0.5) * (one - delta));
  const Dual8 vmx = dsqrt(vx * vx), vmy = dsqrt(vy * vy);
  const Dual8 cmax = dmax(dmax(dmax(cs, va), vfast), vslow);
  return one / ((vmx + cmax) / dconst(M.x.d[i]) + (vmy + cmax) / dconst(M.y.d[j]));
}

}  // namespace spruce
