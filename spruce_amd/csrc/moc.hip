// Open characteristic (open_moc) boundary kernels for gfx950 (reference:
// source/equationsets/idealmhd.cpp:107-224, 308-615; source/mhd/derivs.cpp:10-73, 77-117,
// 167-214, 223-264, 317-404, 417-455, 477-499).
//
// Only the ghost cells of open_moc sides are evaluated, straight from the stage's input set, in
// two passes: k_moc_ops evaluates every derivative the characteristic terms take, one thread per
// (derivative, ghost cell) in derivative-major order, and k_moc_rhs combines them, one thread per
// ghost cell.  Every whole-grid expression of singleBoundaryTermsMOC is evaluated at the cell with
// the reference's operation order, and every derivative the reference takes over a sub-range is
// zero outside that range, as the reference's zero-initialised result grids are.
#include <cfloat>

#include "../../include/spruce_gpu.h"
#include "arith.hpp"
#include "moc.hpp"

namespace spruce {
namespace {

// fields the characteristic terms differentiate
enum Fld : int {
  F_RHO, F_TE, F_P, F_VX, F_VY, F_VZ, F_BEX, F_BEY, F_BEZ, F_BIX, F_BIY, F_BIZ, F_BX, F_BY, F_BZ, F_RVX, F_RVY, F_RVZ
};

struct Ctx {
  MhdSet S;
  MhdStatic st;
  MhdConst K;
  Mesh M;
};

__device__ __forceinline__ double fv(const Ctx& c, int f, int i, int j) {
  const size_t k = (size_t)i * c.M.ny + j;
  switch (f) {
    case F_RHO: return c.S.rho[k];
    case F_TE: return c.S.te[k];
    case F_P: return c.K.gm1 * c.S.te[k];                       // idealmhd.cpp:253
    case F_VX: return c.S.mx[k] / c.S.rho[k];                   // idealmhd.cpp:248-250
    case F_VY: return c.S.my[k] / c.S.rho[k];
    case F_VZ: return c.S.mz[k] / c.S.rho[k];
    case F_BEX: return c.st.bex[k];
    case F_BEY: return c.st.bey[k];
    case F_BEZ: return c.st.bez[k];
    case F_BIX: return c.S.bx[k];
    case F_BIY: return c.S.by[k];
    case F_BIZ: return c.S.bz[k];
    case F_BX: return c.st.bex[k] + c.S.bx[k];                  // idealmhd.cpp:255-257 (be + bi)
    case F_BY: return c.st.bey[k] + c.S.by[k];
    case F_BZ: return c.st.bez[k] + c.S.bz[k];
    case F_RVX: return c.S.rho[k] * (c.S.mx[k] / c.S.rho[k]);  // grids[rho]*grids[v_x]
    case F_RVY: return c.S.rho[k] * (c.S.my[k] / c.S.rho[k]);
    default: return c.S.rho[k] * (c.S.mz[k] / c.S.rho[k]);
  }
}

// the axis by value: a reference to an axis picked at run time would place the context in private
// memory
__device__ __forceinline__ Axis axis_of(const Ctx& c, int a) { return a == 0 ? c.M.x : c.M.y; }
__device__ __forceinline__ int wrapc(const Axis& A, int k) { return A.per ? (k + A.n) % A.n : k; }
// value of field f at coordinate k along axis a through cell (i, j)
__device__ __forceinline__ double at(const Ctx& c, int f, int a, int i, int j, int k) {
  return a == 0 ? fv(c, f, k, j) : fv(c, f, i, k);
}
__device__ __forceinline__ bool in_range(const int* r, int i, int j) {
  return i >= r[0] && i <= r[2] && j >= r[1] && j <= r[3];
}

// boundaryInterpolate between coordinates k1 and k2 (derivs.cpp:477-487)
__device__ __forceinline__ double interp(const Ctx& c, int f, int a, int i, int j, int k1, int k2) {
  const Axis& A = axis_of(c, a);
  const double x1 = at(c, f, a, i, j, k1), x2 = at(c, f, a, i, j, k2);
  const double d1 = A.h[k1], d2 = A.h[k2];
  return (x1 * d2 + x2 * d1) / (d2 + d1);
}
// boundaryExtrapolate from k1 through k2 (derivs.cpp:490-499)
__device__ __forceinline__ double extrap(const Ctx& c, int f, int a, int i, int j, int k1, int k2) {
  const Axis& A = axis_of(c, a);
  const double x1 = at(c, f, a, i, j, k1), x2 = at(c, f, a, i, j, k2);
  const double d1 = A.h[k1], d2 = A.h[k2];
  return x1 + (x2 - x1) * (d1 + 2.0 * d2) / (d1 + d2);
}

// derivative1D over range r (derivs.cpp:223-264)
__device__ __forceinline__ double deriv(const Ctx& c, int f, int a, const int* r, int i, int j) {
  if (!in_range(r, i, j)) return 0.0;
  const Axis& A = axis_of(c, a);
  const int k = a == 0 ? i : j, m = wrapc(A, k - 1), p = wrapc(A, k + 1);
  return (interp(c, f, a, i, j, k, p) - interp(c, f, a, i, j, m, k)) / A.d[k];
}
// secondDerivative1D over range r (derivs.cpp:417-455)
__device__ __forceinline__ double second(const Ctx& c, int f, int a, const int* r, int i, int j) {
  if (!in_range(r, i, j)) return 0.0;
  const Axis& A = axis_of(c, a);
  const int k = a == 0 ? i : j, m = wrapc(A, k - 1), p = wrapc(A, k + 1);
  return (interp(c, f, a, i, j, k, p) - 2.0 * at(c, f, a, i, j, k) + interp(c, f, a, i, j, m, k)) / A.hsq[k];
}
// derivative1DBackward over range r (derivs.cpp:317-353)
__device__ __forceinline__ double back1(const Ctx& c, int f, int a, bool pf, const int* r, int i, int j) {
  if (!in_range(r, i, j)) return 0.0;
  const Axis& A = axis_of(c, a);
  const int k = a == 0 ? i : j, k0 = wrapc(A, pf ? k - 1 : k + 1);
  const double den = 0.5 * (A.d[k] + A.d[k0]);
  return (pf ? 1.0 : -1.0) * (at(c, f, a, i, j, k) - at(c, f, a, i, j, k0)) / den;
}
// secondDerivative1DBackward over range r (derivs.cpp:356-404)
__device__ __forceinline__ double second_back(const Ctx& c, int f, int a, bool pf, const int* r, int i, int j) {
  if (!in_range(r, i, j)) return 0.0;
  const Axis& A = axis_of(c, a);
  const int k = a == 0 ? i : j;
  const int k1 = wrapc(A, pf ? k - 1 : k + 1), k0 = wrapc(A, pf ? k - 2 : k + 2);
  const double den01 = 0.5 * (A.d[k1] + A.d[k0]);
  const double den12 = 0.5 * (A.d[k1] + A.d[k]);
  const double t12 = at(c, f, a, i, j, k1 > k ? k1 : k) - at(c, f, a, i, j, k1 < k ? k1 : k);
  const double t01 = at(c, f, a, i, j, k1 > k0 ? k1 : k0) - at(c, f, a, i, j, k1 < k0 ? k1 : k0);
  return (pf ? 1.0 : -1.0) * (t12 / den12 - t01 / den01) / den12;
}

// Barton face value of face k (between k-1 and k) with the upwind side from vel (derivs.cpp:10-73)
__device__ __forceinline__ double upwind_face(const Ctx& c, int f, int vel, int a, int i, int j, int k) {
  const Axis& A = axis_of(c, a);
  const int k1 = wrapc(A, k - 1), k0 = wrapc(A, k - 2), k3 = wrapc(A, k + 1);
  const double d2 = interp(c, f, a, i, j, k1, k);
  const double vf = interp(c, vel, a, i, j, k1, k);
  const double q2 = at(c, f, a, i, j, k), q1 = at(c, f, a, i, j, k1);
  if (vf > 0.0) {
    const double d3 = q1, d1 = extrap(c, f, a, i, j, k0, k1);
    return q2 <= q1 ? smin(d3, smax(d1, d2)) : smax(d3, smin(d1, d2));
  }
  if (vf < 0.0) {
    const double d3 = q2, d1 = extrap(c, f, a, i, j, k3, k);
    return q2 <= q1 ? smax(d3, smin(d1, d2)) : smin(d3, smax(d1, d2));
  }
  return d2;
}
// transportDerivative1D over range r (derivs.cpp:122-162)
__device__ __forceinline__ double tderiv(const Ctx& c, int f, int vel, int a, const int* r, int i, int j) {
  if (!in_range(r, i, j)) return 0.0;
  const Axis& A = axis_of(c, a);
  const int k = a == 0 ? i : j, m = wrapc(A, k - 1), p = wrapc(A, k + 1);
  const double sp = upwind_face(c, f, vel, a, i, j, p), sk = upwind_face(c, f, vel, a, i, j, k);
  return (sp * interp(c, vel, a, i, j, k, p) - sk * interp(c, vel, a, i, j, m, k)) / A.d[k];
}
// upwindSurfaceDirected face computed from cell k (derivs.cpp:77-117): face k for pf, k+1 otherwise
__device__ __forceinline__ double dir_face_of_cell(const Ctx& c, int f, int a, bool pf, int i, int j, int k) {
  const int sg = pf ? 1 : -1;
  const int k1 = k - sg, k0 = k - 2 * sg;
  const double d2 = interp(c, f, a, i, j, k1, k), d3 = at(c, f, a, i, j, k1), d1 = extrap(c, f, a, i, j, k0, k1);
  return at(c, f, a, i, j, k) <= at(c, f, a, i, j, k1) ? smin(d3, smax(d1, d2)) : smax(d3, smin(d1, d2));
}
// characteristicBartonDerivative1D at a cell of the side's ghost rows (derivs.cpp:167-214): the
// outermost face is extrapolated from the cell centre through the next face
__device__ __forceinline__ double char_barton(const Ctx& c, int f, int a, bool pf, int i, int j) {
  const Axis& A = axis_of(c, a);
  const int k = a == 0 ? i : j, n = A.n;
  double lo, hi;   // faces k and k+1
  if (pf) {
    lo = dir_face_of_cell(c, f, a, pf, i, j, k);
    if (k + 1 == n) {
      const double q = at(c, f, a, i, j, n - 1);
      hi = q + (q - lo);
    } else {
      hi = dir_face_of_cell(c, f, a, pf, i, j, k + 1);
    }
  } else {
    hi = dir_face_of_cell(c, f, a, pf, i, j, k);
    if (k == 0) {
      const double q = at(c, f, a, i, j, 0);
      lo = q - (hi - q);
    } else {
      lo = dir_face_of_cell(c, f, a, pf, i, j, k - 1);
    }
  }
  return (hi - lo) / A.d[k];
}

// The derivatives one side's terms take at one cell of its region, in a fixed order (MOC_NOPS of
// them): evaluated by one thread each in k_moc_ops, then combined per cell in side_terms.
enum MocOp : int {
  OP_CB = 0,        // 11 characteristic Barton derivatives across the side
  OP_ALT = 11,      // 5 fields x {central on the padded range, backward at the upper/lower corner} along it
  OP_SB = 26,       // 3 one-sided second derivatives across the side
  OP_DI = 29,       // 10 central derivatives along the side (parallel range)
  OP_TI = 39,       // 11 transport derivatives along the side (parallel range, velocity v_para)
  OP_SI = 50,       // 3 second derivatives along the side (parallel range)
  MOC_NOPS = 53
};
// element q of the list, as a chain of selects (a local array indexed at run time would live in
// private memory)
template <typename... T>
__device__ __forceinline__ int pick(int q, T... v) {
  int r = 0, n = 0;
  ((r = (n++ == q) ? (int)v : r), ...);
  return r;
}
__device__ __forceinline__ double eval_op(const Ctx& c, const MocSide& sd, int k, int i, int j) {
  const int a = sd.axis, pa = 1 - a;
  const bool pf = sd.lower == 0;
  const int Fvperp = a == 0 ? F_VX : F_VY, Fvpara = a == 0 ? F_VY : F_VX;
  const int Fbe_perp = a == 0 ? F_BEX : F_BEY, Fbe_para = a == 0 ? F_BEY : F_BEX;
  const int Fbi_perp = a == 0 ? F_BIX : F_BIY, Fbi_para = a == 0 ? F_BIY : F_BIX;
  const int Fb_perp = a == 0 ? F_BX : F_BY;
  const int Frv_perp = a == 0 ? F_RVX : F_RVY, Frv_para = a == 0 ? F_RVY : F_RVX;
  if (k < OP_ALT) {
    return char_barton(c, pick(k - OP_CB, Fbe_para, Fbi_para, F_BEZ, F_BIZ, Fvpara, Fvperp, F_VZ, F_P, Fb_perp, F_RHO,
                               F_TE), a, pf, i, j);
  }
  if (k < OP_SB) {
    const int q = k - OP_ALT, fld = pick(q / 3, F_P, Fbe_perp, Fbi_perp, F_BEZ, F_BIZ), w = q % 3;
    if (w == 0) return deriv(c, fld, pa, sd.rp, i, j);
    return w == 1 ? back1(c, fld, pa, true, sd.ru, i, j) : back1(c, fld, pa, false, sd.rl, i, j);
  }
  if (k < OP_DI) {
    return second_back(c, pick(k - OP_SB, Fvpara, F_VZ, Fvperp), a, pf, sd.rb, i, j);
  }
  if (k < OP_TI) {
    return deriv(c, pick(k - OP_DI, F_P, Fbe_para, Fbi_para, Fbe_perp, Fbi_perp, F_BEZ, F_BIZ, Fvpara, F_VZ, Fvperp),
                 pa, sd.ri, i, j);
  }
  if (k < OP_SI) {
    return tderiv(c, pick(k - OP_TI, F_RHO, Frv_para, F_RVZ, Frv_perp, F_TE, Fbe_para, Fbi_para, F_BEZ, F_BIZ,
                          Fbe_perp, Fbi_perp), Fvpara, pa, sd.ri, i, j);
  }
  return second(c, pick(k - OP_SI, Fvpara, F_VZ, Fvperp), pa, sd.ri, i, j);
}

// normal + parallel terms of one open_moc side at a cell of its region, in the order
// {rho, v_x, v_y, v_z, thermal_energy, b_x, b_y, b_z} (idealmhd.cpp:336-615)
__device__ void side_terms(const Ctx& c, const MocSide& sd, double vc, int i, int j, const double* op, double* out) {
  const int a = sd.axis, pa = 1 - a;
  const bool lower = sd.lower != 0, pf = !lower;
  const size_t cell = (size_t)i * c.M.ny + j;
  const int Fvperp = a == 0 ? F_VX : F_VY, Fvpara = a == 0 ? F_VY : F_VX, Fvgd = F_VZ;
  const int Fbe_perp = a == 0 ? F_BEX : F_BEY, Fbe_para = a == 0 ? F_BEY : F_BEX, Fbe_gd = F_BEZ;
  const int Fbi_perp = a == 0 ? F_BIX : F_BIY, Fbi_para = a == 0 ? F_BIY : F_BIX, Fbi_gd = F_BIZ;
  const int Fb_perp = a == 0 ? F_BX : F_BY;
  const int Frv_perp = a == 0 ? F_RVX : F_RVY, Frv_para = a == 0 ? F_RVY : F_RVX, Frv_gd = F_RVZ;
  const double gperp = a == 0 ? c.st.gx[cell] : c.st.gy[cell];
  const double gpara = a == 0 ? c.st.gy[cell] : c.st.gx[cell];

  const double bperp = fv(c, Fb_perp, i, j), bpara = fv(c, a == 0 ? F_BY : F_BX, i, j), bgd = fv(c, F_BZ, i, j);
  const double bh = sqrt(bpara * bpara + bgd * bgd);
  double s_perp = bperp > 0.0 ? 1.0 : -1.0;
  if (bperp == 0.0) s_perp = 0.0;
  const double R_para = bpara / bh, R_gd = bgd / bh;
  const double r = fv(c, F_RHO, i, j), te = fv(c, F_TE, i, j), p = fv(c, F_P, i, j);
  const double bx = fv(c, F_BX, i, j), by = fv(c, F_BY, i, j), bz = fv(c, F_BZ, i, j);
  const double bm = sqrt(bx * bx + by * by + bz * bz);
  const double css = c.K.gamma * p / r;
  const double cas = bm * bm / (4.0 * PI * r);
  const double cps = bperp * bperp / (4.0 * PI * r);
  const double sum = cas + css;
  const double root = sqrt(fabs(sum * sum / 4.0 - cps * css));
  const double cpl = sum / 2.0 + root, cmi = sum / 2.0 - root;
  const double aps = smax((css - cmi) / (cpl - cmi), 0.0), ams = smax((cpl - css) / (cpl - cmi), 0.0);
  const double c_s = sqrt(css), c_perp = sqrt(cps), c_plus = sqrt(cpl), c_minus = sqrt(cmi);
  const double al_p = sqrt(aps), al_m = sqrt(ams);

  const double* cb = op + OP_CB;   // be_para, bi_para, be_gd, bi_gd, v_para, v_perp, v_gd, press, b_perp, rho, te
  const double bpg = cb[0] + cb[1];
  const double bgg = cb[2] + cb[3];
  const double vpg = cb[4], vperpg = cb[5], vgg = cb[6], pg = cb[7];
  const double vp = fv(c, Fvperp, i, j);
  const double sg = a == 0 ? 1.0 : -1.0;
  const double sq4 = sqrt(4.0 * PI * r);
  double d1 = vp * cb[8];
  double d2 = vp * ((te + p) * cb[9] - r * cb[10]);
  double d3 = s_perp * sg * (vp + c_perp) *
              (-s_perp * R_gd * vpg + s_perp * R_para * vgg + R_gd / sq4 * bpg - R_para / sq4 * bgg);
  double d4 = s_perp * sg * (vp - c_perp) *
              (s_perp * R_gd * vpg - s_perp * R_para * vgg + R_gd / sq4 * bpg - R_para / sq4 * bgg);
  const double A1 = al_p / r * pg, A2 = al_m / r * pg;
  const double Bpm = s_perp * R_para * c_minus * al_m * vpg, Bgm = s_perp * R_gd * c_minus * al_m * vgg;
  const double Bpp = s_perp * R_para * c_plus * al_p * vpg, Bgp = s_perp * R_gd * c_plus * al_p * vgg;
  const double Cp = c_plus * al_p * vperpg, Cm = c_minus * al_m * vperpg;
  const double Epm = R_para * c_s * al_m / sq4 * bpg, Egm = R_gd * c_s * al_m / sq4 * bgg;
  const double Epp = R_para * c_s * al_p / sq4 * bpg, Egp = R_gd * c_s * al_p / sq4 * bgg;
  double d5 = (vp + c_plus) * (A1 - Bpm - Bgm + Cp + Epm + Egm);
  double d6 = (vp - c_plus) * (A1 + Bpm + Bgm - Cp + Epm + Egm);
  double d7 = (vp + c_minus) * (A2 + Bpp + Bgp + Cm - Epp - Egp);
  double d8 = (vp - c_minus) * (A2 - Bpp - Bgp - Cm - Epp - Egp);

  // alternate terms for inflowing characteristics: derivatives along the side
  const double* al = op + OP_ALT;   // press, be_perp, bi_perp, be_gd, bi_gd: central, upper, lower
  const double p_alt = al[0] + al[1] + al[2];
  const double bp_alt = al[3] + al[4] + al[5] + al[6] + al[7] + al[8];
  const double bg_alt = al[9] + al[10] + al[11] + al[12] + al[13] + al[14];
  const double bps_alt = 2.0 * bperp * bp_alt;
  const double imask = in_range(sd.ri, i, j) ? 1.0 : 0.0;
  const double d3a =
      s_perp * s_perp * sg * ((-1.0 * imask * R_gd * gpara + (R_gd * (p_alt + bps_alt / (8.0 * PI)) + bh * (bg_alt) / (4.0 * PI)) / r));
  const double d5a = c_plus * al_p * (imask * gperp + bpara * bp_alt / (4.0 * PI * r)) +
                     c_minus * al_m * s_perp * (-imask * R_para * gpara + R_para * (p_alt + bps_alt / (8.0 * PI)) / r);
  const double d7a = c_minus * al_m * (imask * gperp + bpara * bp_alt / (4.0 * PI * r)) -
                     c_plus * al_p * s_perp * (-imask * R_para * gpara + R_para * (p_alt + bps_alt / (8.0 * PI)) / r);
  const double in_dir = lower ? 1.0 : -1.0;
  if (in_dir * vp > 0.0) { d1 = 0.0; d2 = 0.0; }
  if (in_dir * (vp + c_perp) > 0.0) d3 = d3a;
  if (in_dir * (vp - c_perp) > 0.0) d4 = -d3a;
  if (in_dir * (vp + c_plus) > 0.0) d5 = d5a;
  if (in_dir * (vp - c_plus) > 0.0) d6 = -d5a;
  if (in_dir * (vp + c_minus) > 0.0) d7 = d7a;
  if (in_dir * (vp - c_minus) > 0.0) d8 = -d7a;

  // normal terms
  const double sqp = sqrt(PI * r);
  const double rho_n = -(c.K.gamma / r * d2 + 0.5 * r * al_p * (d5 + d6) + 0.5 * r * al_m * (d7 + d8)) / css;
  const double vpara_n = -0.5 * s_perp * (sg * R_gd * (-d3 + d4) + (c_minus * al_m / css * R_para * (-d5 + d6) +
                                                                      c_plus * al_p / css * R_para * (d7 - d8))) +
                         vc * op[OP_SB + 0];
  const double vgd_n = -0.5 * s_perp * (sg * R_para * (d3 - d4) + (c_minus * al_m / css * R_gd * (-d5 + d6) +
                                                                     c_plus * al_p / css * R_gd * (d7 - d8))) +
                       vc * op[OP_SB + 1];
  const double vperp_n = -0.5 / css * (c_plus * al_p * (d5 - d6) + c_minus * al_m * (d7 - d8)) +
                         vc * op[OP_SB + 2];
  const double te_n = -(0.5 * (te + p) * al_p * (d5 + d6) + 0.5 * (te + p) * al_m * (d7 + d8)) / css;
  const double bpara_n = -sqp * (sg * R_gd * (d3 + d4) + al_m / c_s * R_para * (d5 + d6) - al_p / c_s * R_para * (d7 + d8));
  const double bgd_n = -sqp * (sg * -1.0 * R_para * (d3 + d4) + al_m / c_s * R_gd * (d5 + d6) - al_p / c_s * R_gd * (d7 + d8));
  const double bperp_n = -d1;

  // parallel terms (transport along the side)
  // dI: press, be_para, bi_para, be_perp, bi_perp, be_gd, bi_gd, v_para, v_gd, v_perp
  // tI: rho, rho v_para, rho v_gd, rho v_perp, te, be_para, bi_para, be_gd, bi_gd, be_perp, bi_perp
  const double* dI = op + OP_DI;
  const double* tI = op + OP_TI;
  const double* sI = op + OP_SI;   // v_para, v_gd, v_perp
  const double vpara = fv(c, Fvpara, i, j), vgd = fv(c, Fvgd, i, j);
  const double rho_p = -tI[0];
  const double vpara_p = -1.0 / r * (dI[0] + 2.0 * bpara / (8.0 * PI) * (dI[1] + dI[2]) +
                                     2.0 * bperp / (8.0 * PI) * (dI[3] + dI[4]) +
                                     2.0 * bgd / (8.0 * PI) * (dI[5] + dI[6])) -
                         (tI[1] - vpara * tI[0]) / r + bpara / (4.0 * PI * r) * (dI[1] + dI[2]) +
                         imask * gpara + vc * sI[0];
  const double vgd_p = -(tI[2] - vgd * tI[0]) / r + bpara / (4.0 * PI * r) * (dI[5] + dI[6]) + vc * sI[1];
  const double vperp_p = -(tI[3] - vp * tI[0]) / r + bpara / (4.0 * PI * r) * (dI[3] + dI[4]) + imask * gperp +
                         vc * sI[2];
  const double te_p = -tI[4] - p * dI[7];
  const double bpara_p = -(tI[5] + tI[6]) + bpara * dI[7];
  const double bgd_p = -(tI[7] + tI[8]) + bpara * dI[8];
  const double bperp_p = -(tI[9] + tI[10]) + bpara * dI[9];

  out[0] = rho_n + rho_p;
  out[3] = vgd_n + vgd_p;
  out[4] = te_n + te_p;
  out[7] = bgd_n + bgd_p;
  if (a == 0) {
    out[1] = vperp_n + vperp_p;
    out[2] = vpara_n + vpara_p;
    out[5] = bperp_n + bperp_p;
    out[6] = bpara_n + bpara_p;
  } else {
    out[1] = vpara_n + vpara_p;
    out[2] = vperp_n + vperp_p;
    out[5] = bpara_n + bpara_p;
    out[6] = bperp_n + bperp_p;
  }
}

// The cells of the ghost rows x in {0, 1, nx-2, nx-1} (every column), then of the ghost columns
// y in {0, 1, ny-2, ny-1} outside those rows: each cell once.
__device__ __forceinline__ bool strip_cell(int t, int nx, int ny, int& i, int& j) {
  if (t < 4 * ny) {
    const int r = t / ny;
    i = r < 2 ? r : nx - 4 + r;
    j = t % ny;
    return true;
  }
  const int u = t - 4 * ny, q = u / nx;
  j = q < 2 ? q : ny - 4 + q;
  i = u % nx;
  return !(i < 2 || i >= nx - 2);
}

__device__ __forceinline__ int slot_of(const MocArgs& a, int b, int i, int j) {
  const int* r = a.side[b].rb;
  return a.slot_base[b] + (i - r[0]) * (r[3] - r[1] + 1) + (j - r[1]);
}

// one thread per (derivative, slot), derivative-major so that a wavefront evaluates one derivative
// at 64 neighbouring cells (no divergence between the derivative kinds); ops is derivative-major
// too (ops[k * slots + slot]), so a wavefront's stores are contiguous
__global__ void __launch_bounds__(256) k_moc_ops(MocArgs a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int total = a.slot_base[4];
  if (t >= (long long)total * MOC_NOPS) return;
  const int k = (int)(t / total), slot = (int)(t % total);
  int b = 0;
  while (b < 3 && slot >= a.slot_base[b + 1]) ++b;
  // the side's description by selects from constant indices (indexing the kernel argument's array
  // at run time would copy it to private memory)
  MocSide sd;
  const MocSide &s0 = a.side[0], &s1 = a.side[1], &s2 = a.side[2], &s3 = a.side[3];
  auto sel = [&](int v0, int v1, int v2, int v3) { return b == 0 ? v0 : b == 1 ? v1 : b == 2 ? v2 : v3; };
  sd.on = sel(s0.on, s1.on, s2.on, s3.on);
  sd.axis = sel(s0.axis, s1.axis, s2.axis, s3.axis);
  sd.lower = sel(s0.lower, s1.lower, s2.lower, s3.lower);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    sd.rb[q] = sel(s0.rb[q], s1.rb[q], s2.rb[q], s3.rb[q]);
    sd.ri[q] = sel(s0.ri[q], s1.ri[q], s2.ri[q], s3.ri[q]);
    sd.rp[q] = sel(s0.rp[q], s1.rp[q], s2.rp[q], s3.rp[q]);
    sd.ru[q] = sel(s0.ru[q], s1.ru[q], s2.ru[q], s3.ru[q]);
    sd.rl[q] = sel(s0.rl[q], s1.rl[q], s2.rl[q], s3.rl[q]);
  }
  const int* r = sd.rb;
  const int w = r[3] - r[1] + 1, loc = slot - a.slot_base[b];
  const int i = r[0] + loc / w, j = r[1] + loc % w;
  const Ctx c{a.S, a.st, a.K, a.M};
  a.ops[(size_t)k * total + slot] = eval_op(c, sd, k, i, j);
}

__global__ void __launch_bounds__(64) k_moc_rhs(MocArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = a.M.nx, ny = a.M.ny;
  if (t >= 4 * nx + 4 * ny) return;
  int i, j;
  if (!strip_cell(t, nx, ny, i, j)) return;
  bool any = false;
  for (int b = 0; b < 4; ++b) any = any || (a.side[b].on && in_range(a.side[b].rb, i, j));
  if (!any) return;
  const Ctx c{a.S, a.st, a.K, a.M};
  const double vc = a.vc ? *a.vc : 0.0;
  double res[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) res[w] = 0.0;
  for (int b = 0; b < 4; ++b) {
    if (!a.side[b].on || !in_range(a.side[b].rb, i, j)) continue;
    double t8[8], op[MOC_NOPS];
    const int sl = slot_of(a, b, i, j), total = a.slot_base[4];
    for (int k = 0; k < MOC_NOPS; ++k) op[k] = a.ops[(size_t)k * total + sl];
    side_terms(c, a.side[b], vc, i, j, op, t8);
#pragma unroll
    for (int w = 0; w < 8; ++w) res[w] = res[w] + t8[w];
  }
  // momentum form (idealmhd.cpp:323-329)
  const size_t cell = (size_t)i * ny + j, N = (size_t)nx * ny;
  const double rho = a.S.rho[cell];
  const double mx = rho * res[1] + fv(c, F_VX, i, j) * res[0];
  const double my = rho * res[2] + fv(c, F_VY, i, j) * res[0];
  const double mz = rho * res[3] + fv(c, F_VZ, i, j) * res[0];
  res[1] = mx;
  res[2] = my;
  res[3] = mz;
#pragma unroll
  for (int w = 0; w < 8; ++w) a.out[w * N + cell] = res[w];
}

// ---- ghost-cell limiters (idealmhd.cpp:107-224) ----
__device__ __forceinline__ double clamp_ref(double x, double fl, double ce, double ref) {
  return ref >= 0.0 ? smin(smax(x, fl), ce) : smax(smin(x, fl), ce);
}
struct ThrArgs {
  MhdSet G;
  MhdStatic st;
  int nx, ny;
  int moc[4];
  MocLimits lim;
};
__device__ __forceinline__ void thr_var(const ThrArgs& a, int v, double*& q, const double*& e) {
  q = v == 0 ? a.G.mx : v == 1 ? a.G.my : v == 2 ? a.G.mz : v == 3 ? a.G.bx : v == 4 ? a.G.by : a.G.bz;
  e = v == 3 ? a.st.bex : v == 4 ? a.st.bey : v == 5 ? a.st.bez : nullptr;
}
// one thread per (variable, column j): x_bound_1 then x_bound_2, each k = 0..2 in order
__global__ void __launch_bounds__(64) k_moc_thr_x(ThrArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * a.ny) return;
  const int v = t / a.ny, j = t % a.ny;
  if ((v < 3 && !a.lim.mom) || (v >= 3 && !a.lim.b)) return;
  double* q;
  const double* e;
  thr_var(a, v, q, e);
  const double lo = v < 3 ? a.lim.mom_lo : a.lim.b_lo, hi = v < 3 ? a.lim.mom_hi : a.lim.b_hi;
  const int ny = a.ny, nx = a.nx;
  auto fix = [&](int ci, int ri) {
    const size_t c = (size_t)ci * ny + j, rc = (size_t)ri * ny + j;
    if (e) {
      const double ref = e[rc] + q[rc];
      q[c] = clamp_ref(q[c], lo * ref - e[c], hi * ref - e[c], ref);
    } else {
      const double ref = q[rc];
      q[c] = clamp_ref(q[c], lo * ref, hi * ref, ref);
    }
  };
  if (a.moc[0]) for (int k = 0; k < 3; ++k) fix(k, 3);
  if (a.moc[1]) for (int k = 0; k < 3; ++k) fix(nx - 1 - k, nx - 4);
}
// one thread per (variable, row i): y_bound_1 then y_bound_2; the field limiter of y_bound_2 takes
// its reference value from column xdim-4, as the reference does (idealmhd.cpp:158)
__global__ void __launch_bounds__(64) k_moc_thr_y(ThrArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * a.nx) return;
  const int v = t / a.nx, i = t % a.nx;
  if ((v < 3 && !a.lim.mom) || (v >= 3 && !a.lim.b)) return;
  double* q;
  const double* e;
  thr_var(a, v, q, e);
  const double lo = v < 3 ? a.lim.mom_lo : a.lim.b_lo, hi = v < 3 ? a.lim.mom_hi : a.lim.b_hi;
  const int ny = a.ny, nx = a.nx;
  auto fix = [&](int cj, int rj) {
    const size_t c = (size_t)i * ny + cj, rc = (size_t)i * ny + rj;
    if (e) {
      const double ref = e[rc] + q[rc];
      q[c] = clamp_ref(q[c], lo * ref - e[c], hi * ref - e[c], ref);
    } else {
      const double ref = q[rc];
      q[c] = clamp_ref(q[c], lo * ref, hi * ref, ref);
    }
  };
  if (a.moc[2]) for (int k = 0; k < 3; ++k) fix(k, 3);
  if (a.moc[3]) for (int k = 0; k < 3; ++k) fix(ny - 1 - k, e ? nx - 4 : ny - 4);
}

}  // namespace

MocSide moc_side(int b, const int* bc, int nx, int ny) {
  MocSide s{};
  s.on = bc[b] == SPRUCE_BC_OPEN_MOC ? 1 : 0;
  s.axis = b / 2;
  s.lower = (b % 2) == 0 ? 1 : 0;
  const int NG = 2;
  int xi[2], yi[2];
  if (s.axis == 0) {
    xi[0] = s.lower ? 0 : nx - NG; xi[1] = s.lower ? NG - 1 : nx - 1;
    yi[0] = 0; yi[1] = ny - 1;
  } else {
    yi[0] = s.lower ? 0 : ny - NG; yi[1] = s.lower ? NG - 1 : ny - 1;
    xi[0] = 0; xi[1] = nx - 1;
  }
  int xin[2] = {xi[0], xi[1]}, yin[2] = {yi[0], yi[1]}, xpd[2] = {xi[0], xi[1]}, ypd[2] = {yi[0], yi[1]};
  int xlc[2] = {xi[0], xi[1]}, xuc[2] = {xi[0], xi[1]}, ylc[2] = {yi[0], yi[1]}, yuc[2] = {yi[0], yi[1]};
  if (s.axis == 0 && bc[2] != SPRUCE_BC_PERIODIC) {       // idealmhd.cpp:360-370
    yin[0] = yi[0] + NG; yin[1] = yi[1] - NG;
    ypd[0] = yi[0] + 1; ypd[1] = yi[1] - 1;
    if (bc[2] != SPRUCE_BC_OPEN_MOC) { yi[0] = yin[0]; ypd[0] = yin[0]; }
    if (bc[3] != SPRUCE_BC_OPEN_MOC) { yi[1] = yin[1]; ypd[1] = yin[1]; }
    ylc[0] = yi[0]; ylc[1] = ypd[0] - 1;
    yuc[0] = ypd[1] + 1; yuc[1] = yi[1];
  } else if (s.axis == 1 && bc[0] != SPRUCE_BC_PERIODIC) {  // idealmhd.cpp:371-382
    xin[0] = xi[0] + NG; xin[1] = xi[1] - NG;
    xpd[0] = xi[0] + 1; xpd[1] = xi[1] - 1;
    if (bc[0] != SPRUCE_BC_OPEN_MOC) { xi[0] = xin[0]; xpd[0] = xin[0]; }
    if (bc[1] != SPRUCE_BC_OPEN_MOC) { xi[1] = xin[1]; xpd[1] = xin[1]; }
    xlc[0] = xi[0]; xlc[1] = xpd[0] - 1;
    xuc[0] = xpd[1] + 1; xuc[1] = xi[1];
  }
  auto set = [](int* r, const int* x, const int* y) { r[0] = x[0]; r[1] = y[0]; r[2] = x[1]; r[3] = y[1]; };
  set(s.rb, xi, yi);
  set(s.ri, xin, yin);
  set(s.rp, xpd, ypd);
  set(s.ru, xuc, yuc);
  set(s.rl, xlc, ylc);
  return s;
}

int moc_ops_per_slot() { return MOC_NOPS; }

int moc_slot_bases(const MocSide* sides, int* base) {
  int tot = 0;
  for (int b = 0; b < 4; ++b) {
    base[b] = tot;
    if (sides[b].on) {
      const int* r = sides[b].rb;
      tot += (r[2] - r[0] + 1) * (r[3] - r[1] + 1);
    }
  }
  base[4] = tot;
  return tot;
}

void launch_moc_rhs(const MocArgs& a, hipStream_t s) {
  const long long nops = (long long)a.slot_base[4] * MOC_NOPS;
  hipLaunchKernelGGL(k_moc_ops, dim3((unsigned)((nops + 255) / 256)), dim3(256), 0, s, a);
  const int n = 4 * a.M.nx + 4 * a.M.ny;
  hipLaunchKernelGGL(k_moc_rhs, dim3((n + 63) / 64), dim3(64), 0, s, a);
}

void launch_moc_threshold(MhdSet G, MhdStatic st, Mesh M, const int* moc, MocLimits lim, hipStream_t s) {
  if (!lim.mom && !lim.b) return;
  ThrArgs a{G, st, M.nx, M.ny, {moc[0], moc[1], moc[2], moc[3]}, lim};
  if (moc[0] || moc[1]) hipLaunchKernelGGL(k_moc_thr_x, dim3((6 * M.ny + 63) / 64), dim3(64), 0, s, a);
  if (moc[2] || moc[3]) hipLaunchKernelGGL(k_moc_thr_y, dim3((6 * M.nx + 63) / 64), dim3(64), 0, s, a);
}

}  // namespace spruce
