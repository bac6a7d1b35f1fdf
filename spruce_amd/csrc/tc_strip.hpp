// Building blocks of the wavefront-strip form of the conduction update (k_tc_strip, solar.hip) and
// of its reverse (k_adj_tc_strip, adj_modules.hip): one 64-lane wavefront owns consecutive columns
// (y) and marches down the rows (x) of a strip; the face values the energy derivative takes are
// formed once and shared by the two cells of the face, along y through DPP lane shifts, along x
// by keeping the previous row's faces in registers.  Each value is the same sequence of IEEE
// operations as tc_core's (tc_cell.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "arith.hpp"
#include "solar.hpp"
#include "strip.hpp"

namespace spruce {
namespace tcs {

// temp = max((gamma-1) te / (2 K_B n), temp_min) (thermalconduction.cpp:109; idealmhd.cpp:254)
__device__ __forceinline__ double temp_of(double te, double n, const MhdConst& K) {
  return smax(te * K.gm1 / (n * (2.0 * K_B)), K.tmin);
}

// one row of the strip at this lane: cell values and the quantities that need only y-neighbours
struct Row {
  double T, te, gy, s1, bx, by, dbxy, dbyy;
  double coef, dcy;          // flux saturation: coefficient and its y-derivative
  int i;                     // wrapped row index (clamped into [0, nx) when the row lies outside)
  bool val, inx;             // row exists; row inside the interior bounds
};
struct Raw { double te, n, bx, by, coef; };
struct XFaces { double T, gy, bx, by, coef; };

struct Lane {
  int j;                     // wrapped column (clamped into [0, ny) when it lies outside)
  bool val, iny, out;        // column exists; inside the interior bounds; writes an output
  double hc, hn, fd, rfd;    // upper y-face: weight of this lane's value, of the next lane's, denominator
  double dy, rdy, hsq, rhsq;
};

__device__ __forceinline__ int row_index(int k, int n, int per, bool& val) {
  if (per) { k %= n; if (k < 0) k += n; val = true; return k; }
  val = k >= 0 && k < n;
  return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

// The quotient a/b: the IEEE one (EX, what the forward takes), or the product with the prepared
// reciprocal rb (within an ulp; the adjoint's recompute of the values its partial derivatives take)
template <bool EX>
__device__ __forceinline__ double qd(double a, double b, double rb) {
  return EX ? qdiv(a, b, rb) : a * rb;
}

// boundaryInterpolate across this lane's upper y-face (derivs.cpp:477-487), from its own value and
// the next lane's
template <bool EX = true>
__device__ __forceinline__ double yface(const Lane& L, double q) {
  const double up = strip::shl1(q);
  return qd<EX>(q * L.hc + up * L.hn, L.fd, L.rfd);
}
// derivative1D along y at this lane from the upper faces of this lane and of the previous one
template <bool EX = true>
__device__ __forceinline__ double yderiv(const Lane& L, double fup) {
  return qd<EX>(fup - strip::shr1(fup), L.dy, L.rdy);
}

template <bool SAT>
__device__ __forceinline__ Raw load_raw(const TcStepArgs& a, const Lane& L, int k) {
  Raw r{0.0, 0.0, 0.0, 0.0, 0.0};
  bool v;
  const int i = row_index(k, a.mesh.nx, a.mesh.x.per, v);
  if (v && L.val) {
    const size_t c = (size_t)i * a.mesh.ny + L.j;
    r.te = a.te_src[c];
    r.n = a.n[c];
    r.bx = a.bhx[c];
    r.by = a.bhy[c];
    if (SAT) r.coef = a.coef[c];
  }
  return r;
}

template <bool SAT, bool EX = true>
__device__ __forceinline__ Row make_row(const TcStepArgs& a, const Lane& L, int k, const Raw& r) {
  const Mesh& M = a.mesh;
  Row w;
  w.i = row_index(k, M.nx, M.x.per, w.val);
  w.inx = w.val && w.i >= M.x.lo && w.i <= M.x.hi;
  const bool ok = w.val && L.val;
  w.te = r.te;
  w.T = ok ? temp_of(r.te, r.n, a.k) : 0.0;
  w.bx = r.bx;
  w.by = r.by;
  const double fT = yface<EX>(L, w.T);
  const double fTlo = strip::shr1(fT);
  // dtemp_dy is zero outside the interior (derivs.cpp:229-233)
  w.gy = (w.inx && L.iny) ? qd<EX>(fT - fTlo, L.dy, L.rdy) : 0.0;
  w.s1 = qd<EX>(fT - 2.0 * w.T + fTlo, L.hsq, L.rhsq);      // secondDerivative1D (derivs.cpp:450)
  w.dbxy = yderiv<EX>(L, yface<EX>(L, w.bx));
  w.dbyy = yderiv<EX>(L, yface<EX>(L, w.by));
  if (SAT) {
    w.coef = r.coef;
    w.dcy = yderiv<EX>(L, yface<EX>(L, w.coef));
  }
  return w;
}

// boundaryInterpolate across the x-face between rows A (lower) and B (upper)
template <bool SAT, bool EX = true>
__device__ __forceinline__ XFaces xfaces(const Mesh& M, const Row& A, const Row& B) {
  const double ha = M.x.h[A.i], hb = M.x.h[B.i], fd = M.x.fden[B.i], rfd = M.x.rfden[B.i];
  XFaces f;
  f.T = qd<EX>(A.T * hb + B.T * ha, fd, rfd);
  f.gy = qd<EX>(A.gy * hb + B.gy * ha, fd, rfd);
  f.bx = qd<EX>(A.bx * hb + B.bx * ha, fd, rfd);
  f.by = qd<EX>(A.by * hb + B.by * ha, fd, rfd);
  if (SAT) f.coef = qd<EX>(A.coef * hb + B.coef * ha, fd, rfd);
  return f;
}

}  // namespace tcs
}  // namespace spruce
