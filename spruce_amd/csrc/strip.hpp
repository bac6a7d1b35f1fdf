// Building blocks of the wavefront-strip stencil kernels (mhd2.hip, the ideal_2F strip kernel in
// twofluid.hip): one 64-lane wavefront owns 64 columns (y, contiguous in memory) and marches down
// the rows (x) of a strip; y-neighbours come from neighbouring lanes through DPP shifts, x-neighbours
// from a rotating register window.  Face arithmetic follows derivs.cpp:10-73 and 477-499 operation
// for operation.
#pragma once
#include <hip/hip_runtime.h>

#include "arith.hpp"
#include "mhd.hpp"

namespace spruce {
namespace strip {

constexpr int OUT_COLS = 60;   // lanes 2..61 produce outputs; 0, 1, 62, 63 supply the y-halo

// lane l receives the value of lane l-1 (wave_shr:1) / lane l+1 (wave_shl:1)
__device__ __forceinline__ double shr1(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double shl1(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x130, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int wrapn(int k, int n) {
  k %= n;
  return k < 0 ? k + n : k;
}
__device__ __forceinline__ int clampn(int k, int n) { return k < 0 ? 0 : (k >= n ? n - 1 : k); }

// Geometry of one face: interpolation weights and denominators, and the extrapolation geometry on
// both sides (derivs.cpp:477-499).
struct FaceGeom {
  double hlo, hup, den, rden;
  double gup, denm, rdenm;   // extrapolation through (c-2, c-1)
  double gdn, denp, rdenp;   // extrapolation through (c+1, c)
};

// the face between cells c-1 and c of an axis (wrapped on periodic axes, clamped otherwise)
__device__ __forceinline__ FaceGeom face_geom(const Axis& ax, int c) {
  const int n = ax.n;
  int cm1 = c - 1, cp1 = c + 1, cc = c;
  if (ax.per) { cc = wrapn(c, n); cm1 = wrapn(c - 1, n); cp1 = wrapn(c + 1, n); }
  else { cc = clampn(c, n); cm1 = clampn(c - 1, n); cp1 = clampn(c + 1, n); }
  FaceGeom g;
  g.hlo = ax.h[cm1];
  g.hup = ax.h[cc];
  g.den = ax.fden[cc];
  g.rden = ax.rfden[cc];
  g.gup = ax.gup[cc];
  g.denm = ax.fden[cm1];
  g.rdenm = ax.rfden[cm1];
  g.gdn = ax.gdn[cc];
  g.denp = ax.fden[cp1];
  g.rdenp = ax.rfden[cp1];
  return g;
}

// boundaryInterpolate(a at the lower cell, b at the upper cell) (derivs.cpp:477-487)
__device__ __forceinline__ double interp(const FaceGeom& g, double a, double b) {
  return qdiv(a * g.hup + b * g.hlo, g.den, g.rden);
}

// Upwind side of a face from its velocity (derivs.cpp:47-68).
struct FaceDir {
  double vf, ge, de, rde;
  bool pos, moving;
};
__device__ __forceinline__ FaceDir face_dir(double vf, const FaceGeom& g) {
  FaceDir d;
  d.vf = vf;
  d.pos = vf > 0.0;
  d.moving = d.pos || vf < 0.0;
  d.ge = d.pos ? g.gup : g.gdn;
  d.de = d.pos ? g.denm : g.denp;
  d.rde = d.pos ? g.rdenm : g.rdenp;
  return d;
}

// surf * face velocity * (Barton-limited upwind value) of one transported quantity, from the four
// cells c-2, c-1 (lo), c (up), c+1 along the face's axis; d2 = the plain interpolation.
__device__ __forceinline__ double face_flux(double q0, double q1, double q2, double q3, const FaceGeom& g,
                                            const FaceDir& fd, double& d2) {
  d2 = interp(g, q1, q2);
  const double ea = fd.pos ? q0 : q3, eb = fd.pos ? q1 : q2;
  const double d1 = ea + qdiv((eb - ea) * fd.ge, fd.de, fd.rde);
  const bool le = q2 <= q1;
  const bool outer_min = fd.pos ? le : !le;
  const double inner = outer_min ? smax(d1, d2) : smin(d1, d2);
  const double lim = outer_min ? smin(eb, inner) : smax(eb, inner);
  return (fd.moving ? lim : d2) * fd.vf;
}

// Workgroups go to the 8 XCDs round-robin by launch index (index % 8), and each XCD has its own
// L2: launch index L (of `total`) is mapped to the logical tile t so that XCD x takes the x-th
// eighth of the tiles in order, and tiles that run together on one XCD are neighbours.
__device__ __forceinline__ int xcd_tile(int L, int total) {
  const int xcd = L & 7, k = L >> 3;
  const int q = total >> 3, rem = total & 7;
  return xcd * q + (xcd < rem ? xcd : rem) + k;
}

}  // namespace strip
}  // namespace spruce
