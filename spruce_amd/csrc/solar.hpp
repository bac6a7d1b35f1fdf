// Thermal conduction and radiative losses on the GPU (source/modules/solar/thermalconduction.cpp,
// radiativelosses.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

struct TcParams {
  double k0;          // weakening_factor * KAPPA_0
  double c1;          // K_B / (weakening_factor * KAPPA_0)
  double dsm;         // dt_subcycle_min
  double c_sat;       // (1/6) * (3/2)
  double sqrt_me;     // sqrt(M_ELECTRON)
  int saturation;
};

struct RlParams {
  double cut, ramp, cut_plus_ramp, eps, pd_eps;
  int prevent;
  // the loss curve's branch bounds (radiativelosses.cpp:158-180 compares log10(T) with 4.97, 5.67,
  // 6.18, 6.55, 6.90, 7.63): tb[k] = the largest T whose log10 under the host C library is at most
  // the k-th of them, so that T <= tb[k] selects the branch the reference's log10 selects
  double tb[6];
};

enum TcMode : int {
  TC_UPDATE = 0,      // te_out = max(te_base + (mask*mult*h) * dE(T(te_src)), te_min)
  TC_STORE_K = 1,     // as UPDATE, and k_out = dE
  TC_RK4_FINAL = 2    // te_out = max(te_base + (mask*h) * (k1 + 2 k2 + 2 k3 + k4) / 6, te_min)
};

struct TcStepArgs {
  const double* te_src;   // thermal energy the temperature stencil is built from
  const double* te_base;  // thermal energy the update is applied to
  const double* n;        // number density (frozen)
  const double* bhx;      // b_hat (frozen)
  const double* bhy;
  const double* rx;       // saturation: raw field-aligned flux and coefficient of te_src
  const double* ry;
  const double* coef;
  double* te_out;
  double* k_out;
  const double* k1;
  const double* k2;
  const double* k3;
  const double* dbxx;     // derivatives of the frozen b_hat (d b_hat_x/dx, d b_hat_x/dy, d b_hat_y/dx,
  const double* dbxy;     // d b_hat_y/dy), computed once per step
  const double* dbyx;
  const double* dbyy;
  double h, mult;
  int mode;
  int rows;               // rows per wavefront strip (tc_strip form); 0: one thread per cell
  TcParams p;
  MhdConst k;
  Mesh mesh;
};

// b_hat of the live grids (idealmhd.cpp:259-277).
void launch_bhat(MhdSet L, MhdStatic st, Mesh mesh, double* bhx, double* bhy, hipStream_t s);
// derivative1D of the frozen b_hat along both axes (0 outside the interior): the conduction
// subcycles of a step all use the same values (thermalconduction.cpp:158-176).
void launch_bhat_derivs(const double* bhx, const double* bhy, Mesh mesh, double* dbxx, double* dbxy, double* dbyx,
                        double* dbyy, hipStream_t s);
// TC subcycle count inputs: min over the interior of the conduction time scale and, with saturation,
// max |b_hat . grad T| (thermalconduction.cpp:179-193).  Writes out[0] = min, out[1] = max.
void launch_tc_pre(MhdSet L, const double* bhx, const double* bhy, TcParams p, MhdConst k, Mesh mesh,
                   double* partial, int nblocks, double* out, hipStream_t s);
void launch_tc_sat(const double* te_src, const double* n, const double* rho, const double* bhx, const double* bhy,
                   TcParams p, MhdConst k, Mesh mesh, double* rx, double* ry, double* coef, hipStream_t s);
// One conduction update: the wavefront-strip form with a.rows rows per strip, or one thread per cell
// when a.rows is 0 (the b_hat derivative grids are then read, else they are formed in the kernel).
void launch_tc_step(const TcStepArgs& a, hipStream_t s);
// rows per strip giving at least 4096 wavefronts where the grid allows
int tc_strip_rows_default(int nx, int ny);
// 0 (strips) when 8-row strips give at least 4096 wavefronts, else 1 (one thread per cell)
int tc_kernel_default(int nx, int ny);
// RL subcycle count inputs: out[0] = max losses, out[1] = min |te / losses| (radiativelosses.cpp:200-205).
void launch_rl_pre(MhdSet L, const double* dt_grid, RlParams p, MhdConst k, Mesh mesh, double* partial, int nblocks,
                   double* out, hipStream_t s);
// All RL subcycles of one step in one launch (the losses are pointwise): integrator 0/1/2.
void launch_rl_iterate(MhdSet L, const double* dt_grid, RlParams p, MhdConst k, Mesh mesh, int nsub, double h,
                       int integrator, double* te_out, hipStream_t s);
// propagateChanges on the live grids up to the ghost fill: minimums then derived (pointwise).
void launch_propagate_live(MhdSet L, MhdConst k, int nx, int ny, int write_pre, hipStream_t s);
int solar_blocks(const Mesh& m);

}  // namespace spruce
