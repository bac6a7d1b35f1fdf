// Discrete-adjoint kernels for the parts of the step outside the IdealMHD stages (adj_modules.hip):
// the ghost fill of open, fixed and reflect boundaries, the minimum/derived update of a live
// propagate, thermal conduction (Euler subcycles, no flux saturation), radiative losses, and the
// heating of the parametrised profile.  Adjoint vectors are 8*N doubles in evolved order
// (rho, mom_x, mom_y, mom_z, thermal_energy, bi_x, bi_y, bi_z), updated in place.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"
#include "solar.hpp"

namespace spruce {

int adj_mod_blocks(const Mesh& m);
// Reverse of the ghost fill (evolution.cpp:126-334, GHOST_NORMAL, or GHOST_LIVE_RAW with raw = 1)
// on the two sides of one axis (0: x, 1: y), given the filled state P (raw: its interior).  The
// ghost values a side writes lose their adjoint to the interior cells they were computed from.
// Sides with an open boundary need the other axis periodic.
void launch_adj_ghost_axis(MhdSet P, MhdConst k, Mesh mesh, const int kinds[4], const double s1[4],
                           const double s2[4], int axis, double* lam, hipStream_t s, int raw = 0);
// Reverse of enforceMinimums + recomputeDerivedVarsFromEvolvedVars (idealmhd.cpp:234-252) applied
// to rho_in and te_in = te_base + (*step) * add (add may be null: te_in = te_base): the adjoint of
// rho and thermal_energy is kept where no floor is active.  With add given, partial[block] receives
// the block sums of lam_te * add after the mask (the step-size adjoint of te += step * add).
void launch_adj_minmask(const double* rho_in, const double* te_base, const double* add, const double* step,
                        MhdConst k, Mesh mesh, double* lam, double* partial, hipStream_t s);
// Reverse of one Euler conduction subcycle te_out = max(te_s + h dE(te_s), te_min) on the frozen
// n, b_hat and b_hat derivatives (thermalconduction.cpp:94-97, 158-176): lam_te (in: dJ/dte_out,
// out: the direct part of dJ/dte_s), lam_T += dJ/dT on the 3x3 neighbourhoods (atomics), lam_fz
// (6*N: b_hat_x, b_hat_y, d b_hat_x/dx, /dy, d b_hat_y/dx, /dy) += their adjoints, partial[block]
// = block sums of lam * dE (the adjoint of h).
void launch_adj_tc_sub(const double* te_s, const double* n, const double* bhx, const double* bhy,
                       const double* dbxx, const double* dbxy, const double* dbyx, const double* dbyy, double h,
                       double k0, MhdConst k, Mesh mesh, double* lam_te, double* lam_T, double* lam_fz,
                       double* partial, hipStream_t s);
// T = max((gamma-1) te / (2 K_B n), T_min) reversed: lam_te += lam_T dT/dte, lam_rho += lam_T dT/dn / m
// (n = rho / m), and lam_T is cleared for the next subcycle.
void launch_adj_tc_temp(const double* te_s, const double* n, MhdConst k, size_t N, double* lam_T, double* lam_te,
                        double* lam_rho, hipStream_t s);
// The b_hat derivatives (derivative1D, zero off the interior) and b_hat = b/|b| (idealmhd.cpp:259-277)
// reversed onto bi_x, bi_y, bi_z of the state L the conduction step froze them from.
void launch_adj_bhat(MhdSet L, MhdStatic st, Mesh mesh, double* lam_fz, double* lam, hipStream_t s);
// Reverse of RadiativeLosses::iterateModule on the input state A (radiativelosses.cpp:84-140,
// without prevent_subcycling) fused with the live propagate's minimum/derived reverse: lam updated
// in place, partial[block] = block sums of dJ/dh.
void launch_adj_rl(MhdSet A, RlParams p, MhdConst k, Mesh mesh, int nsub, double h, int integ, double* lam,
                   double* partial, hipStream_t s);
// dJ/dtheta_k += (*step) * sum over interior cells of lam_te * phi_k(y) for the parametrised heating
// profile (basis: K x ny values of phi_k at row j); colsum is ny doubles of scratch.
void launch_adj_heat_params(const double* lam_te, const double* basis, int K, const double* step, Mesh mesh,
                            double* colsum, double* g_theta, hipStream_t s);
// gh[0] += scale * sum of partial[0..n) (one block, fixed order).
void launch_adj_accum(const double* partial, int n, double scale, double* gh, hipStream_t s);

}  // namespace spruce
