// Discrete-adjoint kernels for the parts of the step outside the IdealMHD stages (adj_modules.hip):
// the ghost fill of open, fixed and reflect boundaries, the minimum/derived update of a live
// propagate, thermal conduction (any subcycle integrator, flux saturation too), radiative losses,
// artificial viscosity in the right-hand side, the source modules after the integrator, and the
// heating of the parametrised profile.  Adjoint vectors are 8*N doubles in evolved order
// (rho, mom_x, mom_y, mom_z, thermal_energy, bi_x, bi_y, bi_z), updated in place.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"
#include "solar.hpp"
#include "sources.hpp"
#include "visc.hpp"

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
// = block sums of lam * dE (the adjoint of h), added to the entries already there when accumulate.
// The conduction reverses run adj_tc_blocks() blocks (16 x 64 tiles).
int adj_tc_blocks(const Mesh& m);
// Flux saturation in a conduction reverse (thermalconduction.cpp:171-174, 255-271): the energy
// derivative is coef * R + add with add = -(d coef/dx * rx + d coef/dy * ry).  coef, rx, ry are the
// forward grids at the evaluated temperatures (launch_tc_sat); a stage reverse given them seeds R
// with coef times its seed and writes, per cell, the seed times R (lcoef), minus the seed times rx
// and ry (wx, wy: the weights of the coefficient's derivatives) and minus the seed times the
// coefficient's derivatives (lfx, lfy: the adjoint of the raw flux); launch_adj_tc_satrev then
// carries them to the temperatures, the density and b_hat.  coef == nullptr: no saturation.
struct TcSatRev {
  const double *coef = nullptr, *rx = nullptr, *ry = nullptr;
  double *lcoef = nullptr, *wx = nullptr, *wy = nullptr, *lfx = nullptr, *lfy = nullptr;
};
// An Euler conduction subcycle max(te_s + mask*h*dE(T(te_s)), te_min) reversed: lam_te (the
// output's adjoint) through the clamp in place, dE's transpose scattered onto lam_T (and the frozen
// fields' adjoint lam_fz), partial the step-size sums.
void launch_adj_tc_sub(const double* te_s, const double* n, const double* bhx, const double* bhy,
                       const double* dbxx, const double* dbxy, const double* dbyx, const double* dbyy, double h,
                       double k0, MhdConst k, Mesh mesh, double* lam_te, double* lam_T, double* lam_fz,
                       double* partial, int accumulate, hipStream_t s, TcSatRev sr = TcSatRev());
// The same reverse (Euler, no flux saturation) in wavefront strips of R rows: lam_out (a grid other
// than lam_in, the adjoint of the subcycle's output) receives the adjoint of te_s whole, dJ/dT taken
// through the temperature with it; lam_rho += dJ/drho; lam_bx, lam_by += dJ/d(b_hat_x, b_hat_y), the
// reverse of the b_hat derivatives included.  partial[0 .. adj_tc_strip_blocks) as for
// launch_adj_tc_sub.  adj_tc_strip_blocks returns 0 where the grid is too small for strips.
int adj_tc_strip_blocks(const Mesh& m, int R);
void launch_adj_tc_strip(const double* te_s, const double* n, const double* bhx, const double* bhy, double h, double k0,
                         MhdConst k, Mesh mesh, const double* lam_in, double* lam_out, double* lam_rho,
                         double* lam_bx, double* lam_by, double* partial, int accumulate, int R, hipStream_t s);
// A subcycle stage of the RK2/RK4 subcycle integrators reversed (thermalconduction.cpp:67-124):
// dE is taken at the temperatures of `eval`; with cbase given, the stage output was
// max(cbase + mask*(mult*h)*dE, te_min): lam_out goes through the clamp into lam_base (assigned
// when assign_base, else added) and seeds dE with (mult*h)*mu; add_coef*seed_add[c] (seed_add may
// be null) adds the weight of a later combination.  lam_T += dJ/dT at eval (atomics), lam_fz +=
// frozen-field adjoints, partial[block] = block sums of mu*mult*dE (the adjoint of h).
void launch_adj_tc_stage(const double* eval, const double* n, const double* bhx, const double* bhy,
                         const double* dbxx, const double* dbxy, const double* dbyx, const double* dbyy, double h,
                         double mult, double k0, MhdConst k, Mesh mesh, const double* cbase, const double* lam_out,
                         double* lam_base, int assign_base, const double* seed_add, double add_coef, double* lam_T,
                         double* lam_fz, double* partial, int accumulate, hipStream_t s, TcSatRev sr = TcSatRev());
// The RK4 subcycle's final combination max(te + (h*(k1 + 2k2 + 2k3 + k4))/6, te_min), k4 at the
// temperatures of i3, reversed: lam_base = mu, lamS = mu*h/6 (k_j's seed is its weight times
// lamS), k4's transpose applied, partial = block sums of mu*S/6.
void launch_adj_tc_final4(const double* te, const double* k1, const double* k2, const double* k3, const double* i3,
                          const double* n, const double* bhx, const double* bhy, const double* dbxx,
                          const double* dbxy, const double* dbyx, const double* dbyy, double h, double k0, MhdConst k,
                          Mesh mesh, const double* lam_out, double* lam_base, double* lamS, double* lam_T,
                          double* lam_fz, double* partial, int accumulate, hipStream_t s, TcSatRev sr = TcSatRev());
// The saturation outputs of a stage reverse (TcSatRev) carried back: the coefficient's adjoint
// (its own seed plus the transposed derivative1D of the neighbours' weights) through
// coef = sat / sqrt(sat^2 + |F|^2), sat = (rho/m_i) (1/6)(3/2) (K_B T)^1.5 / sqrt(m_e), and the raw flux
// F = -min(k0 T^2.5, kappa_max) (b_hat . grad T) b_hat (thermalconduction.cpp:196-227) onto lam_T
// (atomics; grad T transposed onto the four neighbours), lam_fz (b_hat_x, b_hat_y) and lam_rho,
// at the temperatures of `eval`.
void launch_adj_tc_satrev(const double* eval, const double* n, const double* rho, const double* bhx,
                          const double* bhy, TcParams p, MhdConst k, Mesh mesh, TcSatRev sr, double* lam_T,
                          double* lam_fz, double* lam_rho, hipStream_t s);
// T = max((gamma-1) te / (2 K_B n), T_min) reversed: lam_te += lam_T dT/dte, lam_rho += lam_T dT/dn / m
// (n = rho / m), and lam_T is cleared for the next subcycle.
void launch_adj_tc_temp(const double* te_s, const double* n, MhdConst k, size_t N, double* lam_T, double* lam_te,
                        double* lam_rho, hipStream_t s);
// The b_hat derivatives (derivative1D, zero off the interior) and b_hat = b/|b| (idealmhd.cpp:259-277)
// reversed onto bi_x, bi_y, bi_z of the state L the conduction step froze them from.
void launch_adj_bhat(MhdSet L, MhdStatic st, Mesh mesh, double* lam_fz, double* lam, hipStream_t s);
// One divergence-cleaning subcycle (divcleaning.cpp:28-46) reversed, given the subcycle's total
// fields bx_s, by_s and their inner derivatives dxb = d(bx)/dx, dyb = d(by)/dy (zero off the
// interior; launch_dc_inner): lbx, lby (the adjoints of bi_x, bi_y) updated in place; scratch: 4*N
// doubles, zero on entry and left zero; partial: block sums of the adjoint of the subcycle size h.
void launch_adj_divclean(const double* bx_s, const double* by_s, const double* dxb, const double* dyb,
                         const double* coeff, double h, Mesh mesh, double* lbx, double* lby, double* scratch,
                         double* partial, hipStream_t s);
// BoundaryOutflow after the integrator (boundaryoutflow.cpp:39-67) reversed on its input state P
// (the propagate's floors and ghost fill reversed beforehand): the acceleration a = sc[1] of the
// forward, its template, and with dynamic_mode the forward's window maxima pv, pk (nb blocks,
// launch_outflow_max on P) for the cell that sets the throttle.  lam updated in place; pa: block
// sums of the acceleration's adjoint (scratch, adj_mod_blocks), ph: of the step size's.
void launch_adj_outflow(MhdSet P, MhdStatic st, const double* tmpl, const double* sc, const double* step,
                        OutflowParams p, Mesh mesh, const double* pv, const double* pk, int nb, double* lam,
                        double* pa, double* ph, hipStream_t s);
// FieldHeating's update after its heating (fieldheating.cpp:34-61) reversed: the heating is taken
// from the step's start state Ln (its b_hat bhx, bhy), the update te += mask*(h*H) from the module's
// input P; lam (the propagated output's adjoint) through the floors in place, the heating's adjoint
// onto lamLn (Ln's adjoint, atomics) and onto lfz (b_hat and its four derivatives, for
// launch_adj_bhat on Ln), partial the block sums of the step-size adjoint.
void launch_adj_fheat(MhdSet Ln, MhdSet P, MhdStatic st, const double* bhx, const double* bhy, FheatParams p,
                      const double* step, MhdConst k, Mesh mesh, double* lam, double* lamLn, double* lfz,
                      double* partial, hipStream_t s);
// Artificial viscosity as a right-hand-side term (strength <= 1; viscosity.cpp:112-121, 185-267),
// reversed at the stage state a.S for the stage derivative's adjoint lamk = sa*ka + kb (kb may be
// null), its component `evol`: the Laplacian's stencil to lq (atomics; cleared by
// launch_adj_visc_vars), the scale n*c to lam_S's density, and the coefficient's timescale: a local
// dt grid's adjoint to ldt (per cell, added), a global minimum's to partial (one block sum each;
// null for a local term).  With gradient_correction the derivatives of q go to lq too and those of
// f = nu*sc to lf (atomics; then launch_adj_visc_f).
void launch_adj_visc_term(const ViscTermArgs& a, int evol, const double* ka, double sa, const double* kb, double* lq,
                          double* lf, double* lam_S, double* ldt, double* partial, hipStream_t s);
// The gradient correction's adjoint lf of f = nu*sc on every cell taken to the density (the scale)
// and to the timescale (ldt, or block sums in partial for a global term); lf cleared.
void launch_adj_visc_f(const ViscTermArgs& a, double* lf, double* lam_S, double* ldt, double* partial, hipStream_t s);
// The Laplacian's adjoint lq of the differentiated variable a.var taken to the evolved grids of a.S
// (idealmhd.cpp:241-277 reversed) on every cell, into lam_S; lq cleared.
void launch_adj_visc_vars(const ViscTermArgs& a, double* lq, double* lam_S, hipStream_t s);
// One update of a hyperviscous subcycle, g = init + (mask*hh)*D followed by a live propagate
// (viscosity.cpp:133-173), reversed: P is the state before the update (its other grids and the
// floors' inputs), g the evolved index updated, hh = c*h.  lam (adjoint of the propagated state)
// goes through the density and energy floors; its g component moves to lam_init (added) and is
// cleared; mask*hh times it is added, weighted by w[t], to each of the nt derivative adjoints ld[t];
// partial (one per block; accumulated when `accumulate`) gets the sum of c * mask * lam_g * D.
void launch_adj_hv_update(MhdSet P, int g, const double* init, const double* D, double c, double hh, MhdConst k,
                          Mesh mesh, double* lam, double* lam_init, int nt, double* const ld[4], const double w[4],
                          double* partial, int accumulate, hipStream_t s);
// The adjoint of the per-cell dt grid of L (idealmhd.cpp:279-304) added to lam_L on the interior;
// ldt cleared.
void launch_adj_dt_grid(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double* ldt, double* lam_L, hipStream_t s);
// Reverse of RadiativeLosses::iterateModule on the input state A (radiativelosses.cpp:84-140,
// with prevent_subcycling through the live dt of A, st its static fields) fused with the live
// propagate's minimum/derived reverse: lam updated in place, partial[block] = block sums of dJ/dh.
void launch_adj_rl(MhdSet A, MhdStatic st, RlParams p, MhdConst k, Mesh mesh, int nsub, double h, int integ,
                   double* lam, double* partial, hipStream_t s);
// Reverse of a source module's update after the integrator and of the propagate that follows it,
// from the state P before it (modes of launch_source, sources.hpp: ADD te, ADD_SCALED rho by m_i,
// ADD_WEIGHTED both momenta by rho with templates t, t2, SUB te by (*step) * t): the added term is
// mask * ((h * g) * t) ... with h = *step and g a factor of the time at the step (dg/dtime = gp).
// lam updated in place; ph[block] / pt[block]: block sums of dJ/dh and dJ/dtime.
void launch_adj_source(MhdSet P, int mode, const double* t, const double* t2, const double* step, double g, double gp,
                       MhdConst k, Mesh mesh, double* lam, double* ph, double* pt, hipStream_t s);
// dJ/dtheta_k += (*step) * sum over interior cells of lam_te * phi_k(y) for the parametrised heating
// profile (basis: K x ny values of phi_k at row j); colsum is ny doubles of scratch.
void launch_adj_heat_params(const double* lam_te, const double* basis, int K, const double* step, Mesh mesh,
                            double* colsum, double* g_theta, hipStream_t s);
// gh[0] += scale * sum of partial[0..n) (one block, fixed order).
void launch_adj_accum(const double* partial, int n, double scale, double* gh, hipStream_t s);

}  // namespace spruce
