// Discrete adjoint of the physical viscosity module (adj_pvisc.hip; forward pvisc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "pvisc.hpp"

namespace spruce {

// Adjoint grids of one reverse subcycle update.  lo_*: the adjoint of the update's output (thermal
// energy, momenta); lb_*: of its base, written per base_mode (0: in place, lb == lo, the thermal
// energy through its floor; 1: assigned; 2: added).  l_d (velocity derivatives in PvDerivs order),
// l_t25, l_bh (b_hat), l_v (velocities), l_T: scatter accumulators of the evaluation set, cleared by
// the kernel that consumes them.  partial: block sums of the step-size adjoint (added when
// accumulate).
struct PvLam {
  const double* lo_te;
  const double* lo_m[3];
  double* lb_te;
  double* lb_m[3];
  int base_mode;
  double* l_d[6];
  double* l_t25;
  double* l_bh[3];
  double* l_v[3];
  double* l_T;
  double* partial;
  int accumulate;
};

// One subcycle update (PV_FULL or PV_HALF, physicalviscosity.cpp:163-245) reversed at every cell:
// the floor of the thermal energy, the base, and the heating and viscous force (:47-64, :82-140) at
// the evaluation set onto l_d, l_t25, l_bh and l_v of the cell and its four neighbours (atomics).
// a.d must hold the forward velocity derivatives and T^2.5 of a.eval.
void launch_adj_pv_apply(const PvArgs& a, const PvLam& L, hipStream_t s);
// l_d onto l_v (derivative1D transposed, interior cells) and l_t25 onto l_T at the evaluation set E.
void launch_adj_pv_derivs(PvSet E, Mesh mesh, const PvLam& L, hipStream_t s);
// l_v, l_T of E onto its thermal energy and momenta (assigned or added) and the frozen density.
void launch_adj_pv_derived(PvSet E, PvFrozen f, MhdConst k, size_t N, const PvLam& L, double* out_te, double* out_mx,
                           double* out_my, double* out_mz, int assign, double* lam_rho, hipStream_t s);
// l_bh onto bi_x, bi_y, bi_z of lam (8*N, evolved order) through b_hat = (be + bi)/|be + bi| of S.
void launch_adj_pv_bhat(MhdSet S, MhdStatic st, size_t N, const PvLam& L, double* lam, hipStream_t s);

}  // namespace spruce
