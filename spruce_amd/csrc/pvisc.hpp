// Physical (Braginskii) viscosity (source/modules/solar/physicalviscosity.cpp) on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

// Subcycle state: thermal energy, temperature, momenta and velocities.
struct PvSet {
  double *te, *T, *mx, *my, *mz, *vx, *vy, *vz;
};

// Velocity derivatives (zero outside the interior) and T^2.5 of an evaluation set.
struct PvDerivs {
  double *xvx, *xvy, *xvz, *yvx, *yvy, *yvz, *t25;
};

struct PvFrozen {              // fixed during a module iteration
  const double *bhx, *bhy, *bhz, *coef, *n, *rho;
};

enum PvMode { PV_FULL = 0, PV_HALF = 1 };

struct PvArgs {
  PvSet eval, base, out;
  PvDerivs d;
  PvFrozen f;
  MhdConst k;
  Mesh mesh;
  double h;                    // subcycle step
  int mode;                    // PV_FULL: out = base + h*rate; PV_HALF: out = base + 0.5*h*rate
  int heating_on, force_on, apply, coeff_zero, grad_corr;
  // file output and multispecies accumulators, added to in PV_FULL evaluations only; null when off
  // (physicalviscosity.cpp:164-178, :218-230)
  double* avg_heat;            // += heating / nsd
  double* avg_f[3];            // += force / nsd
  double nsd;                  // the subcycle count
  double* ms_ion;              // += ((1 - msf) * heating) * h when heating is applied, msf < 1
  double* ms_el;               // += (msf * heating) * h when heating is applied, msf > 0
  double msf;
};

// b_hat, velocities, temperature and the copies of te and momenta that start the iteration
// (physicalviscosity.cpp:156-161), from the live set.
void launch_pv_init(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double* bhx, double* bhy, double* bhz, PvSet out,
                    hipStream_t s);
// per-block minima of the viscous timescale over the interior (:67-80)
int pv_blocks(const Mesh& m);
void launch_pv_timescale(PvSet cur, PvFrozen f, double coeff_floor, Mesh mesh, double* partial, hipStream_t s);
void launch_pv_derivs(PvSet e, Mesh mesh, PvDerivs d, hipStream_t s);
void launch_pv_apply(const PvArgs& a, hipStream_t s);

}  // namespace spruce
