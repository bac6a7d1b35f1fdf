// Artificial viscosity (source/modules/viscosity.cpp) on the IdealMHD grid sets.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

// Variables a viscosity term can differentiate, in IdealMHD naming (idealmhd.hpp:19-26).
enum ViscVar : int {
  VV_RHO = 0, VV_N, VV_MX, VV_MY, VV_MZ, VV_TE, VV_BIX, VV_BIY, VV_BIZ, VV_GX, VV_GY, VV_TEMP, VV_PRESS,
  VV_VX, VV_VY, VV_VZ, VV_KE, VV_BX, VV_BY, VV_BZ, VV_BMAG, VV_BHX, VV_BHY, VV_BHZ, VV_COUNT
};

struct ViscTermArgs {
  MhdSet S;                 // grids the term is evaluated on
  MhdStatic st;
  MhdConst k;
  Mesh mesh;
  const double* strength;   // per-cell strength (boundary options), or null for a constant
  double strength_c;        // constant strength (local and global options)
  const double* dtgrid;     // live dt grid (local and boundary options), or null
  const double* dtmin;      // device scalar: min of the live dt grid over the dt bounds (global options)
  int var;                  // ViscVar differentiated
  int scale;                // 0: 1, 1: n * m_i (momentum from velocity), 2: n * (K_B/(gamma-1)) (energy from temperature)
  double scale_c;           // m_i or K_B/(gamma-1)
  int grad;                 // gradient_correction
  int masked;               // multiply by the ghost-zone mask
  double* out;
  // file output (viscosity.cpp:351-376), null when off: the Laplacian, strength and timescale the
  // evaluation used, on every cell
  double* lap_out;
  double* str_out;
  double* dt_out;
};

// dqdt of one term on every cell (constructSingleViscosityGrid, viscosity.cpp:185-267).
void launch_visc_term(const ViscTermArgs& a, hipStream_t s);
// target = init + (mask * h) * d on every cell (viscosity.cpp:133-171).
void launch_visc_update(double* target, const double* init, const double* d, double h, Mesh mesh, hipStream_t s);
// out = (d1 + 2 d2 + 2 d3 + d4) / 6 (viscosity.cpp:166).
void launch_visc_rk4(const double* d1, const double* d2, const double* d3, const double* d4, double* out, size_t n,
                     hipStream_t s);
// out += add (EquationSet::computeTimeDerivatives adding a module term, equationset.cpp:204-210).
void launch_add(double* out, const double* add, size_t n, hipStream_t s);
// y += (*step) * x with the step size read on the device (AmbientHeating::postIterateModule).
void launch_add_step_times(double* y, const double* x, const double* step, size_t n, hipStream_t s);

}  // namespace spruce
