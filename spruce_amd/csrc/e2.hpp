// IdealMHD2E (source/equationsets/idealmhd2E.cpp) and the set-up part of IdealMHDCons
// (source/equationsets/idealmhdcons.cpp) on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

// One IdealMHD2E grid set: the 7 evolved variables plus n (not recoverable from the re-rounded rho)
// and rho_pre (post-minimum, pre-derived density read by open boundaries).
struct E2Set {
  double *rho, *n, *mx, *my, *tei, *tee, *bx, *by, *rho_pre;
};

struct E2Static {
  const double *bex, *bey, *gx, *gy;
};

struct E2StageArgs {
  E2Set S, B, T;
  E2Static st;
  MhdConst k;
  Mesh mesh;
  const double* step;
  double step_mult;
  double* kbuf;            // 7 * N
  const double* k2buf;
  const double* k3buf;
  int mode;                // StageMode
  int write_pre;
  // eic_thermalization modules (eic_thermalization.cpp:27-43): how many, and (2 m_e / m_ion)
  int eic;
  double eic_nu;
};

void launch_e2_stage(const E2StageArgs& a, hipStream_t s);
void launch_e2_populate(E2Set L, const double* ti, const double* te, MhdConst k, int nx, int ny, hipStream_t s);
void launch_e2_ghost(E2Set G, MhdConst k, Mesh mesh, const int* kinds, const double* s1, const double* s2,
                     hipStream_t s);
void launch_e2_dt(E2Set L, E2Static st, MhdConst k, Mesh mesh, double* partial, hipStream_t s);
// `which` indexes E2_DERIVED of e2_host.hip
void launch_e2_derived(E2Set S, E2Static st, MhdConst k, Mesh mesh, int which, double* out, hipStream_t s);
void launch_e2_rhs(E2Set S, E2Static st, MhdConst k, Mesh mesh, double* out, hipStream_t s, int eic = 0,
                   double eic_nu = 0.0);

// IdealMHDCons set-up: recomputeEvolvedVarsFromStateVars + enforceMinimums (idealmhdcons.cpp:89-109),
// then after the ghost fill recomputeDerivedVarsFromEvolvedVars (:111-128).  Grids in the order
// of ConsGrids.
struct ConsGrids {
  double *rho, *temp, *mx, *my, *bx, *by, *n, *vx, *vy, *ke, *b_x, *b_y, *bmag, *mag, *bhx, *bhy, *te, *press,
      *energy, *dt, *rho_pre;
  const double *bex, *bey;
};
void launch_cons_populate(ConsGrids g, MhdConst k, int nx, int ny, hipStream_t s);
void launch_cons_ghost(ConsGrids g, MhdConst k, Mesh mesh, const int* kinds, const double* s1, const double* s2,
                       hipStream_t s);
void launch_cons_derived(ConsGrids g, MhdConst k, Mesh mesh, hipStream_t s);

}  // namespace spruce
