// Open characteristic (open_moc) boundaries of ideal MHD on the GPU
// (source/equationsets/idealmhd.cpp:107-224, 308-615; source/mhd/derivs.cpp:77-117, 167-214,
// 317-404; plasmadomain.cpp:155-159).
//
// The ghost cells of an open_moc side are not filled from the interior: they evolve by the
// characteristic equations, whose time derivatives are added to the (masked) ideal-MHD RHS of those
// cells.  moc_rhs writes them for every cell of every open_moc side's region (its two ghost rows or
// columns, and the ghost corners it shares with another open_moc side); the stage kernels add them
// to the ghost cells' zero RHS.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

// One open_moc side: which axis it closes, whether it is the lower side, and the index ranges the
// reference evaluates its terms on (x from, y from, x to, y to; inclusive, empty when from > to).
struct MocSide {
  int on, axis, lower;
  int rb[4];   // region: the side's ghost cells                     (x_idx, y_idx)
  int ri[4];   // parallel terms                                     (x_idx_interior, y_idx_interior)
  int rp[4];   // central derivative along the side                  (x_idx_padded, y_idx_padded)
  int ru[4];   // backward derivative at the upper corner            (x/y_idx_upper_corner)
  int rl[4];   // backward derivative at the lower corner            (x/y_idx_lower_corner)
};
struct MocArgs {
  MhdSet S;
  MhdStatic st;
  MhdConst K;
  Mesh M;
  MocSide side[4];        // x_bound_1, x_bound_2, y_bound_1, y_bound_2
  const double* vc;       // global viscosity coefficient on the device (null: 0)
  double* out;            // 8 * nx*ny, evolved order; only region cells are written
  // one slot per (open_moc side, cell of its region), sides in order: side b's slots start at
  // slot_base[b], slot_base[4] is the total; ops holds moc_ops_per_slot() doubles per slot
  int slot_base[5];
  double* ops;
};
int moc_ops_per_slot();
// the slot bases of sides[4] (an off side has no slots); returns the total
int moc_slot_bases(const MocSide* sides, int* base);

// The ranges of singleBoundaryTermsMOC (idealmhd.cpp:338-382) for side b (0..3) of an nx x ny grid
// with boundary kinds bc[4].
MocSide moc_side(int b, const int* bc, int nx, int ny);
// computeTimeDerivativesCharacteristicBoundary (idealmhd.cpp:308-332) on the region cells: every
// derivative the terms take, one thread each, then the terms of each cell.
void launch_moc_rhs(const MocArgs& a, hipStream_t s);
// the ghost-cell limiters of recomputeDerivedVarsFromEvolvedVars (idealmhd.cpp:107-224, 244-245):
// momenta then fields, x sides before y sides, on set G
struct MocLimits {
  int mom, b;
  double mom_lo, mom_hi, b_lo, b_hi;
};
void launch_moc_threshold(MhdSet G, MhdStatic st, Mesh M, const int* moc, MocLimits lim, hipStream_t s);

}  // namespace spruce
