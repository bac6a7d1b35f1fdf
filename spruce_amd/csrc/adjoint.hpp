// Discrete adjoint of the IdealMHD time step on the GPU: kernels (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

constexpr int ADJ_COEFS = 17;

int adj_blocks(const Mesh& m);
// Reverse of one stage T = minimums/derived(B + mult*h*kv) (equationset.cpp:222-230,
// idealmhd.cpp:234-252) fused with the per-cell part of the transposed time derivative at S:
// lam_B (+)= mu with mu = dT/dpre^T lam_T; lamk = mult*h*mu stays in registers and gives the cell's
// face coefficients (cx, cy, 17*N each) and its own part, lam_S (+)= own; partial (one per block)
// holds the block sums of mult * mu . kv (the adjoint of the step size h).  assign_*: first writer
// of that vector in the step, so no clear is needed.
void launch_adj_stage_cells(MhdSet B, const double* kv, const double* step, double mult, MhdSet S, MhdStatic st,
                            MhdConst k, Mesh mesh, const double* lam_T, double* lam_B, int assign_B, double* lam_S,
                            int assign_S, double* cx, double* cy, double* partial, hipStream_t s);
// The two halves of launch_adj_stage_cells, for stages whose derivative also feeds a later
// combination (RK4): the pointwise reverse (lam_B (+)= mu, lamk = mult*h*mu, partial of mult*mu.kv)
// and the cell part of the transposed time derivative at S applied to sa*ka + kb (kb may be null).
void launch_adj_update(MhdSet B, const double* kv, const double* step, double mult, MhdConst k, Mesh mesh,
                       const double* lam_T, double* lam_B, int assign_B, double* lamk, double* partial,
                       hipStream_t s);
void launch_adj_rhs_cells(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, const double* ka, double sa,
                          const double* kb, double* lam_S, int assign_S, double* cx, double* cy, hipStream_t s);
// The face part of the transposed time derivative at S (both axes, one launch): FP64 atomics into lam_S.
void launch_adj_faces(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, const double* cx, const double* cy,
                      double* lam_S, hipStream_t s);
// step-size adjoint: g = sum of the partials p1 and p2 (each may be null) (+ gh[0] when not null, the
// modules' part); h = eps * min dt(L) over the interior; lam_L(c*) += eps * g * grad dt(c*)
void launch_adj_dt(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double eps, const double* p1, const double* p2,
                   const double* gh, double* pmin, int* pidx, double* lam_L, hipStream_t s);
// adjoint of populateVariablesFromState (idealmhd.cpp:226-252) from the state grids rho and temp:
// g_state (8*N: rho, temp, mom_x, mom_y, mom_z, bi_x, bi_y, bi_z) from lam (8*N, evolved order)
void launch_adj_populate(const double* rho_s, const double* temp_s, const double* lam, MhdConst k, size_t N,
                         double* g_state, hipStream_t s);
void launch_fill(double* p, double v, size_t n, hipStream_t s);

}  // namespace spruce
