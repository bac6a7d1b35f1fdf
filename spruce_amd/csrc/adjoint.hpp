// Discrete adjoint of the IdealMHD time step on the GPU: kernels (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

constexpr int ADJ_COEFS = 17;

// lam_S += d(k . lamk)/dS: the transpose of the per-cell time derivative applied to lamk (8*N),
// using cx, cy (17*N each) as scratch.
void launch_adj_rhs(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, const double* lamk, double* cx, double* cy,
                    double* lam_S, hipStream_t s);
// Reverse of T = minimums/derived(B + mult*h*kv) (equationset.cpp:222-230, idealmhd.cpp:234-252):
// lam_B += mu, lamk = mult*h*mu with mu = dT/dpre^T lam_T, and per-block partial sums of
// mult * mu . kv (the adjoint of the step size h).
int adj_blocks(const Mesh& m);
void launch_adj_phi(MhdSet B, const double* kv, const double* step, double mult, MhdConst k, Mesh mesh,
                    const double* lam_T, double* lam_B, double* lamk, double* partial, hipStream_t s);
// gh += sum of the partials (one block)
void launch_adj_sum(const double* partial, int n, double* gh, hipStream_t s);
// step-size adjoint: h = eps * min dt(L) over the interior; lam_L(c*) += eps * gh * grad dt(c*)
void launch_adj_dt(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double eps, const double* gh, double* pmin,
                   int* pidx, double* lam_L, hipStream_t s);
// adjoint of populateVariablesFromState (idealmhd.cpp:226-252) from the state grids rho and temp:
// g_state (8*N: rho, temp, mom_x, mom_y, mom_z, bi_x, bi_y, bi_z) from lam (8*N, evolved order)
void launch_adj_populate(const double* rho_s, const double* temp_s, const double* lam, MhdConst k, size_t N,
                         double* g_state, hipStream_t s);
void launch_fill(double* p, double v, size_t n, hipStream_t s);

}  // namespace spruce
