// Device-side data layout and kernel launchers of the IdealMHD time step.
#pragma once
#include <hip/hip_runtime.h>

namespace spruce {

// One axis of the (separable) mesh: d_x depends on x only and d_y on y only, so each is a 1-D array
// (the .state file stores them as full grids; setup checks they are separable).
// Face c lies between cells c-1 and c (wrapping on a periodic axis).
struct Axis {
  const double* d;      // cell size
  const double* rd;     // RN(1/d)
  const double* h;      // 0.5*d                      (derivs.cpp:482-484)
  const double* hsq;    // (0.5*d)^2                   (derivs.cpp:423)
  const double* rhsq;   // RN(1/hsq)
  const double* fden;   // face c: h[c] + h[c-1]       (derivs.cpp:486)
  const double* rfden;  // RN(1/fden)
  const double* gup;    // face c: h[c-2] + 2 h[c-1]   (derivs.cpp:499, extrapolation from below)
  const double* gdn;    // face c: h[c+1] + 2 h[c]     (extrapolation from above)
  int n, lo, hi, per;   // cells, interior bounds, periodic on both sides
  int dlo, dhi;         // bounds of the step-size minimum (plasmadomain.cpp:155-159)
};

// Per-cell state of one IdealMHD grid set.  rho and n are both kept because after
// recomputeDerivedVarsFromEvolvedVars rho = n*m_i (idealmhd.cpp:246-247) and n is not recoverable
// from rho; rho_pre holds the post-minimum, pre-derived density the open boundary reads.
struct MhdSet {
  double *rho, *n, *mx, *my, *mz, *te, *bx, *by, *bz, *rho_pre;
};

struct MhdStatic {
  const double *bex, *bey, *bez, *gx, *gy;
};

struct MhdConst {
  double m, rm;          // ion mass and RN(1/m)
  double gamma, gm1, rgm1;
  double nmin, temin, tmin, eps;
  double obs;            // open_boundary_strength
  double r4pi;           // RN(1/(4 pi))
  double r6, r3;         // RN(1/6), RN(1/3)
};

struct Mesh {
  Axis x, y;
  int nx, ny;
};

enum StageMode : int {
  STAGE_PLAIN = 0,       // T = B + step*k(S)
  STAGE_STORE_K = 1,     // as PLAIN, and store k (RK4 stages 1-3)
  STAGE_RK4_FINAL = 2    // k = (k1+k4)/6 + (k2+k3)/3, T = B + step*k
};

// Boundary modes of the ghost fill.
enum GhostMode : int {
  GHOST_NORMAL = 0,      // propagateChanges on the target set: write ghosts, then derived
  GHOST_LIVE_RAW = 1     // ghost writes into the live set during an intermediate propagate, no derived
};

struct StageArgs {
  MhdSet S;              // set the derivatives are evaluated on
  MhdSet B;              // base of the update
  MhdSet T;              // target
  MhdStatic st;
  MhdConst k;
  Mesh mesh;
  const double* step;    // device step size (epsilon * min dt)
  double step_mult;      // 0.5 for half steps, 1.0 otherwise (applied as mult*step, one rounding)
  double* kbuf;          // 8 * nx*ny: k store (STORE_K) or k1 (RK4_FINAL)
  const double* k2buf;   // RK4_FINAL: k2
  const double* k3buf;   // RK4_FINAL: k3
  double* kf_out;        // RK4_FINAL, when not null: the combined (k1 + k4)/6 + (k2 + k3)/3 (the adjoint's record)
  int mode;
  int write_pre;         // write rho_pre (open boundaries present)
  int write_n;
  // Module RHS terms (computeTimeDerivativesModule, equationset.cpp:204-210): k[extra_var[e]] +=
  // extra[e][c] in order, already multiplied by the ghost mask.
  const double* extra[4];
  int extra_var[4];
  int n_extra;
  // open_moc characteristic terms (8 * nx*ny, evolved order), added to the ghost cells' RHS; null
  // without open_moc sides
  const double* moc;
  // tile form only, null otherwise: one minimum per block of the target's per-cell dt over the dt
  // bounds (the final stage of a step that no ghost fill, limiter or module follows)
  double* dt_partial;
};
constexpr int MAX_RHS_TERMS = 4;

void launch_mhd_stage(const StageArgs& a, hipStream_t s);
// Wavefront-strip form of the same stage (mhd2.hip); R rows per wavefront, a multiple of 4.
void launch_mhd_stage_strip(const StageArgs& a, int R, hipStream_t s);
int mhd_strip_rows_default(int nx, int ny);
// 0 (strips) when the grid gives the strip kernel enough wavefronts, else 2 (tiles)
int mhd_stage_kernel_default(int nx, int ny);
// Tile form (mhd.hip): a block of 192 threads owns an 8x8 tile; each face of the tile is computed
// once, by one thread, into LDS, then one thread per cell combines its four faces.  For grids too
// small to fill the GPU with strips.  mhd_tile_blocks() blocks (and dt_partial entries).
void launch_mhd_stage_tile(const StageArgs& a, hipStream_t s);
int mhd_tile_blocks(const Mesh& m);
void launch_mhd_populate(MhdSet L, const double* temp, MhdConst k, int nx, int ny, hipStream_t s);
// Ghost fill of the sides whose kinds[] entry is non-zero; s1/s2 are the open-boundary decay
// factors of each side (evolution.cpp:166-167), computed on the host with the C library's pow.
void launch_mhd_ghost(MhdSet G, MhdConst k, Mesh mesh, const int* kinds, const double* s1, const double* s2,
                      int mode, hipStream_t s);
// Per-cell dt (also written to dt_grid when non-null), reduced over the dt bounds to one minimum
// per block (NaN dropped as std::min does); mhd_dt_blocks() blocks are launched.
int mhd_dt_blocks(const Mesh& m);
void launch_mhd_dt(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double* dt_grid, double* partial, hipStream_t s);
// global_visc_coeff of IdealMHD::computeTimeDerivativesDerived (idealmhd.cpp:90):
// *vc = (global_viscosity * 0.5) * min over the dt bounds of (1/(1/dx^2 + 1/dy^2)) / dt(S); *err = 1
// when the minimum is not finite.  partial holds mhd_dt_blocks() values.
void launch_moc_visc(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, double gv, double* partial, double* vc, int* err,
                     hipStream_t s);
// out[0..n) = scal, out[n] = *step, out[n+1] = *err (out: host memory from hipHostMalloc)
void launch_gather_scalars(const double* scal, int n, const double* step, const int* err, double* host_out,
                           hipStream_t s);
void launch_reduce_step(const double* partial, int n, double eps, double* step, double* history, int idx,
                        int* err, hipStream_t s);
// with moc non-null the open_moc characteristic terms are added to the ghost cells (idealmhd.cpp:99-103)
void launch_mhd_rhs_full(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, double* out, hipStream_t s,
                         const double* moc = nullptr);
// Derived variable `which` (see kDerived in domain.cpp) of a set, written to out.
void launch_mhd_derived(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, int which, double* out, hipStream_t s);

}  // namespace spruce
