// Discrete-adjoint kernels of the ghost fill, live propagate and modules for gfx950 (reverse of
// mhd.hip's ghost fill, solar.hip's conduction and radiation kernels and the heating update).
// Conduction is reversed by hand (tc_cell.hpp, checked against the same function in forward mode);
// the ghost fill of open sides and the radiative-loss subcycles take their local Jacobians in
// forward mode (dual.hpp).  Contributions to neighbouring cells use FP64 atomics, so results are
// reproducible to rounding.
#include <cfloat>

#include "adj_cell.hpp"
#include "adj_modules.hpp"
#include "host_common.hpp"
#include "arith.hpp"
#include "dual.hpp"
#include "sources.hpp"
#include "tc_cell.hpp"
#include "tc_strip.hpp"

namespace spruce {
namespace {

using tcc::inside;
using tcc::wrp;

__device__ __forceinline__ double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double part[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  if ((lane & 63) == 0) part[lane >> 6] = v;
  __syncthreads();
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    double m = 0.0;
    for (int k = 0; k < nw; ++k) m += part[k];
    v = m;
  }
  return v;
}

// ---- ghost fill ------------------------------------------------------------------------------

struct AdjGhostArgs {
  MhdSet P;
  MhdConst k;
  Mesh mesh;
  int kind[4];
  double s1[4], s2[4];
  int raw;   // GHOST_LIVE_RAW: open sides read rho (not rho_pre) and write without floors
};

// Open boundary ghost values of one boundary cell row (mesh.hip ghost_side, kind 1, GHOST_NORMAL)
// from x = (rho_pre, te, mom_x, mom_y) of the first interior cell: out = rho, te, mom_x, mom_y of
// the two ghost cells (order c1, c2).
__device__ void open_ghost_dual(const MhdConst& K, const Axis& ax, bool xside, bool upper, double s1, double s2,
                                const double x[4], bool raw, Dn<4> out[8]) {
  const int n = ax.n;
  const int e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
  const Dn<4> rho3 = dn_var<4>(x[0], 0), te3 = dn_var<4>(x[1], 1), mx3 = dn_var<4>(x[2], 2), my3 = dn_var<4>(x[3], 3);
  const Dn<4> rg1 = s1 * rho3, rg2 = s2 * rho3, tg1 = s1 * te3, tg2 = s2 * te3;
  Dn<4> cs = dn_const<4>(0.0);
  const Dn<4> csn = dn_sqrt(K.gamma * (te3 * K.gm1) / rho3);
  if (csn.v > cs.v) cs = csn;
  const Dn<4> vx = mx3 / rho3, vy = my3 / rho3;
  Dn<4> boost = K.obs * cs;
  if (!upper) boost = boost * -1.0;
  const double dist = 0.5 * (ax.d[e2] + ax.d[e3]);
  const Dn<4> vn = xside ? vx : vy;
  const Dn<4> bv = upper ? dn_max(dn_const<4>(0.0), vn + boost) : dn_min(dn_const<4>(0.0), vn + boost);
  const Dn<4> gv = (dist * bv - (0.5 * ax.d[e2]) * vn) / (0.5 * ax.d[e3]);
  if (raw) {
    out[0] = rg1;
    out[1] = rg2;
    out[2] = tg1;
    out[3] = tg2;
  } else {
    const Dn<4> n1 = dn_max(rg1 / K.m, K.nmin), n2 = dn_max(rg2 / K.m, K.nmin);
    out[0] = n1 * K.m;
    out[1] = n2 * K.m;
    out[2] = dn_max(tg1, K.temin);
    out[3] = dn_max(tg2, K.temin);
  }
  out[4] = rg1 * (xside ? gv : vx);
  out[5] = rg2 * (xside ? gv : vx);
  out[6] = rg1 * (xside ? vy : gv);
  out[7] = rg2 * (xside ? vy : gv);
}

__device__ void adj_ghost_side(const AdjGhostArgs& a, int side, int t, double* lam) {
  const int kind = a.kind[side];
  if (kind != 1 && kind != 2 && kind != 3 && kind != 5) return;
  const Mesh& M = a.mesh;
  const bool xside = side < 2, upper = side & 1;
  const Axis& ax = xside ? M.x : M.y;
  const Axis& ot = xside ? M.y : M.x;
  if (kind == 2) {
    if (t < 0 || t >= ot.n) return;
  } else {
    if (t < ot.lo || t > ot.hi) return;
  }
  const int n = ax.n;
  const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
  auto at = [&](int e) -> size_t { return xside ? (size_t)e * M.ny + t : (size_t)t * M.ny + e; };
  const size_t c1 = at(e1), c2 = at(e2), c3 = at(e3);
  const size_t N = (size_t)M.nx * M.ny;
  double* lr = lam + 0 * N; double* lmx = lam + 1 * N; double* lmy = lam + 2 * N; double* lmz = lam + 3 * N;
  double* lte = lam + 4 * N;
  if (kind == 5) {                 // open_ucnp: both ghost cells copy the first interior cell
    const size_t ci = at(upper ? ax.hi : ax.lo);
    for (int g = N_GHOST - 1; g >= 0; --g) {
      const size_t cb = at(upper ? ax.hi + 1 + g : g);
      for (int v = 0; v < 8; ++v) {
        lam[v * N + ci] += lam[v * N + cb];
        lam[v * N + cb] = 0.0;
      }
    }
    return;
  }
  if (kind == 2) {                 // fixed: momenta of c1, c2, c3 set to zero
    lmx[c1] = 0.0; lmx[c2] = 0.0; lmx[c3] = 0.0;
    lmy[c1] = 0.0; lmy[c2] = 0.0; lmy[c3] = 0.0;
    lmz[c1] = 0.0; lmz[c2] = 0.0; lmz[c3] = 0.0;
  } else if (kind == 3) {          // reflect: rho, te (and n) copied from c3, momenta zeroed
    lte[c3] += lte[c1] + lte[c2];
    lte[c1] = 0.0; lte[c2] = 0.0;
    lr[c3] += lr[c1] + lr[c2];
    lr[c1] = 0.0; lr[c2] = 0.0;
    lmx[c1] = 0.0; lmx[c2] = 0.0; lmx[c3] = 0.0;
    lmy[c1] = 0.0; lmy[c2] = 0.0; lmy[c3] = 0.0;
    lmz[c1] = 0.0; lmz[c2] = 0.0; lmz[c3] = 0.0;
  } else {                         // open
    const double x[4] = {a.raw ? a.P.rho[c3] : a.P.rho_pre[c3], a.P.te[c3], a.P.mx[c3], a.P.my[c3]};
    Dn<4> o[8];
    open_ghost_dual(a.k, ax, xside, upper, a.s1[side], a.s2[side], x, a.raw != 0, o);
    const double lo[8] = {lr[c1], lr[c2], lte[c1], lte[c2], lmx[c1], lmx[c2], lmy[c1], lmy[c2]};
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < 8; ++q)
      for (int k = 0; k < 4; ++k) g[k] += lo[q] * o[q].d[k];
    lr[c1] = 0.0; lr[c2] = 0.0; lte[c1] = 0.0; lte[c2] = 0.0;
    lmx[c1] = 0.0; lmx[c2] = 0.0; lmy[c1] = 0.0; lmy[c2] = 0.0;
    // normal fill: rho_pre of c3 and its rho come from the same pre-update density with the same
    // floors; raw fill: rho of c3 itself
    lr[c3] += g[0];
    lte[c3] += g[1];
    lmx[c3] += g[2];
    lmy[c3] += g[3];
  }
}

__global__ void __launch_bounds__(256) k_adj_ghost_axis(AdjGhostArgs a, int xsides, double* lam) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s0 = xsides ? 0 : 2;
  // the forward fills the lower side, then the upper one
  adj_ghost_side(a, s0 + 1, t, lam);
  adj_ghost_side(a, s0, t, lam);
}

// ---- minimums / derived ------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_adj_minmask(const double* rho_in, const double* te_base, const double* add,
                                                     const double* step, MhdConst K, Mesh M, double* lam,
                                                     double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double x = rho_in[c] / K.m;
    const double rho1 = smax(x, K.nmin) * K.m;
    const double drho = (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
    const double te_in = add ? te_base[c] + (*step) * add[c] : te_base[c];
    const double dte = te_in < K.temin ? 0.0 : 1.0;
    lam[c] *= drho;
    const double lt = lam[4 * N + c] * dte;
    lam[4 * N + c] = lt;
    if (add) dot = lt * add[c];
  }
  if (partial) {
    dot = block_sum(dot);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
  }
}

// ---- thermal conduction ----------------------------------------------------------------------

__device__ __forceinline__ double temp_of(double te, double n, const MhdConst& K) {
  return smax(te * K.gm1 / (n * (2.0 * K_B)), K.tmin);
}

// The conduction reverses scatter dJ/dT onto the 3x3 neighbourhood of every cell.  A block of 256
// threads (64 x 4, one cell each) owns a tile of 64 columns x 4 rows; the scatter goes into an LDS
// accumulator over the tile plus a one-cell ring, which is then added to lam_T: plainly for the
// cells whose whole neighbourhood lies in the tile (no other tile reaches them), atomically on the
// tile's border and ring (about one global atomic per cell instead of nine).  Every cell's
// contributions are the same; only their summation order differs from a scatter of global atomics.
constexpr int TCR_COLS = 64, TCR_ROWS = 4, TCR_W = TCR_COLS + 2, TCR_N = (TCR_ROWS + 2) * TCR_W;

__device__ __forceinline__ void tcr_zero(double* acc) {
  for (int t = threadIdx.y * 64 + threadIdx.x; t < TCR_N; t += 256) acc[t] = 0.0;
}
// cell at tile row li, column lj: its neighbour (i-1+a, j-1+b) is accumulator entry (li+a, lj+b)
__device__ __forceinline__ void tcr_add(double* acc, int li, int lj, const double gT[3][3], double ld) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) atomicAdd(&acc[(li + a) * TCR_W + lj + b], ld * gT[a][b]);
}
__device__ __forceinline__ void tcr_flush(const double* acc, const Mesh& M, int i0, int j0, double* lam_T) {
  const int vr = min(TCR_ROWS, M.nx - i0), vc = min(TCR_COLS, M.ny - j0);
  for (int t = threadIdx.y * 64 + threadIdx.x; t < TCR_N; t += 256) {
    const double v = acc[t];
    if (v == 0.0) continue;
    const int la = t / TCR_W, lb = t % TCR_W;
    int gi = i0 - 1 + la, gj = j0 - 1 + lb;
    if (gi < 0 || gi >= M.nx) {
      if (!M.x.per) continue;
      gi = gi < 0 ? gi + M.nx : gi - M.nx;
    }
    if (gj < 0 || gj >= M.ny) {
      if (!M.y.per) continue;
      gj = gj < 0 ? gj + M.ny : gj - M.ny;
    }
    const size_t g = (size_t)gi * M.ny + gj;
    if (la >= 2 && la <= vr - 1 && lb >= 2 && lb <= vc - 1) lam_T[g] += v;
    else atomicAdd(&lam_T[g], v);
  }
}
// The temperatures of the block's tile and its one-cell ring, each computed once (a division) into
// LDS; ring cells outside a closed grid are never read (interior cells have all neighbours inside).
__device__ __forceinline__ void tcr_load_temps(const double* te, const double* n, const MhdConst& K, const Mesh& M,
                                               int i0, int j0, double* Ts) {
  for (int t = threadIdx.y * 64 + threadIdx.x; t < TCR_N; t += 256) {
    const int a = t / TCR_W, b = t - a * TCR_W;
    const int ii = wrp(i0 - 1 + a, M.nx, M.x.per), jj = wrp(j0 - 1 + b, M.ny, M.y.per);
    double v = 0.0;
    if (ii >= 0 && ii < M.nx && jj >= 0 && jj < M.ny) {
      const size_t q = (size_t)ii * M.ny + jj;
      v = temp_of(te[q], n[q], K);
    }
    Ts[t] = v;
  }
}
// the 3x3 temperatures of tile cell (li, lj) from the LDS tile
__device__ __forceinline__ void tcr_temps(const double* Ts, int li, int lj, double T[3][3]) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) T[a][b] = Ts[(li + a) * TCR_W + lj + b];
}
__device__ __forceinline__ void tcr_partial(double dot, double* partial, int accumulate) {
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    partial[b] = accumulate ? partial[b] + dot : dot;
  }
}

// The saturated energy derivative from the unsaturated R at interior cell (i, j) (as the forward
// forms it, thermalconduction.cpp:171-174); cc receives the coefficient, dcx/dcy its derivatives.
__device__ __forceinline__ double tc_sat_apply(const TcSatRev& sr, const Mesh& M, int i, int j, double R, double& cc,
                                               double& dcx, double& dcy) {
  const int ny = M.ny;
  const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
  const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
  const size_t c = (size_t)i * ny + j;
  cc = sr.coef[c];
  dcx = qdiv(tcc::fint(M.x, ip, i, cc, sr.coef[(size_t)ip * ny + j]) - tcc::fint(M.x, i, im, sr.coef[(size_t)im * ny + j], cc),
             M.x.d[i], M.x.rd[i]);
  dcy = qdiv(tcc::fint(M.y, jp, j, cc, sr.coef[(size_t)i * ny + jp]) - tcc::fint(M.y, j, jm, sr.coef[(size_t)i * ny + jm], cc),
             M.y.d[j], M.y.rd[j]);
  const double add = -1.0 * (dcx * sr.rx[c] + dcy * sr.ry[c]);
  return cc * R + add;
}
// the saturation outputs of one cell for the seed ld of its saturated derivative (zero seed: zeros)
__device__ __forceinline__ void tc_sat_out(const TcSatRev& sr, size_t c, double ld, double R, double dcx,
                                           double dcy) {
  sr.lcoef[c] = ld * R;
  sr.wx[c] = -ld * sr.rx[c];
  sr.wy[c] = -ld * sr.ry[c];
  sr.lfx[c] = -ld * dcx;
  sr.lfy[c] = -ld * dcy;
}
__device__ __forceinline__ void tc_sat_zero(const TcSatRev& sr, size_t c) {
  if (sr.coef) tc_sat_out(sr, c, 0.0, 0.0, 0.0, 0.0);
}

template <bool SAT>
__global__ void __launch_bounds__(256) k_adj_tc_sub(const double* te_s, const double* n, const double* bhx,
                                                    const double* bhy, const double* dbxx, const double* dbxy,
                                                    const double* dbyx, const double* dbyy, double h, double k0,
                                                    MhdConst K, Mesh M, double* lam_te, double* lam_T,
                                                    double* lam_fz, double* partial, int accumulate, TcSatRev sr) {
  __shared__ double acc[TCR_N];
  __shared__ double Ts[TCR_N];
  const int i0 = blockIdx.y * TCR_ROWS, j0 = blockIdx.x * TCR_COLS;
  tcr_zero(acc);
  tcr_load_temps(te_s, n, K, M, i0, j0, Ts);
  __syncthreads();
  double dot = 0.0;
  do {
    const int li = threadIdx.y, lj = threadIdx.x;
    const int i = i0 + li, j = j0 + lj;
    if (i >= M.nx || j >= M.ny) break;
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double base = te_s[c];
    if (!inside(M, i, j)) {
      lam_te[c] *= (base + 0.0 < K.temin) ? 0.0 : 1.0;
      if (SAT) tc_sat_zero(sr, c);
      continue;
    }
    double T[3][3];
    tcr_temps(Ts, li, lj, T);
    // the clamp of the update needs dE first; its reverse is linear in the seed, so it runs with
    // a unit seed and is scaled afterwards
    double gT[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double gfz[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double R = tcc::tc_core_reverse(M, i, j, T, bhx[c], bhy[c], dbxx[c], dbxy[c], dbyx[c], dbyy[c], k0, 1.0,
                                          gT, gfz);
    double cc = 1.0, dcx = 0.0, dcy = 0.0;
    const double dE = SAT ? tc_sat_apply(sr, M, i, j, R, cc, dcx, dcy) : R;
    const double v = base + 1.0 * (1.0 * h) * dE;
    const double lv = (v < K.temin) ? 0.0 : lam_te[c];
    lam_te[c] = lv;
    dot += lv * dE;
    const double ld = lv * h;
    if (SAT) tc_sat_out(sr, c, ld, R, dcx, dcy);
    if (ld != 0.0) {
      tcr_add(acc, li, lj, gT, ld * cc);
      for (int k = 0; k < 6; ++k) lam_fz[k * N + c] += (ld * cc) * gfz[k];
    }
  } while (false);   // one cell per thread; `continue`/`break` leave the cell
  __syncthreads();
  tcr_flush(acc, M, i0, j0, lam_T);
  tcr_partial(dot, partial, accumulate);
}

// Generalised subcycle stage reverse: the energy derivative dE is evaluated at the temperatures of
// `eval`; with cbase given the stage output was max(cbase + mask*(mult*h)*dE, te_min), whose
// adjoint lam_out is reversed through the clamp into lam_base (assigned or added) and seeds dE with
// (mult*h)*mu; seed_add (may be null) adds add_coef * seed_add[c] to that seed (the weights a later
// RK4 combination gives this stage's dE).
template <bool SAT>
__global__ void __launch_bounds__(256) k_adj_tc_stage(const double* eval, const double* n, const double* bhx,
                                                      const double* bhy, const double* dbxx, const double* dbxy,
                                                      const double* dbyx, const double* dbyy, double h, double mult,
                                                      double k0, MhdConst K, Mesh M, const double* cbase,
                                                      const double* lam_out, double* lam_base, int assign_base,
                                                      const double* seed_add, double add_coef, double* lam_T,
                                                      double* lam_fz, double* partial, int accumulate, TcSatRev sr) {
  __shared__ double acc[TCR_N];
  __shared__ double Ts[TCR_N];
  const int i0 = blockIdx.y * TCR_ROWS, j0 = blockIdx.x * TCR_COLS;
  tcr_zero(acc);
  tcr_load_temps(eval, n, K, M, i0, j0, Ts);
  __syncthreads();
  double dot = 0.0;
  do {
    const int li = threadIdx.y, lj = threadIdx.x;
    const int i = i0 + li, j = j0 + lj;
    if (i >= M.nx || j >= M.ny) break;
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    if (!inside(M, i, j)) {
      if (cbase) {
        const double mu = (cbase[c] + 0.0 < K.temin) ? 0.0 : lam_out[c];
        lam_base[c] = assign_base ? mu : lam_base[c] + mu;
      }
      if (SAT) tc_sat_zero(sr, c);
      continue;
    }
    double T[3][3];
    tcr_temps(Ts, li, lj, T);
    double gT[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double gfz[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double R = tcc::tc_core_reverse(M, i, j, T, bhx[c], bhy[c], dbxx[c], dbxy[c], dbyx[c], dbyy[c], k0, 1.0,
                                          gT, gfz);
    double cc = 1.0, dcx = 0.0, dcy = 0.0;
    const double dE = SAT ? tc_sat_apply(sr, M, i, j, R, cc, dcx, dcy) : R;
    double ld = 0.0;
    if (cbase) {
      const double v = cbase[c] + 1.0 * (mult * h) * dE;
      const double mu = (v < K.temin) ? 0.0 : lam_out[c];
      lam_base[c] = assign_base ? mu : lam_base[c] + mu;
      dot += mu * mult * dE;
      ld = mu * (mult * h);
    }
    if (seed_add) ld += add_coef * seed_add[c];
    if (SAT) tc_sat_out(sr, c, ld, R, dcx, dcy);
    if (ld != 0.0) {
      tcr_add(acc, li, lj, gT, ld * cc);
      for (int k = 0; k < 6; ++k) lam_fz[k * N + c] += (ld * cc) * gfz[k];
    }
  } while (false);   // one cell per thread; `continue`/`break` leave the cell
  __syncthreads();
  tcr_flush(acc, M, i0, j0, lam_T);
  tcr_partial(dot, partial, accumulate);
}

// Reverse of an RK4 subcycle's final combination out = max(te + (1.0*h*S)/6, te_min) with
// S = k1 + 2 k2 + 2 k3 + k4, k4 = dE at the temperatures of i3 (thermalconduction.cpp:122-124):
// lam_base = mu (assigned), lamS = mu*h/6 (the seed of each k, times its weight), k4's transpose
// applied here, partial = mu*S/6.
template <bool SAT>
__global__ void __launch_bounds__(256) k_adj_tc_final4(const double* te, const double* k1, const double* k2,
                                                       const double* k3, const double* i3, const double* n,
                                                       const double* bhx, const double* bhy, const double* dbxx,
                                                       const double* dbxy, const double* dbyx, const double* dbyy,
                                                       double h, double k0, MhdConst K, Mesh M,
                                                       const double* lam_out, double* lam_base, double* lamS,
                                                       double* lam_T, double* lam_fz, double* partial,
                                                       int accumulate, TcSatRev sr) {
  __shared__ double acc[TCR_N];
  __shared__ double Ts[TCR_N];
  const int i0 = blockIdx.y * TCR_ROWS, j0 = blockIdx.x * TCR_COLS;
  tcr_zero(acc);
  tcr_load_temps(i3, n, K, M, i0, j0, Ts);
  __syncthreads();
  double dot = 0.0;
  do {
    const int li = threadIdx.y, lj = threadIdx.x;
    const int i = i0 + li, j = j0 + lj;
    if (i >= M.nx || j >= M.ny) break;
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    if (!inside(M, i, j)) {
      lam_base[c] = (te[c] + 0.0 < K.temin) ? 0.0 : lam_out[c];
      lamS[c] = 0.0;
      if (SAT) tc_sat_zero(sr, c);
      continue;
    }
    double T[3][3];
    tcr_temps(Ts, li, lj, T);
    double gT[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double gfz[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double R = tcc::tc_core_reverse(M, i, j, T, bhx[c], bhy[c], dbxx[c], dbxy[c], dbyx[c], dbyy[c], k0, 1.0,
                                          gT, gfz);
    double cc = 1.0, dcx = 0.0, dcy = 0.0;
    const double dE = SAT ? tc_sat_apply(sr, M, i, j, R, cc, dcx, dcy) : R;
    const double sum = k1[c] + 2.0 * k2[c] + 2.0 * k3[c] + dE;
    const double v = te[c] + qdiv(1.0 * h * sum, 6.0, K.r6);
    const double mu = (v < K.temin) ? 0.0 : lam_out[c];
    lam_base[c] = mu;
    const double ls = mu * h / 6.0;
    lamS[c] = ls;
    dot += mu * sum / 6.0;
    if (SAT) tc_sat_out(sr, c, ls, R, dcx, dcy);
    if (ls != 0.0) {
      tcr_add(acc, li, lj, gT, ls * cc);
      for (int k = 0; k < 6; ++k) lam_fz[k * N + c] += (ls * cc) * gfz[k];
    }
  } while (false);   // one cell per thread; `continue`/`break` leave the cell
  __syncthreads();
  tcr_flush(acc, M, i0, j0, lam_T);
  tcr_partial(dot, partial, accumulate);
}

__global__ void __launch_bounds__(256) k_adj_tc_temp(const double* te_s, const double* n, MhdConst K, size_t N,
                                                     double* lam_T, double* lam_te, double* lam_rho) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double lT = lam_T[c];
  if (lT == 0.0) return;
  lam_T[c] = 0.0;
  const double te = te_s[c], nn = n[c];
  const double den = nn * (2.0 * K_B);
  const double x = te * K.gm1 / den;
  if (x < K.tmin) return;
  lam_te[c] += lT * (K.gm1 / den);
  lam_rho[c] += lT * (-x / nn) * K.rm;
}

// Strip form of the reverse of one Euler conduction subcycle without flux saturation,
// te_out = max(te_s + h dE(te_s), te_min) (thermalconduction.cpp:60-66, 158-176).  A 64-lane
// wavefront holds columns 60*tile - 2 + lane and marches rows r0-1 .. r1 of its strip, recomputing
// every cell's face and row values as the forward strip kernel forms them (tc_strip.hpp), with the
// geometric divisions taken as products with the prepared reciprocals and the powers of T as
// T sqrt(T), T^2 sqrt(T): within a few ulps of the forward's values, which only the clamp's test
// (v < te_min) and the step-size sum take; lanes 1..62 reverse their cells.  The reverse of tc_combine at a cell goes to the values it was formed
// from, and these go to the temperatures and b_hat of the rows and lanes around the cell: along y
// through the transposes of the forward's DPP lane shifts, along x through the face between two
// rows, held in registers until both of its cells have added to it.  A row's adjoint is complete
// once the next row is done; lanes 2..61 of rows r0..r1-1 then own it: they write the adjoint of
// te_s (the clamp's seed plus dJ/dT through the temperature) and add dJ/drho and dJ/d b_hat.  No
// atomics, and dJ/dT and the b_hat derivatives' adjoints never reach memory.
__global__ void __launch_bounds__(64) k_adj_tc_strip(TcStepArgs a, const double* lam_in, double* lam_out,
                                                     double* lam_rho, double* lam_bx, double* lam_by, double* partial,
                                                     int accumulate, int R, int nstrips, int total) {
  const Mesh& M = a.mesh;
  const MhdConst& K = a.k;
  const int t = strip::xcd_tile(blockIdx.x, total);
  const int tile = t / nstrips, sidx = t - tile * nstrips;
  const int lane = threadIdx.x;
  const int jc = tile * 60 + lane - 2;
  tcs::Lane L;
  L.j = tcs::row_index(jc, M.ny, M.y.per, L.val);
  L.iny = L.val && L.j >= M.y.lo && L.j <= M.y.hi;
  L.out = lane >= 2 && lane <= 61 && jc < M.ny;
  {
    bool vp;
    const int jp = tcs::row_index(L.j + 1, M.ny, M.y.per, vp);
    L.hc = M.y.h[jp];
    L.hn = M.y.h[L.j];
    L.fd = M.y.fden[jp];
    L.rfd = M.y.rfden[jp];
    L.dy = M.y.d[L.j];
    L.rdy = M.y.rd[L.j];
    L.hsq = M.y.hsq[L.j];
    L.rhsq = M.y.rhsq[L.j];
  }
  const bool cmp = lane >= 1 && lane <= 62 && L.val;   // reverses its cells
  const double wc = L.hc * L.rfd, wn = L.hn * L.rfd;   // d(upper y-face)/d(own value, next lane's)
  const int r0 = sidx * R, r1 = min(r0 + R, M.nx);
  tcs::Raw rw = tcs::load_raw<false>(a, L, r0 - 2);
  const tcs::Row first = tcs::make_row<false, false>(a, L, r0 - 2, rw);
  rw = tcs::load_raw<false>(a, L, r0 - 1);
  tcs::Row cur = tcs::make_row<false, false>(a, L, r0 - 1, rw);
  tcs::XFaces lo = tcs::xfaces<false, false>(M, first, cur);
  double cn = rw.n;
  tcs::Raw ahead = tcs::load_raw<false>(a, L, r0);
  // adjoints of the previous row (p*) and of the face between it and the current row (f*): T, gy
  // (as masked), b_hat; ps1 = the previous row's seed of secondDerivative1D along y times 1/h^2
  double pT = 0.0, pgy = 0.0, pbx = 0.0, pby = 0.0, ps1 = 0.0, plv = 0.0, pte = 0.0, pn = 1.0;
  double fT = 0.0, fgy = 0.0, fbx = 0.0, fby = 0.0;
  int pi = first.i;
  bool pown = false, pgm = false;
  double dot = 0.0;
  for (int r = r0 - 1; r <= r1; ++r) {
    const tcs::Raw raw = ahead;
    if (r < r1) ahead = tcs::load_raw<false>(a, L, r + 2);
    const tcs::Row nxt = tcs::make_row<false, false>(a, L, r + 1, raw);
    const tcs::XFaces hi = tcs::xfaces<false, false>(M, cur, nxt);
    // the cell of row r, as the forward forms it
    const int i = cur.i;
    const double rdx = M.x.rd[i];
    const bool in = cur.inx && L.iny;
    const double gx = (hi.T - lo.T) * rdx;
    const double dgx_y = tcs::yderiv<false>(L, tcs::yface<false>(L, in ? gx : 0.0));
    const bool live = cmp && cur.val;
    double lv = 0.0;
    tcc::TcCombAdj G{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (live) {
      const size_t c = (size_t)i * M.ny + L.j;
      const double lin = lam_in[c];
      if (in) {
        // dE and its partial derivatives for a unit seed of the update (seed h), scaled by the
        // clamp's seed afterwards
        const double s0 = (hi.T - 2.0 * cur.T + lo.T) * M.x.rhsq[i];
        const double dgy_x = (hi.gy - lo.gy) * rdx;
        const double dbxx = (hi.bx - lo.bx) * rdx;
        const double dbyx = (hi.by - lo.by) * rdx;
        const double dE = tcc::tc_combine_reverse<false>(cur.T, gx, cur.gy, s0, cur.s1, dgx_y, dgy_x, cur.bx, cur.by,
                                                         dbxx, cur.dbxy, dbyx, cur.dbyy, a.p.k0, a.h, G);
        const double v = cur.te + 1.0 * (1.0 * a.h) * dE;
        lv = (v < K.temin) ? 0.0 : lin;
        if (L.out && r >= r0 && r < r1) dot += lv * dE;
        G.T *= lv; G.gx *= lv; G.gy *= lv; G.s0 *= lv; G.s1 *= lv; G.dgx_y *= lv; G.dgy_x *= lv;
        G.bx *= lv; G.by *= lv; G.dbx_x *= lv; G.dbx_y *= lv; G.dby_x *= lv; G.dby_y *= lv;
      } else {
        lv = (cur.te + 0.0 < K.temin) ? 0.0 : lin;   // mask = 0: dE = 0
      }
    }
    // this row's transposes along y: dgx_y onto gx of lanes l-1..l+1 (as masked), the y-derivatives
    // of b_hat onto b_hat
    double lgx;
    {
      const double u = G.dgx_y * L.rdy;
      const double lf = u - strip::shl1(u);
      const double lg = lf * wc + strip::shr1(lf * wn);
      lgx = G.gx + (in ? lg : 0.0);
    }
    double cT, cgy = G.gy, cbx = G.bx, cby = G.by;
    {
      const double ub = G.dbx_y * L.rdy, vb = G.dby_y * L.rdy;
      const double fb = ub - strip::shl1(ub), gb = vb - strip::shl1(vb);
      cbx += fb * wc + strip::shr1(fb * wn);
      cby += gb * wc + strip::shr1(gb * wn);
    }
    const double cs1 = G.s1 * L.rhsq;
    const double a_gx = lgx * rdx, a_s0 = G.s0 * M.x.rhsq[i];
    cT = G.T - 2.0 * a_s0 - 2.0 * cs1;
    const double a_gyx = G.dgy_x * rdx, a_bxx = G.dbx_x * rdx, a_byx = G.dby_x * rdx;
    // the face between rows r-1 and r is complete: xfaces(A = row r-1, B = row r) transposed
    fT += a_s0 - a_gx;
    fgy -= a_gyx;
    fbx -= a_bxx;
    fby -= a_byx;
    {
      const double rf = M.x.rfden[i];
      const double wa = M.x.h[i] * rf, wb = M.x.h[pi] * rf;
      pT += fT * wa;
      cT += fT * wb;
      pgy += fgy * wa;
      cgy += fgy * wb;
      pbx += fbx * wa;
      cbx += fbx * wb;
      pby += fby * wa;
      cby += fby * wb;
    }
    // row r-1 is complete: its gy (masked as the forward masks it) and secondDerivative1D along y
    // onto the temperatures of lanes l-1..l+1, then through the temperature onto te_s and rho
    {
      const double u = (pgm ? pgy : 0.0) * L.rdy;
      const double lf = (u + ps1) + strip::shl1(ps1 - u);
      pT += lf * wc + strip::shr1(lf * wn);
      if (pown) {
        const size_t q = (size_t)pi * M.ny + L.j;
        const double den = pn * (2.0 * K_B);
        const double x = pte * K.gm1 / den;
        double lte = plv;
        if (!(x < K.tmin)) {
          lte += pT * (K.gm1 / den);
          lam_rho[q] += pT * (-x / pn) * K.rm;
        }
        lam_out[q] = lte;
        lam_bx[q] += pbx;
        lam_by[q] += pby;
      }
    }
    // row r becomes the previous row; the face above it holds row r's part
    pT = cT; pgy = cgy; pbx = cbx; pby = cby; ps1 = cs1; plv = lv; pte = cur.te; pn = cn; pi = i;
    pgm = in;
    pown = L.out && cur.val && r >= r0 && r < r1;
    fT = a_s0 + a_gx;
    fgy = a_gyx;
    fbx = a_bxx;
    fby = a_byx;
    cur = nxt;
    lo = hi;
    cn = raw.n;
  }
  for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off);
  if (lane == 0) partial[blockIdx.x] = accumulate ? partial[blockIdx.x] + dot : dot;
}

// derivative1D(q, axis) at interior cell (i, j) transposed: d(c) depends on q(c-1), q(c), q(c+1)
__device__ __forceinline__ void scatter_deriv(const Mesh& M, int axis, int i, int j, double ld, double* lq) {
  const Axis& ax = axis == 0 ? M.x : M.y;
  const int c = axis == 0 ? i : j;
  const int cm = wrp(c - 1, ax.n, ax.per), cp = wrp(c + 1, ax.n, ax.per);
  auto at = [&](int k) -> size_t { return axis == 0 ? (size_t)k * M.ny + j : (size_t)i * M.ny + k; };
  const double r = ld / ax.d[c];
  // upper face (c, cp): (q(c) h[cp] + q(cp) h[c]) / fden[cp]; lower face (cm, c): (q(cm) h[c] + q(c) h[cm]) / fden[c]
  atomicAdd(&lq[at(cp)], r * (ax.h[c] / ax.fden[cp]));
  atomicAdd(&lq[at(c)], r * (ax.h[cp] / ax.fden[cp] - ax.h[cm] / ax.fden[c]));
  atomicAdd(&lq[at(cm)], -r * (ax.h[c] / ax.fden[c]));
}


// d derivative1D(q, axis)(at cell k) / d q(c) for the three cells k = c-1, c, c+1 along the axis
// (derivs.cpp:229-233, 477-487): the weights the coefficient's adjoint gathers
__device__ __forceinline__ double gather_deriv(const Mesh& M, int axis, int i, int j, const double* w) {
  const Axis& ax = axis == 0 ? M.x : M.y;
  const int c = axis == 0 ? i : j;
  const int cm = wrp(c - 1, ax.n, ax.per), cp = wrp(c + 1, ax.n, ax.per);
  auto at = [&](int k) -> size_t { return axis == 0 ? (size_t)k * M.ny + j : (size_t)i * M.ny + k; };
  double g = w[at(c)] * ((ax.h[cp] / ax.fden[cp] - ax.h[cm] / ax.fden[c]) / ax.d[c]);
  if (cm >= 0 && cm < ax.n) g += w[at(cm)] * (ax.h[cm] / ax.fden[c] / ax.d[cm]);
  if (cp >= 0 && cp < ax.n) g -= w[at(cp)] * (ax.h[cp] / ax.fden[cp] / ax.d[cp]);
  return g;
}

__global__ void __launch_bounds__(256) k_adj_tc_satrev(const double* ev, const double* n, const double* rho,
                                                       const double* bhx, const double* bhy, TcParams p, MhdConst K,
                                                       Mesh M, TcSatRev sr, double* lam_T, double* lam_fz,
                                                       double* lam_rho) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny || !inside(M, i, j)) return;   // the coefficient is 1 off the interior
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  // the coefficient's adjoint: its own seed and the add terms of the cell and its four neighbours
  const double lc = sr.lcoef[c] + gather_deriv(M, 0, i, j, sr.wx) + gather_deriv(M, 1, i, j, sr.wy);
  double lfx = sr.lfx[c], lfy = sr.lfy[c];
  if (lc == 0.0 && lfx == 0.0 && lfy == 0.0) return;
  // the forward quantities of the cell (saturateConductiveFlux / fieldAlignedConductiveFlux)
  auto Tat = [&](int a, int b) {
    const size_t q = (size_t)wrp(a, M.nx, M.x.per) * M.ny + wrp(b, M.ny, M.y.per);
    return temp_of(ev[q], n[q], K);
  };
  const double T = Tat(i, j);
  const int ip = wrp(i + 1, M.nx, M.x.per), im = wrp(i - 1, M.nx, M.x.per);
  const int jp = wrp(j + 1, M.ny, M.y.per), jm = wrp(j - 1, M.ny, M.y.per);
  const double gx = qdiv(tcc::fint(M.x, ip, i, T, Tat(i + 1, j)) - tcc::fint(M.x, i, im, Tat(i - 1, j), T), M.x.d[i], M.x.rd[i]);
  const double gy = qdiv(tcc::fint(M.y, jp, j, T, Tat(i, j + 1)) - tcc::fint(M.y, j, jm, Tat(i, j - 1), T), M.y.d[j], M.y.rd[j]);
  const double r = rho[c], bx = bhx[c], by = bhy[c];
  const double kmax = M.x.d[i] * M.y.d[j] * K_B * (r / K.m) / p.dsm;
  const double kt = pow_2p5(T) * p.k0;
  const bool capped = kmax < kt;
  const double kap = capped ? kmax : kt;
  const double bdg = gx * bx + gy * by;
  const double fmag = -kap * bdg;
  const double fx = fmag * bx, fy = fmag * by;
  const double sat = (r / K.m) * p.c_sat * pow_1p5(T * K_B) / p.sqrt_me;
  const double fm2 = fx * fx + fy * fy;
  double lT = 0.0, lr = 0.0;
  if (fm2 > 0.0) {
    // coef = sat / sqrt(sat^2 + |F|^2)
    const double D = sat * sat + fm2, D32 = D * sqrt(D);
    const double lsat = lc * (fm2 / D32);
    lfx += lc * (-sat * fx / D32);
    lfy += lc * (-sat * fy / D32);
    lT += lsat * 1.5 * sat / T;
    lr += lsat * sat / r;
  }
  // F = fmag b_hat, fmag = -kap (b_hat . grad T)
  const double lfm = lfx * bx + lfy * by;
  const double lbx = lfx * fmag + lfm * (-kap * gx), lby = lfy * fmag + lfm * (-kap * gy);
  const double lkap = -lfm * bdg;
  if (capped) lr += lkap * kmax / r;
  else lT += lkap * p.k0 * 2.5 * pow_1p5(T);
  const double lgx = -lfm * kap * bx, lgy = -lfm * kap * by;
  if (lgx != 0.0) scatter_deriv(M, 0, i, j, lgx, lam_T);
  if (lgy != 0.0) scatter_deriv(M, 1, i, j, lgy, lam_T);
  if (lT != 0.0) atomicAdd(&lam_T[c], lT);
  lam_fz[c] += lbx;
  lam_fz[N + c] += lby;
  lam_rho[c] += lr;
}

__global__ void __launch_bounds__(256) k_adj_bhat_derivs(Mesh M, double* lam_fz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny || !inside(M, i, j)) return;
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  double* lbx = lam_fz;
  double* lby = lam_fz + N;
  const double lxx = lam_fz[2 * N + c], lxy = lam_fz[3 * N + c], lyx = lam_fz[4 * N + c], lyy = lam_fz[5 * N + c];
  if (lxx != 0.0) scatter_deriv(M, 0, i, j, lxx, lbx);
  if (lxy != 0.0) scatter_deriv(M, 1, i, j, lxy, lbx);
  if (lyx != 0.0) scatter_deriv(M, 0, i, j, lyx, lby);
  if (lyy != 0.0) scatter_deriv(M, 1, i, j, lyy, lby);
}

__global__ void __launch_bounds__(256) k_adj_bhat_field(MhdSet L, MhdStatic st, Mesh M, const double* lam_fz,
                                                        double* lam) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  const double gx = lam_fz[c], gy = lam_fz[N + c];
  if (gx == 0.0 && gy == 0.0) return;
  const double bx = st.bex[c] + L.bx[c], by = st.bey[c] + L.by[c], bz = st.bez[c] + L.bz[c];
  const double bm = sqrt(bx * bx + by * by + bz * bz);
  if (bm == 0.0) return;
  const double hx = bx / bm, hy = by / bm, hz = bz / bm;
  // d b_hat_a / d b_b = (delta_ab - b_hat_a b_hat_b) / |b|
  const double s = gx * hx + gy * hy;
  lam[5 * N + c] += (gx - s * hx) / bm;
  lam[6 * N + c] += (gy - s * hy) / bm;
  lam[7 * N + c] += (-s * hz) / bm;
}

// ---- radiative losses --------------------------------------------------------------------------

template <int K>
__device__ Dn<K> rl_loss_dn(const Dn<K>& t, const Dn<K>& n, const Dn<K>& te_live, const Dn<K>& dt_live,
                           const RlParams& p) {
  if (t.v < p.cut) return dn_const<K>(0.0);
  double chi;
  Dn<K> pw;
  if (t.v <= p.tb[0]) { chi = 1.09e-31; pw = t * t; }
  else if (t.v <= p.tb[1]) { chi = 8.87e-17; pw = 1.0 / t; }
  else if (t.v <= p.tb[2]) { chi = 1.90e-22; pw = dn_const<K>(1.0); }
  else if (t.v <= p.tb[3]) { chi = 3.53e-13; pw = dn_pow_m1p5(t); }
  else if (t.v <= p.tb[4]) { chi = 3.46e-25; pw = dn_pow_third(t); }
  else if (t.v <= p.tb[5]) { chi = 5.49e-16; pw = 1.0 / t; }
  else { chi = 1.96e-27; pw = dn_sqrt(t); }
  Dn<K> v = n * n * chi * pw;
  if (t.v < p.cut_plus_ramp) v = v * ((t - p.cut) / p.ramp);
  // prevent_subcycling: the cap by the live thermal energy over the live dt grid (:148-152)
  if (p.prevent && 0.1 * p.eps * (te_live.v / v.v) < p.pd_eps * dt_live.v)
    v = 0.1 * p.eps * te_live / (p.pd_eps * dt_live);
  return v;
}

// The subcycles of one cell in a 4-component forward dual: the cell's thermal energy (which is also
// the live energy prevent_subcycling reads), n, the subcycle step and, with prevent_subcycling,
// the cell's live dt, whose own gradient (dt_cell_dual, of the input state A) carries that part on.
__global__ void __launch_bounds__(256) k_adj_rl(MhdSet A, MhdStatic st, RlParams p, MhdConst K, Mesh M, int nsub,
                                                double h, int integ, double* lam, double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    // the live propagate after the update: rho floors on A's density, te floor on the output
    const double x = A.rho[c] / K.m;
    const double rho1 = smax(x, K.nmin) * K.m;
    const double drho = (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
    double lrho = lam[c] * drho;
    const double te0 = A.te[c];
    if (!inside(M, i, j)) {
      lam[4 * N + c] *= (te0 < K.temin) ? 0.0 : 1.0;
    } else {
      const Dn<4> nn = dn_var<4>(A.n[c], 1), hv = dn_var<4>(h, 2);
      Dn<4> te = dn_var<4>(te0, 0);
      const Dn<4> te_live = te;
      Dual8 dtd = dconst(0.0);
      Dn<4> dtl = dn_const<4>(0.0);
      if (p.prevent) {
        dtd = dt_cell_dual(A, st, K, M, i, j);
        dtl = dn_var<4>(dtd.v, 3);
      }
      const Dn<4> hh = 1.0 * hv, half = 1.0 * (0.5 * hv);
      auto loss = [&](const Dn<4>& e) {
        const Dn<4> T = dn_max(e * K.gm1 / (nn * (2.0 * K_B)), K.tmin);
        return rl_loss_dn(T, nn, te_live, dtl, p);
      };
      for (int s = 0; s < nsub; ++s) {
        if (integ == 0) {
          te = dn_max(te - hh * loss(te), K.temin);
        } else if (integ == 1) {
          const Dn<4> hs = dn_max(te - half * loss(te), K.temin);
          te = dn_max(te - hh * loss(hs), K.temin);
        } else {
          const Dn<4> k1 = -loss(te);
          Dn<4> im = dn_max(te + half * k1, K.temin);
          const Dn<4> k2 = -loss(im);
          im = dn_max(te + half * k2, K.temin);
          const Dn<4> k3 = -loss(im);
          im = dn_max(te + hh * k3, K.temin);
          const Dn<4> k4 = -loss(im);
          te = dn_max(te + (hh * (k1 + 2.0 * k2 + 2.0 * k3 + k4)) / 6.0, K.temin);
        }
      }
      const double lo = (te.v < K.temin) ? 0.0 : lam[4 * N + c];
      lam[4 * N + c] = lo * te.d[0];
      lrho += lo * te.d[1] * K.rm;
      dot = lo * te.d[2];
      if (p.prevent) {
        const double ldt = lo * te.d[3];
        lrho += ldt * dtd.d[AV_RHO];
        for (int k = 1; k < 8; ++k) lam[(size_t)k * N + c] += ldt * dtd.d[k];
      }
    }
    lam[c] = lrho;
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
}

// ---- source-term modules after the integrator ------------------------------------------------

// Reverse of one source module's post-integrator update and the live propagate's floors
// (sources.hip k_source; localizedheating.cpp:52-66, massinjection.cpp:49-53,
// momentuminjection.cpp:70-76, ambientheatingsink.cpp:35-37), P the state before it.  The added
// term is mask * ((h * g) * t) [* m_i | * rho] for ADD / ADD_SCALED / ADD_WEIGHTED (the last on
// both momenta, t and t2), or -(*step) * t for SUB; g depends on the time at the step (dg/dtime =
// gp).  lam is updated in place; ph / pt: block sums of dJ/dh and dJ/dtime.
__global__ void __launch_bounds__(256) k_adj_source(MhdSet P, int mode, const double* t, const double* t2,
                                                    const double* step, double g, double gp, MhdConst K, Mesh M,
                                                    double* lam, double* ph, double* pt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dh = 0.0, dtm = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double mask = inside(M, i, j) ? 1.0 : 0.0;
    const double h = *step;
    double rho_in = P.rho[c], te_in = P.te[c];
    if (mode == SRC_ADD) te_in = P.te[c] + mask * ((h * g) * t[c]);
    else if (mode == SRC_ADD_SCALED) rho_in = P.rho[c] + mask * (((h * g) * t[c]) * K.m);
    else if (mode == SRC_SUB) te_in = P.te[c] - h * t[c];
    const double x = rho_in / K.m;
    const double rho1 = smax(x, K.nmin) * K.m;
    const double drho = (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
    const double dte = te_in < K.temin ? 0.0 : 1.0;
    const double lr = lam[c] * drho, lt = lam[4 * N + c] * dte;
    lam[c] = lr;
    lam[4 * N + c] = lt;
    if (mode == SRC_ADD) {
      dh = lt * mask * (g * t[c]);
      dtm = lt * mask * ((h * gp) * t[c]);
    } else if (mode == SRC_ADD_SCALED) {
      dh = lr * mask * ((g * t[c]) * K.m);
      dtm = lr * mask * (((h * gp) * t[c]) * K.m);
    } else if (mode == SRC_ADD_WEIGHTED) {
      const double lmx = lam[N + c], lmy = lam[2 * N + c], rho = P.rho[c];
      const double sx = lmx * t[c] + lmy * t2[c];
      lam[c] = lr + mask * ((h * g) * sx);
      dh = mask * (g * sx) * rho;
      dtm = mask * ((h * gp) * sx) * rho;
    } else {
      dh = -lt * t[c];
    }
  }
  dh = block_sum(dh);
  dtm = block_sum(dtm);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    ph[b] = dh;
    pt[b] = dtm;
  }
}

// ---- parametrised heating ----------------------------------------------------------------------

// interior column sums of lam_te: 64 columns per block, 16 rows in flight per column
__global__ void __launch_bounds__(1024) k_adj_colsum(const double* lam_te, Mesh M, double* colsum) {
  __shared__ double part[16][64];
  const int j = blockIdx.x * 64 + threadIdx.x;
  double s = 0.0;
  if (j < M.ny && j >= M.y.lo && j <= M.y.hi)
    for (int i = M.x.lo + (int)threadIdx.y; i <= M.x.hi; i += 16) s += lam_te[(size_t)i * M.ny + j];
  part[threadIdx.y][threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.y == 0 && j < M.ny) {
    double t = 0.0;
    for (int r = 0; r < 16; ++r) t += part[r][threadIdx.x];
    colsum[j] = t;
  }
}

__global__ void __launch_bounds__(256) k_adj_theta(const double* colsum, const double* basis, const double* step,
                                                   Mesh M, double* g_theta) {
  const int k = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < M.ny; j += blockDim.x) s += basis[(size_t)k * M.ny + j] * colsum[j];
  s = block_sum(s);
  if (threadIdx.x == 0) g_theta[k] += (*step) * s;
}

__global__ void __launch_bounds__(256) k_adj_accum(const double* partial, int n, double scale, double* gh) {
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) s += partial[k];
  s = block_sum(s);
  if (threadIdx.x == 0) gh[0] += scale * s;
}

// ---- divergence cleaning ------------------------------------------------------------------------

// weights of derivative1D (d) and secondDerivative1D (s2) at cell c of an axis with respect to the
// values at c-1, c, c+1 (derivs.cpp:229-264, 450-462, 477-487)
struct Wm { double m, c, p; };
__device__ __forceinline__ Wm dw(const Axis& A, int c, int cm, int cp) {
  const double up_c = A.h[cp] / A.fden[cp], up_p = A.h[c] / A.fden[cp];
  const double lo_m = A.h[c] / A.fden[c], lo_c = A.h[cm] / A.fden[c];
  return Wm{-lo_m / A.d[c], (up_c - lo_c) / A.d[c], up_p / A.d[c]};
}
__device__ __forceinline__ Wm s2w(const Axis& A, int c, int cm, int cp) {
  const double up_c = A.h[cp] / A.fden[cp], up_p = A.h[c] / A.fden[cp];
  const double lo_m = A.h[c] / A.fden[c], lo_c = A.h[cm] / A.fden[c];
  return Wm{lo_m / A.hsq[c], (up_c - 2.0 + lo_c) / A.hsq[c], up_p / A.hsq[c]};
}

// One divergence-cleaning subcycle (divcleaning.cpp:28-46), linear in b: bx += mask*h*coeff*
// (d/dx d(by)/dy + d2(bx)/dx2), by += mask*h*coeff*(d/dy d(bx)/dx + d2(by)/dy2), the inner
// derivatives zero off the interior.  Reversed in two passes: this one takes g = mask*h*coeff*lam
// of each interior cell through the second derivatives onto ax, ay and through the outer first
// derivatives onto ux, uy (interior targets only: the inner derivatives vanish elsewhere), and sums
// the step-size partial lam . mask*coeff*(the stencil at the subcycle's b, bx_s, by_s).
__global__ void __launch_bounds__(256) k_adj_dc_a(const double* bx_s, const double* by_s, const double* dxb,
                                                  const double* dyb, const double* coeff, double h, Mesh M,
                                                  const double* lbx, const double* lby, double* ax, double* ay,
                                                  double* ux, double* uy, double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dot = 0.0;
  if (i < M.nx && j < M.ny && inside(M, i, j)) {
    const size_t c = (size_t)i * M.ny + j;
    const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
    const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
    const size_t cxm = (size_t)im * M.ny + j, cxp = (size_t)ip * M.ny + j;
    const size_t cym = (size_t)i * M.ny + jm, cyp = (size_t)i * M.ny + jp;
    const Wm wx = dw(M.x, i, im, ip), wy = dw(M.y, j, jm, jp), sx = s2w(M.x, i, im, ip), sy = s2w(M.y, j, jm, jp);
    const double lx = lbx[c], ly = lby[c];
    // the subcycle's stencil values (as k_dc_update forms them) for the step-size partial
    const double stx = (wx.m * dyb[cxm] + wx.c * dyb[c] + wx.p * dyb[cxp]) +
                       (sx.m * bx_s[cxm] + sx.c * bx_s[c] + sx.p * bx_s[cxp]);
    const double sty = (wy.m * dxb[cym] + wy.c * dxb[c] + wy.p * dxb[cyp]) +
                       (sy.m * by_s[cym] + sy.c * by_s[c] + sy.p * by_s[cyp]);
    dot = lx * coeff[c] * stx + ly * coeff[c] * sty;
    const double gx = h * coeff[c] * lx, gy = h * coeff[c] * ly;
    if (gx != 0.0) {
      atomicAdd(&ax[cxm], gx * sx.m);
      atomicAdd(&ax[c], gx * sx.c);
      atomicAdd(&ax[cxp], gx * sx.p);
      if (inside(M, im, j)) atomicAdd(&uy[cxm], gx * wx.m);
      atomicAdd(&uy[c], gx * wx.c);
      if (inside(M, ip, j)) atomicAdd(&uy[cxp], gx * wx.p);
    }
    if (gy != 0.0) {
      atomicAdd(&ay[cym], gy * sy.m);
      atomicAdd(&ay[c], gy * sy.c);
      atomicAdd(&ay[cyp], gy * sy.p);
      if (inside(M, i, jm)) atomicAdd(&ux[cym], gy * wy.m);
      atomicAdd(&ux[c], gy * wy.c);
      if (inside(M, i, jp)) atomicAdd(&ux[cyp], gy * wy.p);
    }
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
}

// second pass: the inner derivatives transposed (uy through d/dy onto ay, ux through d/dx onto ax)
__global__ void __launch_bounds__(256) k_adj_dc_b(Mesh M, double* ux, double* uy, double* ax, double* ay) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double u1 = ux[c], u2 = uy[c];
  ux[c] = 0.0;
  uy[c] = 0.0;
  if (!inside(M, i, j)) return;
  const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
  const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
  if (u1 != 0.0) {   // ux = adjoint of d(bx)/dx at c
    const Wm w = dw(M.x, i, im, ip);
    atomicAdd(&ax[(size_t)im * M.ny + j], u1 * w.m);
    atomicAdd(&ax[c], u1 * w.c);
    atomicAdd(&ax[(size_t)ip * M.ny + j], u1 * w.p);
  }
  if (u2 != 0.0) {   // uy = adjoint of d(by)/dy at c
    const Wm w = dw(M.y, j, jm, jp);
    atomicAdd(&ay[(size_t)i * M.ny + jm], u2 * w.m);
    atomicAdd(&ay[c], u2 * w.c);
    atomicAdd(&ay[(size_t)i * M.ny + jp], u2 * w.p);
  }
}

// lam_bx += ax, lam_by += ay; ax, ay cleared
__global__ void __launch_bounds__(256) k_adj_dc_c(size_t N, double* ax, double* ay, double* lbx, double* lby) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  lbx[c] += ax[c];
  lby[c] += ay[c];
  ax[c] = 0.0;
  ay[c] = 0.0;
}

// ---- boundary outflow -------------------------------------------------------------------------

// b_hat_x, b_hat_y = b/|b| (b = be + bi, zero where |b| = 0) of a cell and the map of their adjoint
// (gx, gy) onto bi_x, bi_y, bi_z of lam
__device__ __forceinline__ void of_bhat_rev(const MhdSet& P, const MhdStatic& st, size_t c, size_t N, double gx,
                                            double gy, double* lam) {
  const double b[3] = {st.bex[c] + P.bx[c], st.bey[c] + P.by[c], st.bez[c] + P.bz[c]};
  const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  if (bm == 0.0) return;
  const double gb = (gx * b[0] + gy * b[1]) / (bm * bm * bm);
  lam[5 * N + c] += gx / bm - gb * b[0];
  lam[6 * N + c] += gy / bm - gb * b[1];
  lam[7 * N + c] -= gb * b[2];
}

// BoundaryOutflow::postIterateModule (boundaryoutflow.cpp:39-67) reversed on its input state P:
// mom += step*(a*tmpl)*(b_hat or the boundary normal)*rho.  lam updated in place (density, b_hat
// onto bi); pa / ph: block sums of the adjoint of the acceleration a and of the step size.
__global__ void __launch_bounds__(256) k_adj_outflow(MhdSet P, MhdStatic st, const double* tmpl, const double* sc,
                                                     const double* step, OutflowParams p, Mesh M, double* lam,
                                                     double* pa, double* ph) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double da = 0.0, dh = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double a = sc[1], h = *step, t = tmpl[c];
    const double A = h * (a * t);
    const double rho = P.rho[c], lmx = lam[N + c], lmy = lam[2 * N + c];
    double lA = 0.0;
    if (p.aligned) {
      const double b[3] = {st.bex[c] + P.bx[c], st.bey[c] + P.by[c], st.bez[c] + P.bz[c]};
      const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      const double hx = bm == 0.0 ? 0.0 : b[0] / bm, hy = bm == 0.0 ? 0.0 : b[1] / bm;
      lam[c] += lmx * (A * hx) + lmy * (A * hy);
      lA = lmx * hx * rho + lmy * hy * rho;
      if (A != 0.0) of_bhat_rev(P, st, c, N, lmx * A * rho, lmy * A * rho, lam);
    } else {
      const double lm = (p.side == OF_X1 || p.side == OF_X2) ? lmx : lmy;
      lam[c] += lm * A;
      lA = lm * rho;
    }
    da = lA * h * t;
    dh = lA * a * t;
  }
  da = block_sum(da);
  dh = block_sum(dh);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    pa[b] = da;
    ph[b] = dh;
  }
}

// dynamic_mode (:41-45): a = clamp((target - m)/dynamic_time, 0, max_accel) with m the largest
// oriented outflow speed b_hat . v over the window (launch_outflow_max's block results pv, pk);
// where the clamp does not bind, the acceleration's adjoint reaches the cell that sets m.  One block.
__global__ void __launch_bounds__(256) k_adj_outflow_dyn(MhdSet P, MhdStatic st, OutflowParams p, Mesh M,
                                                         const double* pv, const double* pk, int nb, const double* pa,
                                                         int npa, double* lam) {
  __shared__ double sv[256], sa[256];
  __shared__ long long sk[256];
  double v = -INFINITY, ga = 0.0;
  long long k = -1;
  auto pick = [](double& v0, long long& k0, double ov, long long ok) {
    if (ok >= 0 && (ov > v0 || (ov == v0 && (k0 < 0 || ok < k0)))) { v0 = ov; k0 = ok; }
  };
  for (int b = threadIdx.x; b < nb; b += blockDim.x) pick(v, k, pv[b], (long long)pk[b]);
  for (int b = threadIdx.x; b < npa; b += blockDim.x) ga += pa[b];
  sv[threadIdx.x] = v;
  sk[threadIdx.x] = k;
  sa[threadIdx.x] = ga;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      pick(sv[threadIdx.x], sk[threadIdx.x], sv[threadIdx.x + off], sk[threadIdx.x + off]);
      sa[threadIdx.x] += sa[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double init = -1.0 * p.target;
  if (!(sk[0] >= 0 && init < sv[0])) return;          // m is the fold's start value: no cell sets it
  const double m = sv[0];
  const double a = (p.target - m) / p.dtime;
  if (!(a < p.amax) || a < 0.0) return;               // the clamp binds
  const double lmv = -sa[0] / p.dtime;                // d a / d m = -1/dynamic_time
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)sk[0];
  const double b[3] = {st.bex[c] + P.bx[c], st.bey[c] + P.by[c], st.bez[c] + P.bz[c]};
  const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double hx = bm == 0.0 ? 0.0 : b[0] / bm, hy = bm == 0.0 ? 0.0 : b[1] / bm;
  double sg = 1.0;
  if (p.side == OF_X1 && hx > 0.0) sg = -1.0;
  else if (p.side == OF_X2 && hx < 0.0) sg = -1.0;
  else if (p.side == OF_Y2 && hy < 0.0) sg = -1.0;
  else if (p.side == OF_Y1 && hy > 0.0) sg = -1.0;
  const double rho = P.rho[c], vx = P.mx[c] / rho, vy = P.my[c] / rho;
  const double g = lmv * sg;
  of_bhat_rev(P, st, c, N, g * vx, g * vy, lam);
  lam[N + c] += g * hx / rho;
  lam[2 * N + c] += g * hy / rho;
  lam[c] -= g * (hx * vx + hy * vy) / rho;
}

// ---- field heating ---------------------------------------------------------------------------

// FieldHeating (fieldheating.cpp:34-61) reversed: the heating H is taken from the step's start state
// Ln (b = be + bi, n, and b_hat of Ln), te += mask*(h*H) is applied to the module's input state P and
// propagated.  lam (the propagated state's adjoint) goes through the propagate's floors in place;
// H's adjoint lam_te*mask*h is carried to Ln: the current through derivative1D transposed onto bi_y
// (x neighbours) and bi_x (y neighbours), |b| onto bi of the cell, n onto rho, and the curvature of
// b_hat onto lfz (b_hat and its four derivatives at the cell, for launch_adj_bhat) -- all into lamLn
// (atomics).  partial: block sums of lam_te*mask*H (the step-size adjoint).
__global__ void __launch_bounds__(256) k_adj_fheat(MhdSet Ln, MhdSet P, MhdStatic st, const double* bhx,
                                                   const double* bhy, FheatParams p, const double* step, MhdConst K,
                                                   Mesh M, double* lam, double* lamLn, double* lfz, double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const bool in = inside(M, i, j);
    const double h = *step;
    // the heating as the forward forms it
    double H = p.coeff, curl = 0.0, A = 1.0, B = 1.0, C = 1.0, D = 1.0, bm = 0.0, kx = 0.0, ky = 0.0, km = 0.0;
    double dbx_x = 0.0, dbx_y = 0.0, dby_x = 0.0, dby_y = 0.0;
    const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
    const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
    const size_t cxm = (size_t)im * M.ny + j, cxp = (size_t)ip * M.ny + j;
    const size_t cym = (size_t)i * M.ny + jm, cyp = (size_t)i * M.ny + jp;
    auto dx_of = [&](double qm, double qc, double qp) {
      return qdiv(tcc::fint(M.x, ip, i, qc, qp) - tcc::fint(M.x, i, im, qm, qc), M.x.d[i], M.x.rd[i]);
    };
    auto dy_of = [&](double qm, double qc, double qp) {
      return qdiv(tcc::fint(M.y, jp, j, qc, qp) - tcc::fint(M.y, j, jm, qm, qc), M.y.d[j], M.y.rd[j]);
    };
    const double b[3] = {st.bex[c] + Ln.bx[c], st.bey[c] + Ln.by[c], st.bez[c] + Ln.bz[c]};
    if (p.coeff != 0.0 && in) {
      if (p.cpow != 0.0) {
        curl = dx_of(st.bey[cxm] + Ln.by[cxm], b[1], st.bey[cxp] + Ln.by[cxp]) -
               dy_of(st.bex[cym] + Ln.bx[cym], b[0], st.bex[cyp] + Ln.bx[cyp]);
        A = pow(p.cc * fabs(curl), p.cpow);
      }
      if (p.bpow != 0.0) {
        bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        B = pow(bm, p.bpow);
      }
      if (p.npow != 0.0) C = pow(Ln.n[c], p.npow);
      if (p.rpow != 0.0) {
        dbx_x = dx_of(bhx[cxm], bhx[c], bhx[cxp]);
        dbx_y = dy_of(bhx[cym], bhx[c], bhx[cyp]);
        dby_x = dx_of(bhy[cxm], bhy[c], bhy[cxp]);
        dby_y = dy_of(bhy[cym], bhy[c], bhy[cyp]);
        kx = bhx[c] * dbx_x + bhy[c] * dbx_y;
        ky = bhx[c] * dby_x + bhy[c] * dby_y;
        km = sqrt(kx * kx + ky * ky);
        D = pow(smax(km, 1.0e-16), p.rpow);          // h / roc^rpow = h * max(|k|, 1e-16)^rpow
      }
      H = p.coeff * A * B * C * D;
    } else {
      H = 0.0;
    }
    // the update te + mask*(h*H) and the propagate's floors
    const double mask = in ? 1.0 : 0.0;
    const double te_in = P.te[c] + mask * (h * H);
    const double x = P.rho[c] / K.m;
    const double rho1 = smax(x, K.nmin) * K.m;
    lam[c] *= (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
    const double lt = lam[4 * N + c] * (te_in < K.temin ? 0.0 : 1.0);
    lam[4 * N + c] = lt;
    dot = lt * mask * H;
    const double lH = lt * mask * h;
    if (lH != 0.0 && H != 0.0) {
      if (p.cpow != 0.0 && curl != 0.0) {
        const double lc = lH * H * p.cpow / curl;          // d|c|^p/dc = p |c|^p / c
        scatter_deriv(M, 0, i, j, lc, lamLn + 6 * N);      // curl = d(by)/dx - d(bx)/dy
        scatter_deriv(M, 1, i, j, -lc, lamLn + 5 * N);
      }
      if (p.bpow != 0.0 && bm != 0.0) {
        const double lb = lH * H * p.bpow / (bm * bm);
        for (int k = 0; k < 3; ++k) atomicAdd(&lamLn[(5 + k) * N + c], lb * b[k]);
      }
      if (p.npow != 0.0) atomicAdd(&lamLn[c], lH * H * p.npow / Ln.n[c] * K.rm);
      if (p.rpow != 0.0 && km > 1.0e-16) {
        const double lk = lH * H * p.rpow / (km * km);     // d|k|^p/dk = p |k|^p k / |k|^2
        const double lkx = lk * kx, lky = lk * ky;
        lfz[c] += lkx * dbx_x + lky * dby_x;                // b_hat_x
        lfz[N + c] += lkx * dbx_y + lky * dby_y;            // b_hat_y
        lfz[2 * N + c] += lkx * bhx[c];                     // d b_hat_x / dx
        lfz[3 * N + c] += lkx * bhy[c];                     // d b_hat_x / dy
        lfz[4 * N + c] += lky * bhx[c];                     // d b_hat_y / dx
        lfz[5 * N + c] += lky * bhy[c];                     // d b_hat_y / dy
      }
    }
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
}

// ---- artificial viscosity ----------------------------------------------------------------------

// the differentiated variable at one cell, as the forward term reads it (idealmhd.cpp:241-277)
__device__ double visc_var(const MhdSet& S, const MhdStatic& st, const MhdConst& K, int var, size_t c) {
  switch (var) {
    case VV_RHO: return S.rho[c];
    case VV_N: return S.n[c];
    case VV_MX: return S.mx[c];
    case VV_MY: return S.my[c];
    case VV_MZ: return S.mz[c];
    case VV_TE: return S.te[c];
    case VV_BIX: return S.bx[c];
    case VV_BIY: return S.by[c];
    case VV_BIZ: return S.bz[c];
    case VV_GX: return st.gx[c];
    case VV_GY: return st.gy[c];
    case VV_TEMP: return smax((S.te[c] * K.gm1) / (S.n[c] * (2.0 * K_B)), K.tmin);
    case VV_PRESS: return S.te[c] * K.gm1;
    case VV_VX: return S.mx[c] / S.rho[c];
    case VV_VY: return S.my[c] / S.rho[c];
    case VV_VZ: return S.mz[c] / S.rho[c];
    case VV_KE: {
      const double r = S.rho[c], vx = S.mx[c] / r, vy = S.my[c] / r;
      return r * 0.5 * (vx * vx + vy * vy);
    }
    case VV_BX: return st.bex[c] + S.bx[c];
    case VV_BY: return st.bey[c] + S.by[c];
    case VV_BZ: return st.bez[c] + S.bz[c];
    default: {
      const double bx = st.bex[c] + S.bx[c], by = st.bey[c] + S.by[c], bz = st.bez[c] + S.bz[c];
      const double bm = sqrt(bx * bx + by * by + bz * bz);
      if (var == VV_BMAG) return bm;
      if (bm == 0.0) return 0.0;
      return var == VV_BHX ? bx / bm : (var == VV_BHY ? by / bm : bz / bm);
    }
  }
}

// nu at one cell (viscosity.cpp:212): ((s * 1) / (1/d_x^2 + 1/d_y^2)) / 2 / dt
__device__ __forceinline__ double visc_nu(const ViscTermArgs& a, int i, int j, size_t c, double& dt) {
  const double dx = a.mesh.x.d[i], dy = a.mesh.y.d[j];
  const double st = a.strength ? a.strength[c] : a.strength_c;
  dt = a.dtgrid ? a.dtgrid[c] : *a.dtmin;
  return st / (1.0 / (dx * dx) + 1.0 / (dy * dy)) / 2. / dt;
}

__global__ void __launch_bounds__(256) k_adj_visc_term(ViscTermArgs a, int evol, const double* ka, double sa,
                                                       const double* kb, double* lq, double* lf, double* lam_S,
                                                       double* ldt, double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const Mesh& M = a.mesh;
  double dot = 0.0;
  if (i < M.nx && j < M.ny && inside(M, i, j)) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double lo = sa * ka[evol * N + c] + (kb ? kb[evol * N + c] : 0.0);
    if (lo != 0.0) {
      const Axis& X = M.x;
      const Axis& Y = M.y;
      double dt;
      const double nu = visc_nu(a, i, j, c, dt);
      const double sc = a.scale == 0 ? 1.0 : a.S.n[c] * a.scale_c;
      const int im = wrp(i - 1, M.nx, X.per), ip = wrp(i + 1, M.nx, X.per);
      const int jm = wrp(j - 1, M.ny, Y.per), jp = wrp(j + 1, M.ny, Y.per);
      const size_t cxm = (size_t)im * M.ny + j, cxp = (size_t)ip * M.ny + j;
      const size_t cym = (size_t)i * M.ny + jm, cyp = (size_t)i * M.ny + jp;
      const double q = visc_var(a.S, a.st, a.k, a.var, c);
      const double qxm = visc_var(a.S, a.st, a.k, a.var, cxm), qxp = visc_var(a.S, a.st, a.k, a.var, cxp);
      const double qym = visc_var(a.S, a.st, a.k, a.var, cym), qyp = visc_var(a.S, a.st, a.k, a.var, cyp);
      // second derivatives through the face interpolants (derivs.cpp:450-462, 477-487)
      const double xu_c = X.h[ip] / X.fden[ip], xu_p = X.h[i] / X.fden[ip];
      const double xl_m = X.h[i] / X.fden[i], xl_c = X.h[im] / X.fden[i];
      const double yu_c = Y.h[jp] / Y.fden[jp], yu_p = Y.h[j] / Y.fden[jp];
      const double yl_m = Y.h[j] / Y.fden[j], yl_c = Y.h[jm] / Y.fden[j];
      const double sx = ((q * xu_c + qxp * xu_p) - 2.0 * q + (qxm * xl_m + q * xl_c)) / X.hsq[i];
      const double sy = ((q * yu_c + qyp * yu_p) - 2.0 * q + (qym * yl_m + q * yl_c)) / Y.hsq[j];
      const double lap = sx + sy;
      // r = nu * lap * sc: the Laplacian's stencil
      const double ll = lo * nu * sc;
      const double rx = ll / X.hsq[i], ry = ll / Y.hsq[j];
      atomicAdd(&lq[c], rx * (xu_c - 2.0 + xl_c) + ry * (yu_c - 2.0 + yl_c));
      atomicAdd(&lq[cxp], rx * xu_p);
      atomicAdd(&lq[cxm], rx * xl_m);
      atomicAdd(&lq[cyp], ry * yu_p);
      atomicAdd(&lq[cym], ry * yl_m);
      // the scale n * c (n = rho / m_i)
      if (a.scale != 0) lam_S[c] += lo * nu * lap * a.scale_c * a.k.rm;
      // the coefficient's timescale: d nu / d dt = -nu / dt
      const double ld = -(lo * lap * sc) * nu / dt;
      if (a.dtgrid) ldt[c] += ld;
      else dot = ld;
      if (a.grad) {
        // + d(nu*sc)/dx * dq/dx + d(nu*sc)/dy * dq/dy (derivative1D of both, derivs.cpp:223-264)
        auto fof = [&](int ii, int jj) {
          const size_t cc = (size_t)ii * M.ny + jj;
          double t;
          return visc_nu(a, ii, jj, cc, t) * (a.scale == 0 ? 1.0 : a.S.n[cc] * a.scale_c);
        };
        const double f = nu * sc, fxm = fof(im, j), fxp = fof(ip, j), fym = fof(i, jm), fyp = fof(i, jp);
        const double gx = ((q * xu_c + qxp * xu_p) - (qxm * xl_m + q * xl_c)) / X.d[i];
        const double gy = ((q * yu_c + qyp * yu_p) - (qym * yl_m + q * yl_c)) / Y.d[j];
        const double fx = ((f * xu_c + fxp * xu_p) - (fxm * xl_m + f * xl_c)) / X.d[i];
        const double fy = ((f * yu_c + fyp * yu_p) - (fym * yl_m + f * yl_c)) / Y.d[j];
        // derivative1D transposed onto the three cells of each axis, for q (seed fx, fy) and f (gx, gy)
        auto spread = [&](double* to, double sx_, double sy_) {
          const double ax = sx_ / X.d[i], ay = sy_ / Y.d[j];
          atomicAdd(&to[c], ax * (xu_c - xl_c) + ay * (yu_c - yl_c));
          atomicAdd(&to[cxp], ax * xu_p);
          atomicAdd(&to[cxm], -ax * xl_m);
          atomicAdd(&to[cyp], ay * yu_p);
          atomicAdd(&to[cym], -ay * yl_m);
        };
        spread(lq, lo * fx, lo * fy);
        spread(lf, lo * gx, lo * gy);
      }
    }
  }
  if (partial) {
    dot = block_sum(dot);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
  }
}

// the gradient correction's adjoint of f = nu * sc on every cell: the scale onto the density, nu
// onto its timescale (per cell, or block sums for a global minimum); lf cleared
__global__ void __launch_bounds__(256) k_adj_visc_f(ViscTermArgs a, double* lf, double* lam_S, double* ldt,
                                                    double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const Mesh& M = a.mesh;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t c = (size_t)i * M.ny + j;
    const double l = lf[c];
    if (l != 0.0) {
      lf[c] = 0.0;
      double dt;
      const double nu = visc_nu(a, i, j, c, dt);
      const double sc = a.scale == 0 ? 1.0 : a.S.n[c] * a.scale_c;
      if (a.scale != 0) lam_S[c] += l * nu * a.scale_c * a.k.rm;
      const double ld = -(l * sc) * nu / dt;
      if (a.dtgrid) ldt[c] += ld;
      else dot = ld;
    }
  }
  if (partial) {
    dot = block_sum(dot);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
  }
}

__global__ void __launch_bounds__(256) k_adj_visc_vars(ViscTermArgs a, double* lq, double* lam_S) {
  const size_t N = (size_t)a.mesh.nx * a.mesh.ny;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double l = lq[c];
  if (l == 0.0) return;
  lq[c] = 0.0;
  const MhdSet& S = a.S;
  const MhdConst& K = a.k;
  double* lr = lam_S;
  double* lmx = lam_S + N;
  double* lmy = lam_S + 2 * N;
  double* lmz = lam_S + 3 * N;
  double* lte = lam_S + 4 * N;
  double* lb[3] = {lam_S + 5 * N, lam_S + 6 * N, lam_S + 7 * N};
  switch (a.var) {
    case VV_RHO: lr[c] += l; break;
    case VV_N: lr[c] += l * K.rm; break;
    case VV_MX: lmx[c] += l; break;
    case VV_MY: lmy[c] += l; break;
    case VV_MZ: lmz[c] += l; break;
    case VV_TE: lte[c] += l; break;
    case VV_BIX: case VV_BX: lb[0][c] += l; break;
    case VV_BIY: case VV_BY: lb[1][c] += l; break;
    case VV_BIZ: case VV_BZ: lb[2][c] += l; break;
    case VV_GX: case VV_GY: break;
    case VV_TEMP: {
      const double n = S.n[c], den = n * (2.0 * K_B);
      const double x = (S.te[c] * K.gm1) / den;
      if (x < K.tmin) break;
      lte[c] += l * (K.gm1 / den);
      lr[c] += l * (-x / n) * K.rm;
      break;
    }
    case VV_PRESS: lte[c] += l * K.gm1; break;
    case VV_VX: case VV_VY: case VV_VZ: {
      const double r = S.rho[c];
      const double m = a.var == VV_VX ? S.mx[c] : (a.var == VV_VY ? S.my[c] : S.mz[c]);
      double* lm = a.var == VV_VX ? lmx : (a.var == VV_VY ? lmy : lmz);
      lm[c] += l / r;
      lr[c] -= l * m / (r * r);
      break;
    }
    case VV_KE: {
      const double r = S.rho[c], vx = S.mx[c] / r, vy = S.my[c] / r;
      lmx[c] += l * vx;
      lmy[c] += l * vy;
      lr[c] -= l * 0.5 * (vx * vx + vy * vy);
      break;
    }
    default: {   // |b| and b_hat
      const double b[3] = {a.st.bex[c] + S.bx[c], a.st.bey[c] + S.by[c], a.st.bez[c] + S.bz[c]};
      const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      if (bm == 0.0) break;
      if (a.var == VV_BMAG) {
        for (int k = 0; k < 3; ++k) lb[k][c] += l * b[k] / bm;
      } else {
        const int w = a.var == VV_BHX ? 0 : (a.var == VV_BHY ? 1 : 2);
        const double bm3 = bm * bm * bm;
        for (int k = 0; k < 3; ++k) lb[k][c] += l * ((k == w ? 1.0 / bm : 0.0) - b[w] * b[k] / bm3);
      }
    }
  }
}

struct HvTargets {
  double* ld[4];
  double w[4];
  int n;
};

__global__ void __launch_bounds__(256) k_adj_hv_update(MhdSet P, int g, const double* init, const double* D, double cf,
                                                       double hh, MhdConst K, Mesh M, double* lam, double* lam_init,
                                                       HvTargets tg, double* partial, int accumulate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double mask = inside(M, i, j) ? 1.0 : 0.0;
    const double gn = init[c] + mask * hh * D[c];   // the update as k_visc_update forms it
    const double rho_in = g == 0 ? gn : P.rho[c], te_in = g == 4 ? gn : P.te[c];
    const double x = rho_in / K.m;
    const double rho1 = smax(x, K.nmin) * K.m;
    lam[c] *= (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
    lam[4 * N + c] *= te_in < K.temin ? 0.0 : 1.0;
    const double lg = lam[(size_t)g * N + c];
    lam[(size_t)g * N + c] = 0.0;
    lam_init[c] += lg;
    const double ldk = mask * hh * lg;
    for (int t = 0; t < tg.n; ++t) tg.ld[t][c] += tg.w[t] * ldk;
    dot = cf * mask * lg * D[c];
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    partial[b] = accumulate ? partial[b] + dot : dot;
  }
}

__global__ void __launch_bounds__(256) k_adj_dt_grid(MhdSet L, MhdStatic st, MhdConst K, Mesh M, double* ldt,
                                                     double* lam_L) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  const double g = ldt[c];
  if (g == 0.0) return;
  ldt[c] = 0.0;
  const Dual8 dt = dt_cell_dual(L, st, K, M, i, j);
  for (int v = 0; v < 8; ++v) lam_L[v * N + c] += g * dt.d[v];
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }
inline dim3 tcr_grid(const Mesh& m) { return dim3((m.ny + TCR_COLS - 1) / TCR_COLS, (m.nx + TCR_ROWS - 1) / TCR_ROWS); }

}  // namespace

int adj_mod_blocks(const Mesh& m) {
  const dim3 g = g2(m, dim3(64, 4));
  return (int)(g.x * g.y);
}

void launch_adj_ghost_axis(MhdSet P, MhdConst k, Mesh mesh, const int kinds[4], const double s1[4],
                           const double s2[4], int axis, double* lam, hipStream_t s, int raw) {
  AdjGhostArgs a;
  a.P = P; a.k = k; a.mesh = mesh; a.raw = raw;
  for (int q = 0; q < 4; ++q) { a.kind[q] = kinds[q]; a.s1[q] = s1[q]; a.s2[q] = s2[q]; }
  const int s0 = axis == 0 ? 0 : 2;
  if (!(a.kind[s0] || a.kind[s0 + 1])) return;
  const int len = axis == 0 ? mesh.ny : mesh.nx;
  hipLaunchKernelGGL(k_adj_ghost_axis, dim3((len + 255) / 256), dim3(256), 0, s, a, axis == 0 ? 1 : 0, lam);
}

void launch_adj_minmask(const double* rho_in, const double* te_base, const double* add, const double* step,
                        MhdConst k, Mesh mesh, double* lam, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_minmask, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, rho_in, te_base, add, step, k, mesh, lam,
                     partial);
}

void launch_adj_tc_sub(const double* te_s, const double* n, const double* bhx, const double* bhy,
                       const double* dbxx, const double* dbxy, const double* dbyx, const double* dbyy, double h,
                       double k0, MhdConst k, Mesh mesh, double* lam_te, double* lam_T, double* lam_fz,
                       double* partial, int accumulate, hipStream_t s, TcSatRev sr) {
  if (sr.coef)
    hipLaunchKernelGGL(k_adj_tc_sub<true>, tcr_grid(mesh), dim3(64, 4), 0, s, te_s, n, bhx, bhy, dbxx, dbxy, dbyx,
                       dbyy, h, k0, k, mesh, lam_te, lam_T, lam_fz, partial, accumulate, sr);
  else
    hipLaunchKernelGGL(k_adj_tc_sub<false>, tcr_grid(mesh), dim3(64, 4), 0, s, te_s, n, bhx, bhy, dbxx, dbxy, dbyx,
                       dbyy, h, k0, k, mesh, lam_te, lam_T, lam_fz, partial, accumulate, sr);
}

int adj_tc_strip_blocks(const Mesh& m, int R) {
  if (R < 1 || m.ny < 64 || m.nx < 2 * R) return 0;
  return ((m.ny + 59) / 60) * ((m.nx + R - 1) / R);
}

void launch_adj_tc_strip(const double* te_s, const double* n, const double* bhx, const double* bhy, double h, double k0,
                         MhdConst k, Mesh mesh, const double* lam_in, double* lam_out, double* lam_rho,
                         double* lam_bx, double* lam_by, double* partial, int accumulate, int R, hipStream_t s) {
  const int total = adj_tc_strip_blocks(mesh, R);
  if (!total) throw Error(SPRUCE_E_ARG, "conduction reverse strips: grid too small for the strip height");
  TcStepArgs a{};
  a.te_src = te_s; a.te_base = te_s; a.n = n; a.bhx = bhx; a.bhy = bhy;
  a.h = h; a.mult = 1.0; a.mode = TC_UPDATE; a.p.k0 = k0; a.k = k; a.mesh = mesh;
  hipLaunchKernelGGL(k_adj_tc_strip, dim3(total), dim3(64), 0, s, a, lam_in, lam_out, lam_rho, lam_bx, lam_by,
                     partial, accumulate, R, (mesh.nx + R - 1) / R, total);
}

int adj_tc_blocks(const Mesh& m) {
  const dim3 g = tcr_grid(m);
  return (int)(g.x * g.y);
}

void launch_adj_tc_stage(const double* eval, const double* n, const double* bhx, const double* bhy,
                         const double* dbxx, const double* dbxy, const double* dbyx, const double* dbyy, double h,
                         double mult, double k0, MhdConst k, Mesh mesh, const double* cbase, const double* lam_out,
                         double* lam_base, int assign_base, const double* seed_add, double add_coef, double* lam_T,
                         double* lam_fz, double* partial, int accumulate, hipStream_t s, TcSatRev sr) {
  if (sr.coef)
    hipLaunchKernelGGL(k_adj_tc_stage<true>, tcr_grid(mesh), dim3(64, 4), 0, s, eval, n, bhx, bhy, dbxx, dbxy, dbyx,
                       dbyy, h, mult, k0, k, mesh, cbase, lam_out, lam_base, assign_base, seed_add, add_coef, lam_T,
                       lam_fz, partial, accumulate, sr);
  else
    hipLaunchKernelGGL(k_adj_tc_stage<false>, tcr_grid(mesh), dim3(64, 4), 0, s, eval, n, bhx, bhy, dbxx, dbxy, dbyx,
                       dbyy, h, mult, k0, k, mesh, cbase, lam_out, lam_base, assign_base, seed_add, add_coef, lam_T,
                       lam_fz, partial, accumulate, sr);
}

void launch_adj_tc_final4(const double* te, const double* k1, const double* k2, const double* k3, const double* i3,
                          const double* n, const double* bhx, const double* bhy, const double* dbxx,
                          const double* dbxy, const double* dbyx, const double* dbyy, double h, double k0, MhdConst k,
                          Mesh mesh, const double* lam_out, double* lam_base, double* lamS, double* lam_T,
                          double* lam_fz, double* partial, int accumulate, hipStream_t s, TcSatRev sr) {
  if (sr.coef)
    hipLaunchKernelGGL(k_adj_tc_final4<true>, tcr_grid(mesh), dim3(64, 4), 0, s, te, k1, k2, k3, i3, n, bhx, bhy, dbxx,
                       dbxy, dbyx, dbyy, h, k0, k, mesh, lam_out, lam_base, lamS, lam_T, lam_fz, partial, accumulate,
                       sr);
  else
    hipLaunchKernelGGL(k_adj_tc_final4<false>, tcr_grid(mesh), dim3(64, 4), 0, s, te, k1, k2, k3, i3, n, bhx, bhy,
                       dbxx, dbxy, dbyx, dbyy, h, k0, k, mesh, lam_out, lam_base, lamS, lam_T, lam_fz, partial,
                       accumulate, sr);
}

void launch_adj_divclean(const double* bx_s, const double* by_s, const double* dxb, const double* dyb,
                         const double* coeff, double h, Mesh mesh, double* lbx, double* lby, double* scratch,
                         double* partial, hipStream_t s) {
  const size_t N = (size_t)mesh.nx * mesh.ny;
  double *ax = scratch, *ay = scratch + N, *ux = scratch + 2 * N, *uy = scratch + 3 * N;
  hipLaunchKernelGGL(k_adj_dc_a, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, bx_s, by_s, dxb, dyb, coeff, h, mesh,
                     (const double*)lbx, (const double*)lby, ax, ay, ux, uy, partial);
  hipLaunchKernelGGL(k_adj_dc_b, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, mesh, ux, uy, ax, ay);
  hipLaunchKernelGGL(k_adj_dc_c, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, ax, ay, lbx, lby);
}

void launch_adj_outflow(MhdSet P, MhdStatic st, const double* tmpl, const double* sc, const double* step,
                        OutflowParams p, Mesh mesh, const double* pv, const double* pk, int nb, double* lam,
                        double* pa, double* ph, hipStream_t s) {
  const dim3 g = g2(mesh, dim3(64, 4));
  hipLaunchKernelGGL(k_adj_outflow, g, dim3(64, 4), 0, s, P, st, tmpl, sc, step, p, mesh, lam, pa, ph);
  if (p.dynamic)
    hipLaunchKernelGGL(k_adj_outflow_dyn, dim3(1), dim3(256), 0, s, P, st, p, mesh, pv, pk, nb, (const double*)pa,
                       (int)(g.x * g.y), lam);
}

void launch_adj_fheat(MhdSet Ln, MhdSet P, MhdStatic st, const double* bhx, const double* bhy, FheatParams p,
                      const double* step, MhdConst k, Mesh mesh, double* lam, double* lamLn, double* lfz,
                      double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_fheat, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, Ln, P, st, bhx, bhy, p, step, k, mesh, lam,
                     lamLn, lfz, partial);
}

void launch_adj_visc_term(const ViscTermArgs& a, int evol, const double* ka, double sa, const double* kb, double* lq,
                          double* lf, double* lam_S, double* ldt, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_visc_term, g2(a.mesh, dim3(64, 4)), dim3(64, 4), 0, s, a, evol, ka, sa, kb, lq, lf, lam_S,
                     ldt, partial);
}

void launch_adj_visc_f(const ViscTermArgs& a, double* lf, double* lam_S, double* ldt, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_visc_f, g2(a.mesh, dim3(64, 4)), dim3(64, 4), 0, s, a, lf, lam_S, ldt, partial);
}

void launch_adj_visc_vars(const ViscTermArgs& a, double* lq, double* lam_S, hipStream_t s) {
  const size_t N = (size_t)a.mesh.nx * a.mesh.ny;
  hipLaunchKernelGGL(k_adj_visc_vars, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, a, lq, lam_S);
}

void launch_adj_hv_update(MhdSet P, int g, const double* init, const double* D, double c, double hh, MhdConst k,
                          Mesh mesh, double* lam, double* lam_init, int nt, double* const ld[4], const double w[4],
                          double* partial, int accumulate, hipStream_t s) {
  HvTargets tg{};
  tg.n = nt;
  for (int t = 0; t < nt; ++t) { tg.ld[t] = ld[t]; tg.w[t] = w[t]; }
  hipLaunchKernelGGL(k_adj_hv_update, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, P, g, init, D, c, hh, k, mesh, lam,
                     lam_init, tg, partial, accumulate);
}

void launch_adj_dt_grid(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double* ldt, double* lam_L, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_dt_grid, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, st, k, mesh, ldt, lam_L);
}

void launch_adj_tc_satrev(const double* eval, const double* n, const double* rho, const double* bhx,
                          const double* bhy, TcParams p, MhdConst k, Mesh mesh, TcSatRev sr, double* lam_T,
                          double* lam_fz, double* lam_rho, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_tc_satrev, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, eval, n, rho, bhx, bhy, p, k, mesh,
                     sr, lam_T, lam_fz, lam_rho);
}

void launch_adj_tc_temp(const double* te_s, const double* n, MhdConst k, size_t N, double* lam_T, double* lam_te,
                        double* lam_rho, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_tc_temp, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, te_s, n, k, N, lam_T, lam_te,
                     lam_rho);
}

void launch_adj_bhat(MhdSet L, MhdStatic st, Mesh mesh, double* lam_fz, double* lam, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_bhat_derivs, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, mesh, lam_fz);
  hipLaunchKernelGGL(k_adj_bhat_field, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, st, mesh, (const double*)lam_fz,
                     lam);
}

void launch_adj_rl(MhdSet A, MhdStatic st, RlParams p, MhdConst k, Mesh mesh, int nsub, double h, int integ,
                   double* lam, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_rl, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, A, st, p, k, mesh, nsub, h, integ, lam,
                     partial);
}

void launch_adj_source(MhdSet P, int mode, const double* t, const double* t2, const double* step, double g, double gp,
                       MhdConst k, Mesh mesh, double* lam, double* ph, double* pt, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_source, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, P, mode, t, t2, step, g, gp, k, mesh, lam,
                     ph, pt);
}

void launch_adj_heat_params(const double* lam_te, const double* basis, int K, const double* step, Mesh mesh,
                            double* colsum, double* g_theta, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_colsum, dim3((mesh.ny + 63) / 64), dim3(64, 16), 0, s, lam_te, mesh, colsum);
  hipLaunchKernelGGL(k_adj_theta, dim3(K), dim3(256), 0, s, (const double*)colsum, basis, step, mesh, g_theta);
}

void launch_adj_accum(const double* partial, int n, double scale, double* gh, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_accum, dim3(1), dim3(256), 0, s, partial, n, scale, gh);
}

}  // namespace spruce
