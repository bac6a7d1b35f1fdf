// Two-fluid (ion + electron fluids with electric and magnetic fields) time-step kernels for gfx950,
// following source/equationsets/ideal2F.cpp operation for operation.
//
// Every launch covers a batch of domains: block b handles tile b of the concatenated tile lists of
// all domains (tile_dom maps it to its domain), so one domain and 64 domains of different sizes use
// the same kernels, and a whole step of the batch is a fixed sequence of launches that a HIP graph
// can capture.  Each domain keeps its own step size on the device.
#include <cfloat>

#include <type_traits>

#include "arith.hpp"
#include "strip.hpp"
#include <hip/hip_cooperative_groups.h>

#include "host_common.hpp"
#include "twofluid.hpp"

namespace spruce {
namespace {

constexpr double E_CH = E_CHARGE;   // 4.80320425e-10 (constants.hpp:19)
constexpr double C_L = C_LIGHT;     // 29979245800 (constants.hpp:18)

__device__ __forceinline__ int wrp(int k, int n, int per) {
  if (per) {
    if (k < 0) k += n;
    else if (k >= n) k -= n;
  }
  return k;
}
__device__ __forceinline__ bool inside(const Mesh& M, int i, int j) {
  return i >= M.x.lo && i <= M.x.hi && j >= M.y.lo && j <= M.y.hi;
}
__device__ __forceinline__ const TfSet& pick(const TfDom& D, int sel) {
  return sel == TF_SEL_L ? D.L : (sel == TF_SEL_W ? D.W : D.W2);
}

// Species layout: 0 = ions, 1 = electrons.
__device__ __forceinline__ int v_rho(int s) { return s ? TF_ERHO : TF_IRHO; }
__device__ __forceinline__ int v_mx(int s) { return s ? TF_EMX : TF_IMX; }
__device__ __forceinline__ int v_my(int s) { return s ? TF_EMY : TF_IMY; }
__device__ __forceinline__ int v_te(int s) { return s ? TF_ETE : TF_ITE; }

struct TfFace {
  double P[2][4];   // surf * face velocity of rho, mom_x, mom_y, thermal_energy per species
  double vf[2];     // face velocity per species (boundaryInterpolate of v along the axis)
  double ip[2];     // interpolated pressure per species
  double fl[4];     // interpolated b_z, b_(other), E_z, E_(other): x-faces b_y, E_y; y-faces b_x, E_x
};

template <int AX>
__device__ __forceinline__ void tf_face(const TfDom& D, const TfSet& S, int i, int j, TfFace& F) {
  const Mesh& M = D.mesh;
  const Axis& ax = AX == 0 ? M.x : M.y;
  const int c = AX == 0 ? i : j;
  const int cm2 = wrp(c - 2, ax.n, ax.per), cm1 = wrp(c - 1, ax.n, ax.per), cp1 = wrp(c + 1, ax.n, ax.per);
  auto at = [&](int k) -> size_t { return AX == 0 ? (size_t)k * M.ny + j : (size_t)i * M.ny + k; };
  const size_t i0 = at(cm2), i1 = at(cm1), i2 = at(c), i3 = at(cp1);
  const double hlo = ax.h[cm1], hup = ax.h[c], den = ax.fden[c], rden = ax.rfden[c];
  auto interp = [&](double a, double b) { return qdiv(a * hup + b * hlo, den, rden); };
  const TfConst& K = D.c;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double* rho = S.v[v_rho(s)];
    const double* mom = S.v[AX == 0 ? v_mx(s) : v_my(s)];
    const double* te = S.v[v_te(s)];
    const double vf = interp(mom[i1] / rho[i1], mom[i2] / rho[i2]);
    F.vf[s] = vf;
    F.ip[s] = interp(te[i1] * K.gm1, te[i2] * K.gm1);
    const bool pos = vf > 0.0, neg = vf < 0.0;
    const double g = pos ? ax.gup[c] : ax.gdn[c];
    const int ce = pos ? cm1 : cp1;
    const double eden = ax.fden[ce], reden = ax.rfden[ce];
    const double* src[4] = {rho, S.v[v_mx(s)], S.v[v_my(s)], te};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double q0 = src[q][i0], q1 = src[q][i1], q2 = src[q][i2], q3 = src[q][i3];
      const double d2 = interp(q1, q2);
      const double a = pos ? q0 : q3, b = pos ? q1 : q2;
      const double d1 = a + qdiv((b - a) * g, eden, reden);
      const bool le = q2 <= q1;
      const bool outer_min = pos ? le : !le;
      const double inner = outer_min ? smax(d1, d2) : smin(d1, d2);
      const double lim = outer_min ? smin(b, inner) : smax(b, inner);
      F.P[s][q] = ((pos || neg) ? lim : d2) * vf;
    }
  }
  if (K.curl) {
    const double* bz = S.v[TF_BZ];
    const double* Ez = S.v[TF_EZ];
    F.fl[0] = interp(bz[i1], bz[i2]);
    F.fl[2] = interp(Ez[i1], Ez[i2]);
    if (AX == 0) {  // b_y = be_y + bi_y (ideal2F.cpp:136)
      const double* by = S.v[TF_BY];
      F.fl[1] = interp(D.bey[i1] + by[i1], D.bey[i2] + by[i2]);
      F.fl[3] = interp(S.v[TF_EY][i1], S.v[TF_EY][i2]);
    } else {
      const double* bx = S.v[TF_BX];
      F.fl[1] = interp(D.bex[i1] + bx[i1], D.bex[i2] + bx[i2]);
      F.fl[3] = interp(S.v[TF_EX][i1], S.v[TF_EX][i2]);
    }
  }
}

// Differences of the two faces of a cell along one axis, divided by the cell size.
struct TfAxisTerms {
  double tdiv[2][4];  // transport divergence contribution per species and quantity
  double dp[2];       // pressure derivative per species
  double dv[2];       // velocity derivative per species (one term of divergence2D(v))
  double fl[4];       // field derivatives: d b_z, d b_(other), d E_z, d E_(other)
};

template <int AX>
__device__ __forceinline__ void tf_axis(const TfDom& D, const TfSet& S, int i, int j, TfAxisTerms& T) {
  const Mesh& M = D.mesh;
  TfFace lo, up;
  tf_face<AX>(D, S, i, j, lo);
  if (AX == 0) tf_face<0>(D, S, wrp(i + 1, M.x.n, M.x.per), j, up);
  else tf_face<1>(D, S, i, wrp(j + 1, M.y.n, M.y.per), up);
  const Axis& ax = AX == 0 ? M.x : M.y;
  const int c = AX == 0 ? i : j;
  const double d = ax.d[c], rd = ax.rd[c];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int q = 0; q < 4; ++q) T.tdiv[s][q] = qdiv(up.P[s][q] - lo.P[s][q], d, rd);
    T.dp[s] = qdiv(up.ip[s] - lo.ip[s], d, rd);
    T.dv[s] = qdiv(up.vf[s] - lo.vf[s], d, rd);
  }
  if (D.c.curl) {
#pragma unroll
    for (int q = 0; q < 4; ++q) T.fl[q] = qdiv(up.fl[q] - lo.fl[q], d, rd);
  }
}

// Ideal2F::computeTimeDerivativesDerived at one cell, ghost mask applied (ideal2F.cpp:30-96).
__device__ void tf_rhs_cell(const TfDom& D, const TfSet& S, int i, int j, double k[TF_NE]) {
  const Mesh& M = D.mesh;
  const TfConst& K = D.c;
  const size_t c = (size_t)i * M.ny + j;
  double tdiv[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  double dpx[2] = {0.0, 0.0}, dpy[2] = {0.0, 0.0}, dv[2] = {0.0, 0.0};
  double dbz_x = 0.0, dbz_y = 0.0, dby_x = 0.0, dbx_y = 0.0, dEz_x = 0.0, dEz_y = 0.0, dEy_x = 0.0, dEx_y = 0.0;
  const bool in = inside(M, i, j);
  if (in) {  // derivatives are zero outside the interior (derivs.cpp:229-233)
    TfAxisTerms tx, ty;
    tf_axis<0>(D, S, i, j, tx);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int q = 0; q < 4; ++q) tdiv[s][q] = tx.tdiv[s][q];
      dpx[s] = tx.dp[s];
      dv[s] = tx.dv[s];
    }
    if (K.curl) { dbz_x = tx.fl[0]; dby_x = tx.fl[1]; dEz_x = tx.fl[2]; dEy_x = tx.fl[3]; }
    tf_axis<1>(D, S, i, j, ty);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int q = 0; q < 4; ++q) tdiv[s][q] = tdiv[s][q] + ty.tdiv[s][q];   // transportDivergence2D
      dpy[s] = ty.dp[s];
      dv[s] = dv[s] + ty.dv[s];                                                // divergence2D(v)
    }
    if (K.curl) { dbz_y = ty.fl[0]; dbx_y = ty.fl[1]; dEz_y = ty.fl[2]; dEx_y = ty.fl[3]; }
  }
  const double irho = S.v[TF_IRHO][c], erho = S.v[TF_ERHO][c];
  const double ivx = S.v[TF_IMX][c] / irho, ivy = S.v[TF_IMY][c] / irho;
  const double evx = S.v[TF_EMX][c] / erho, evy = S.v[TF_EMY][c] / erho;
  const double in_ = qdiv(irho, K.mi, K.rmi), en_ = qdiv(erho, K.me, K.rme);
  const double jx = in_ * E_CH * ivx - en_ * E_CH * evx;          // ideal2F.cpp:128-129
  const double jy = in_ * E_CH * ivy - en_ * E_CH * evy;
  const double bz = S.v[TF_BZ][c];
  const double Ex = S.v[TF_EX][c], Ey = S.v[TF_EY][c];
  const double ivb0 = ivy * bz, ivb1 = (-ivx) * bz, evb0 = evy * bz, evb1 = (-evx) * bz;
  const double iFx = E_CH * in_ * (Ex + qdiv(ivb0, C_L, K.rC));
  const double iFy = E_CH * in_ * (Ey + qdiv(ivb1, C_L, K.rC));
  const double eFx = -E_CH * en_ * (Ex + qdiv(evb0, C_L, K.rC));
  const double eFy = -E_CH * en_ * (Ey + qdiv(evb1, C_L, K.rC));
  const double ipc = S.v[TF_ITE][c] * K.gm1, epc = S.v[TF_ETE][c] * K.gm1;
  k[TF_IRHO] = -tdiv[0][0];
  k[TF_ERHO] = -tdiv[1][0];
  k[TF_IMX] = -tdiv[0][1] - dpx[0] + irho * D.gx[c] + iFx;
  k[TF_IMY] = -tdiv[0][2] - dpy[0] + irho * D.gy[c] + iFy;
  k[TF_EMX] = -tdiv[1][1] - dpx[1] + erho * D.gx[c] + eFx;
  k[TF_EMY] = -tdiv[1][2] - dpy[1] + erho * D.gy[c] + eFy;
  k[TF_ITE] = -tdiv[0][3] - ipc * dv[0];
  k[TF_ETE] = -tdiv[1][3] - epc * dv[1];
  if (K.curl) {                                                    // ideal2F.cpp:75-80
    k[TF_EX] = C_L * dbz_y - 4. * PI * jx;
    k[TF_EY] = -C_L * dbz_x - 4. * PI * jy;
    k[TF_EZ] = C_L * (dby_x - dbx_y);
    k[TF_BX] = -C_L * dEz_y;
    k[TF_BY] = C_L * dEz_x;
    k[TF_BZ] = C_L * (dEx_y - dEy_x);
  } else {                                                         // :83-88
    k[TF_EX] = -4. * PI * jx;
    k[TF_EY] = -4. * PI * jy;
    k[TF_EZ] = 0.0;
    k[TF_BX] = 0.0;
    k[TF_BY] = 0.0;
    k[TF_BZ] = 0.0;
  }
  const double mask = in ? 1.0 : 0.0;                              // :94
#pragma unroll
  for (int v = 0; v < TF_NE; ++v) k[v] *= mask;
}

__device__ __forceinline__ bool tile_cell(const TfBatch& b, int& d, int& i, int& j) {
  const int t = blockIdx.x;
  d = b.tile_dom[t];
  const TfDom& D = b.doms[d];
  const int lt = t - D.tile_base;
  const int ty = lt / D.tiles_x, tx = lt - ty * D.tiles_x;
  i = ty * b.by + threadIdx.y;
  j = tx * b.bx + threadIdx.x;
  return i < D.mesh.nx && j < D.mesh.ny;
}

__global__ __launch_bounds__(256) void k_tf_stage(TfBatch b, int Ssel, int Bsel, int Tsel, double mult, int mode, int kslot) {
  int d, i, j;
  if (!tile_cell(b, d, i, j)) return;
  const TfDom& D = b.doms[d];
  const TfSet& S = pick(D, Ssel);
  const TfSet& B = pick(D, Bsel);
  const TfSet& T = pick(D, Tsel);
  const TfConst& K = D.c;
  const size_t N = (size_t)D.mesh.nx * D.mesh.ny;
  const size_t c = (size_t)i * D.mesh.ny + j;
  double k[TF_NE];
  tf_rhs_cell(D, S, i, j, k);
  if (mode == TF_STAGE_STORE_K) {
    for (int v = 0; v < TF_NE; ++v) D.k[kslot][v * N + c] = k[v];
  } else if (mode == TF_STAGE_RK4_FINAL) {                         // evolution.cpp:119-121
    for (int v = 0; v < TF_NE; ++v)
      k[v] = qdiv(D.k[0][v * N + c] + k[v], 6.0, K.r6) + qdiv(D.k[1][v * N + c] + D.k[2][v * N + c], 3.0, K.r3);
  }
  const double step = mult * (*D.step);
  double nv[TF_NE];
#pragma unroll
  for (int v = 0; v < TF_NE; ++v) nv[v] = B.v[v][c] + k[v] * step;
  // enforceMinimums (ideal2F.cpp:98-105)
  nv[TF_IRHO] = smax(qdiv(nv[TF_IRHO], K.mi, K.rmi), K.nmin) * K.mi;
  nv[TF_ERHO] = smax(qdiv(nv[TF_ERHO], K.me, K.rme), K.nmin) * K.me;
  nv[TF_ITE] = smax(nv[TF_ITE], K.temin);
  nv[TF_ETE] = smax(nv[TF_ETE], K.temin);
#pragma unroll
  for (int v = 0; v < TF_NE; ++v) T.v[v][c] = nv[v];
}

// recomputeEvolvedVarsFromStateVars, then enforceMinimums (ideal2F.cpp:107-117, 98-105)
__global__ void k_tf_populate(TfBatch b, const double* const* itemp, const double* const* etemp) {
  int d, i, j;
  if (!tile_cell(b, d, i, j)) return;
  const TfDom& D = b.doms[d];
  const TfConst& K = D.c;
  const size_t c = (size_t)i * D.mesh.ny + j;
  const TfSet& L = D.L;
  const double irho = L.v[TF_IRHO][c], erho = L.v[TF_ERHO][c];
  const double in_ = qdiv(irho, K.mi, K.rmi), en_ = qdiv(erho, K.me, K.rme);
  const double ip = in_ * K_B * itemp[d][c], ep = en_ * K_B * etemp[d][c];
  L.v[TF_ITE][c] = smax(qdiv(ip, K.gm1, K.rgm1), K.temin);
  L.v[TF_ETE][c] = smax(qdiv(ep, K.gm1, K.rgm1), K.temin);
  L.v[TF_IRHO][c] = smax(in_, K.nmin) * K.mi;
  L.v[TF_ERHO][c] = smax(en_, K.nmin) * K.me;
}

// Ghost fill of one side (evolution.cpp:126-334) for a two-species set.  Writes values only: the
// two-fluid derived step does not change evolved variables (ideal2F.cpp:119-154).
__device__ void tf_ghost_side(const TfDom& D, const TfSet& G, int side, int kind, int t) {
  const Mesh& M = D.mesh;
  const bool xside = side < 2, upper = side & 1;
  const Axis& ax = xside ? M.x : M.y;
  const Axis& ot = xside ? M.y : M.x;
  if (kind == 2) { if (t < 0 || t >= ot.n) return; }
  else if (t < ot.lo || t > ot.hi) return;
  const int n = ax.n;
  const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
  auto at = [&](int e) -> size_t { return xside ? (size_t)e * M.ny + t : (size_t)t * M.ny + e; };
  const size_t c1 = at(e1), c2 = at(e2), c3 = at(e3);
  const TfConst& K = D.c;
  if (kind == 1) {                                                   // open, evolution.cpp:158-224
    const double s1 = D.s1[side], s2 = D.s2[side];
    for (int s = 0; s < 2; ++s) {
      const double r3 = G.v[v_rho(s)][c3];
      G.v[v_rho(s)][c1] = s1 * r3;
      G.v[v_rho(s)][c2] = s2 * r3;
    }
    for (int s = 0; s < 2; ++s) {
      const double t3 = G.v[v_te(s)][c3];
      G.v[v_te(s)][c1] = s1 * t3;
      G.v[v_te(s)][c2] = s2 * t3;
    }
    double cs = 0.0;
    for (int s = 0; s < 2; ++s) {
      const double csn = sqrt(K.gamma * (G.v[v_te(s)][c3] * K.gm1) / G.v[v_rho(s)][c3]);
      if (csn > cs) cs = csn;
    }
    const double dist = 0.5 * (ax.d[e2] + ax.d[e3]);
    for (int s = 0; s < 2; ++s) {
      double* mx = G.v[v_mx(s)];
      double* my = G.v[v_my(s)];
      const double* rho = G.v[v_rho(s)];
      const double vx = mx[c3] / rho[c3], vy = my[c3] / rho[c3];
      double boost = K.obs * cs;
      if (!upper) boost *= -1.0;
      const double vn = xside ? vx : vy;
      const double bv = upper ? smax(0.0, vn + boost) : smin(0.0, vn + boost);
      const double gv = (dist * bv - 0.5 * ax.d[e2] * vn) / (0.5 * ax.d[e3]);
      const double r1 = rho[c1], r2 = rho[c2];
      if (xside) { mx[c1] = r1 * gv; mx[c2] = r2 * gv; my[c1] = r1 * vy; my[c2] = r2 * vy; }
      else { mx[c1] = r1 * vx; mx[c2] = r2 * vx; my[c1] = r1 * gv; my[c2] = r2 * gv; }
    }
  } else if (kind == 3) {                                            // reflect, :231-266
    for (int s = 0; s < 2; ++s) {
      const double t3 = G.v[v_te(s)][c3];
      G.v[v_te(s)][c1] = t3; G.v[v_te(s)][c2] = t3;
    }
    for (int s = 0; s < 2; ++s) {
      const double r3 = G.v[v_rho(s)][c3];
      G.v[v_rho(s)][c1] = r3; G.v[v_rho(s)][c2] = r3;
    }
    for (int v : {TF_IMX, TF_IMY, TF_EMX, TF_EMY}) { G.v[v][c1] = 0.0; G.v[v][c2] = 0.0; G.v[v][c3] = 0.0; }
  } else if (kind == 2) {                                            // fixed, :272-282
    for (int v : {TF_IMX, TF_IMY, TF_EMX, TF_EMY}) { G.v[v][c1] = 0.0; G.v[v][c2] = 0.0; G.v[v][c3] = 0.0; }
  } else if (kind == 5) {                                            // open_ucnp, :290-334
    const size_t ci = at(upper ? ax.hi : ax.lo);
    for (int g = 0; g < N_GHOST; ++g) {
      const size_t cb = at(upper ? ax.hi + 1 + g : g);
      for (int v : {TF_IRHO, TF_ERHO, TF_ITE, TF_ETE, TF_EX, TF_EY, TF_EZ, TF_IMX, TF_IMY, TF_EMX, TF_EMY})
        G.v[v][cb] = G.v[v][ci];
    }
  }
}

__global__ void k_tf_ghost(TfBatch b, int Gsel, int kinds_mask, int xsides) {
  const int d = blockIdx.y;
  const TfDom& D = b.doms[d];
  const TfSet& G = pick(D, Gsel);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s0 = xsides ? 0 : 2;
  for (int s = s0; s < s0 + 2; ++s) {
    const int kind = D.bc[s];
    if (kind != 0 && kind != 4 && ((kinds_mask >> kind) & 1)) tf_ghost_side(D, G, s, kind, t);
  }
}

// Ideal2F::recomputeDT at one cell (ideal2F.cpp:169-198).
__device__ double tf_dt_cell(const TfDom& D, const TfSet& L, int i, int j) {
  const Mesh& M = D.mesh;
  const TfConst& K = D.c;
  const size_t c = (size_t)i * M.ny + j;
  const double erho = L.v[TF_ERHO][c];
  const double en = qdiv(erho, K.me, K.rme);
  const double ep = L.v[TF_ETE][c] * K.gm1;
  const double etemp = ep / (K_B * en);
  const double evx = L.v[TF_EMX][c] / erho, evy = L.v[TF_EMY][c] / erho;
  const double vth = sqrt(qdiv(K_B * etemp, K.me, K.rme));
  const double wpe = sqrt(qdiv(4 * PI * en * E_CH * E_CH, K.me, K.rme));
  const double vth2 = vth * vth;
  const double wthx = sqrt(D.kx3[i] * vth2), wthy = sqrt(D.ky3[j] * vth2);
  const double wLx = sqrt(wpe * wpe + wthx * wthx), wLy = sqrt(wpe * wpe + wthy * wthy);
  const double vLx = wLx / D.kxv[i], vLy = wLy / D.kyv[j];
  const double vmx = fabs(evx) + vLx, vmy = fabs(evy) + vLy;
  const double dx = M.x.d[i], dy = M.y.d[j];
  double dt = 1. / (qdiv(vmx, dx, M.x.rd[i]) + qdiv(vmy, dy, M.y.rd[j]));
  if (K.em_dt) {
    const double dtem = qdiv(dx * dy / (dx + dy), C_L, K.rC);
    dt = smin(dt, dtem);
  }
  return dt;
}

__device__ __forceinline__ double block_min(double v) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  __shared__ double part[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  if ((lane & 63) == 0) part[lane >> 6] = v;
  __syncthreads();
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    double m = part[0];
    for (int k = 1; k < nw; ++k) m = (part[k] < m) ? part[k] : m;
    v = m;
  }
  return v;
}

__global__ void k_tf_dt(TfBatch b) {
  int d, i, j;
  const bool ok = tile_cell(b, d, i, j);
  const TfDom& D = b.doms[d];
  double v = DBL_MAX;
  if (ok && inside(D.mesh, i, j)) {
    const double dt = tf_dt_cell(D, D.L, i, j);
    v = (dt < v) ? dt : v;
  }
  v = block_min(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) D.partial[blockIdx.x - D.tile_base] = v;
}

__global__ void k_tf_reduce(TfBatch b) {
  const TfDom& D = b.doms[blockIdx.x];
  double v = DBL_MAX;
  const int n = D.tiles_x * D.tiles_y;
  for (int k = threadIdx.x; k < n; k += blockDim.x) { const double o = D.partial[k]; v = (o < v) ? o : v; }
  for (int off = 32; off > 0; off >>= 1) { const double o = __shfl_xor(v, off, 64); v = (o < v) ? o : v; }
  __shared__ double part[16];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = (part[k] < m) ? part[k] : m;
    const double s = D.c.eps * m;
    *D.step = s;
    const int idx = *D.counter;
    if (idx < D.hist_cap) D.history[idx] = (m < DBL_MAX) ? s : __builtin_nan("");
    *D.counter = idx + 1;
  }
}

__global__ void k_tf_reset(TfBatch b) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < b.ndom) {
    b.doms[d].history[0] = *b.doms[d].step;
    *b.doms[d].counter = 1;
  }
}

__global__ void k_tf_copy(TfBatch b, int from, int to) {
  int d, i, j;
  if (!tile_cell(b, d, i, j)) return;
  const TfDom& D = b.doms[d];
  const TfSet& F = pick(D, from);
  const TfSet& T = pick(D, to);
  const size_t c = (size_t)i * D.mesh.ny + j;
#pragma unroll
  for (int v = 0; v < TF_NE; ++v) T.v[v][c] = F.v[v][c];
}

// Derived variables of ideal2F.cpp:119-154 and recomputeDT for download.
__global__ void k_tf_derived(TfBatch b, int d, int which, double* out) {
  const TfDom& D = b.doms[d];
  const Mesh& M = D.mesh;
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const TfConst& K = D.c;
  const TfSet& L = D.L;
  const size_t c = (size_t)i * M.ny + j;
  const double irho = L.v[TF_IRHO][c], erho = L.v[TF_ERHO][c];
  const double in_ = qdiv(irho, K.mi, K.rmi), en_ = qdiv(erho, K.me, K.rme);
  const double ivx = L.v[TF_IMX][c] / irho, ivy = L.v[TF_IMY][c] / irho;
  const double evx = L.v[TF_EMX][c] / erho, evy = L.v[TF_EMY][c] / erho;
  const double ip = L.v[TF_ITE][c] * K.gm1, ep = L.v[TF_ETE][c] * K.gm1;
  const double bx = D.bex[c] + L.v[TF_BX][c], by = D.bey[c] + L.v[TF_BY][c], bz = L.v[TF_BZ][c];
  const double bmxy = sqrt(bx * bx + by * by);
  auto deriv = [&](int axis, auto f) -> double {   // derivative1D of a cell function f(i, j)
    if (!inside(M, i, j)) return 0.0;
    const Axis& ax = axis == 0 ? M.x : M.y;
    const int cc = axis == 0 ? i : j;
    const int cm = wrp(cc - 1, ax.n, ax.per), cp = wrp(cc + 1, ax.n, ax.per);
    const double fc = f(i, j);
    const double fm = axis == 0 ? f(cm, j) : f(i, cm), fp = axis == 0 ? f(cp, j) : f(i, cp);
    const double up = qdiv(fc * ax.h[cp] + fp * ax.h[cc], ax.fden[cp], ax.rfden[cp]);
    const double lo = qdiv(fm * ax.h[cc] + fc * ax.h[cm], ax.fden[cc], ax.rfden[cc]);
    return qdiv(up - lo, ax.d[cc], ax.rd[cc]);
  };
  auto fld = [&](int v) { return [&, v](int a, int bb) { return L.v[v][(size_t)a * M.ny + bb]; }; };
  double r = 0.0;
  switch (which) {
    case 0: r = ip / (K_B * in_); break;                         // i_temp
    case 1: r = ep / (K_B * en_); break;                         // e_temp
    case 2: r = in_; break;
    case 3: r = en_; break;
    case 4: r = ivx; break;
    case 5: r = ivy; break;
    case 6: r = evx; break;
    case 7: r = evy; break;
    case 8: r = in_ * E_CH * ivx - en_ * E_CH * evx; break;      // j_x
    case 9: r = in_ * E_CH * ivy - en_ * E_CH * evy; break;      // j_y
    case 10: r = ip; break;
    case 11: r = ep; break;
    case 12: r = ip + ep; break;                                 // press
    case 13: r = irho + erho; break;                             // rho
    case 14: r = E_CH * (in_ - en_); break;                      // rho_c
    case 15: r = in_; break;                                     // n
    case 16: r = in_ - en_; break;                               // dn
    case 17: r = tf_dt_cell(D, L, i, j); break;                  // dt
    case 18: {                                                   // dt_i (ideal2F.cpp:201-208)
      const double cs = sqrt(K.gamma * ip / irho);
      r = 1. / (qdiv(fabs(ivx) + cs, M.x.d[i], M.x.rd[i]) + qdiv(fabs(ivy) + cs, M.y.d[j], M.y.rd[j]));
      break;
    }
    case 19: r = bx; break;
    case 20: r = by; break;
    case 21: r = bz; break;
    case 22: r = sqrt(bx * bx + by * by + bz * bz); break;
    case 23: r = bmxy; break;
    case 24: r = bmxy == 0.0 ? 0.0 : bx / bmxy; break;
    case 25: r = bmxy == 0.0 ? 0.0 : by / bmxy; break;
    case 26: r = -(deriv(1, fld(TF_EX)) - deriv(0, fld(TF_EY))); break;                      // curlE_z
    case 27: r = deriv(0, fld(TF_EX)) + deriv(1, fld(TF_EY)); break;                         // divE
    case 28: {                                                                                 // divB
      auto fbx = [&](int a, int bb) { const size_t q = (size_t)a * M.ny + bb; return D.bex[q] + L.v[TF_BX][q]; };
      auto fby = [&](int a, int bb) { const size_t q = (size_t)a * M.ny + bb; return D.bey[q] + L.v[TF_BY][q]; };
      r = deriv(0, fbx) + deriv(1, fby);
      break;
    }
    case 29: r = deriv(0, [&](int a, int bb) { return L.v[TF_ITE][(size_t)a * M.ny + bb] * K.gm1; }); break;
    case 30: r = deriv(0, [&](int a, int bb) { return L.v[TF_ETE][(size_t)a * M.ny + bb] * K.gm1; }); break;
    case 31: r = D.gx[c]; break;
    case 32: r = D.gy[c]; break;
  }
  out[c] = r;
}

__global__ __launch_bounds__(256) void k_tf_rhs(TfBatch b, int d, double* out) {
  const TfDom& D = b.doms[d];
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= D.mesh.nx || j >= D.mesh.ny) return;
  const size_t N = (size_t)D.mesh.nx * D.mesh.ny, c = (size_t)i * D.mesh.ny + j;
  double k[TF_NE];
  tf_rhs_cell(D, D.L, i, j, k);
  for (int v = 0; v < TF_NE; ++v) out[v * N + c] = k[v];
}


// ---------------------------------------------------------------------------------------------
// Strip form of the stage kernel: one 64-lane wavefront per (domain, 60-column tile, strip of R
// rows), every face computed once (x-faces carried down the strip in registers, y-faces shared
// between neighbouring lanes), with the arithmetic of tf_face / tf_axis / tf_rhs_cell above,
// operation for operation.
// ---------------------------------------------------------------------------------------------
// per-cell transported quantities, index s * 4 + q (q: rho, mom_x, mom_y, thermal_energy)
constexpr int TQ = 8;
// per-cell field values: bz, Ez, be_y + bi_y, Ey (x-faces), be_x + bi_x, Ex (y-faces)
enum { TFF_BZ = 0, TFF_EZ, TFF_BYT, TFF_EY, TFF_BXT, TFF_EX, TFF_N };
// per-cell derived values: v_x, v_y and press per species
enum { TFD_VX0 = 0, TFD_VY0, TFD_P0, TFD_VX1, TFD_VY1, TFD_P1, TFD_N };

__device__ __forceinline__ void tf_load_q(const TfSet& S, size_t c, bool ok, double* q) {
  if (ok) {
    q[0] = S.v[TF_IRHO][c]; q[1] = S.v[TF_IMX][c]; q[2] = S.v[TF_IMY][c]; q[3] = S.v[TF_ITE][c];
    q[4] = S.v[TF_ERHO][c]; q[5] = S.v[TF_EMX][c]; q[6] = S.v[TF_EMY][c]; q[7] = S.v[TF_ETE][c];
  } else {
#pragma unroll
    for (int k = 0; k < TQ; ++k) q[k] = 0.0;
    q[0] = 1.0;
    q[4] = 1.0;
  }
}
__device__ __forceinline__ void tf_load_f(const TfDom& D, const TfSet& S, size_t c, bool ok, double* f) {
  if (ok) {
    f[TFF_BZ] = S.v[TF_BZ][c];
    f[TFF_EZ] = S.v[TF_EZ][c];
    f[TFF_BYT] = D.bey[c] + S.v[TF_BY][c];
    f[TFF_EY] = S.v[TF_EY][c];
    f[TFF_BXT] = D.bex[c] + S.v[TF_BX][c];
    f[TFF_EX] = S.v[TF_EX][c];
  } else {
#pragma unroll
    for (int k = 0; k < TFF_N; ++k) f[k] = 0.0;
  }
}
// v = mom/rho per species (ideal2F.cpp:124-127), press = (gamma-1) te (:131-132)
__device__ __forceinline__ void tf_derive(const double* q, const TfConst& K, double* dv) {
  dv[TFD_VX0] = q[1] / q[0];
  dv[TFD_VY0] = q[2] / q[0];
  dv[TFD_P0] = q[3] * K.gm1;
  dv[TFD_VX1] = q[5] / q[4];
  dv[TFD_VY1] = q[6] / q[4];
  dv[TFD_P1] = q[7] * K.gm1;
}

__device__ __forceinline__ void tf_strip_of(const TfBatch& b, int t, int& d, int& col, int& strip) {
  d = b.strip_dom[t];
  const TfDom& D = b.doms[d];
  const int lt = t - D.strip_base;
  col = lt / D.strips_x;
  strip = lt - col * D.strips_x;
}

// One strip tile t (a 64-lane wavefront); qs: the wavefront's 16 KiB of LDS for the rotating rows.
using TfRows = double[4][TQ][64];
__device__ __forceinline__ void tf_stage_strip_tile(const TfBatch& b, int t, int Ssel, int Bsel, int Tsel,
                                                    double mult, int mode, int kslot, TfRows& qs) {
  using namespace strip;
  int d, col, sidx;
  tf_strip_of(b, t, d, col, sidx);
  const TfDom& D = b.doms[d];
  const TfSet& S = pick(D, Ssel);
  const TfSet& B = pick(D, Bsel);
  const TfSet& T = pick(D, Tsel);
  const TfConst& K = D.c;
  const Mesh& M = D.mesh;
  const int nx = M.nx, ny = M.ny, R = b.strip_rows;
  const size_t N = (size_t)nx * ny;
  const int lane = threadIdx.x;
  const int jraw = col * OUT_COLS + lane - 2;
  int j;
  bool jvalid;
  if (M.y.per) { j = wrapn(jraw, ny); jvalid = true; }
  else { jvalid = jraw >= 0 && jraw < ny; j = clampn(jraw, ny); }
  const bool out_lane = lane >= 2 && lane < 62 && jraw >= 0 && jraw < ny;
  const bool yint = jvalid && j >= M.y.lo && j <= M.y.hi;
  const bool curl = K.curl != 0;
  const int r0 = sidx * R;
  auto row_of = [&](int r, bool& ok) -> int {
    if (M.x.per) { ok = true; return wrapn(r, nx); }
    ok = r >= 0 && r < nx;
    return clampn(r, nx);
  };
  auto cell = [&](int i) { return (size_t)i * ny + j; };

  // rotating rows r-1 .. r+2 of the transported quantities, one column per lane: kept in LDS
  // (16 KiB per wavefront) rather than registers, which leaves room for two wavefronts per SIMD
  auto Q = [&](int slot, int k) -> double& { return qs[slot][k][lane]; };
  auto load_q = [&](size_t c, bool ok, int slot) {
    double t[TQ];
    tf_load_q(S, c, ok, t);
#pragma unroll
    for (int k = 0; k < TQ; ++k) Q(slot, k) = t[k];
  };
  auto derive = [&](int slot, double* dvo) {
    double t[TQ];
#pragma unroll
    for (int k = 0; k < TQ; ++k) t[k] = Q(slot, k);
    tf_derive(t, K, dvo);
  };
  double f[2][TFF_N];         // fields of rows r, r+1
  double dv[2][TFD_N];        // derived values of rows r, r+1
  double Px[TQ];              // flux of row r's lower x-face
  double xvf[2][2], xip[2][2], xfl[2][4];   // lower/upper x-face: face velocity, press, fields

  {  // prologue: rows r0-2 .. r0+1, fields of row r0-1 and r0, lower face of row r0
    bool ok;
    int i;
    i = row_of(r0 - 2, ok); load_q(cell(i), ok && jvalid, 3);
    i = row_of(r0 - 1, ok); load_q(cell(i), ok && jvalid, 0);
    double ft[TFF_N], dt[TFD_N];
    if (curl) tf_load_f(D, S, cell(i), ok && jvalid, ft);
    i = row_of(r0, ok);     load_q(cell(i), ok && jvalid, 1);
    if (curl) tf_load_f(D, S, cell(i), ok && jvalid, f[0]);
    i = row_of(r0 + 1, ok); load_q(cell(i), ok && jvalid, 2);
    derive(0, dt);
    derive(1, dv[0]);
    const FaceGeom gx = face_geom(M.x, r0);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int o = s * 3;
      xvf[0][s] = interp(gx, dt[o + TFD_VX0], dv[0][o + TFD_VX0]);
      xip[0][s] = interp(gx, dt[o + TFD_P0], dv[0][o + TFD_P0]);
      const FaceDir fd = face_dir(xvf[0][s], gx);
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        double d2;
        const int k = s * 4 + qq;
        Px[k] = face_flux(Q(3, k), Q(0, k), Q(1, k), Q(2, k), gx, fd, d2);
      }
    }
    if (curl) {
      xfl[0][0] = interp(gx, ft[TFF_BZ], f[0][TFF_BZ]);
      xfl[0][1] = interp(gx, ft[TFF_BYT], f[0][TFF_BYT]);
      xfl[0][2] = interp(gx, ft[TFF_EZ], f[0][TFF_EZ]);
      xfl[0][3] = interp(gx, ft[TFF_EY], f[0][TFF_EY]);
    }
  }

  auto step = [&](auto BC, int r) {
    constexpr int Bi = decltype(BC)::value;
    constexpr int SM = Bi % 4, S0 = (Bi + 1) % 4, S1 = (Bi + 2) % 4, S2 = (Bi + 3) % 4;
    constexpr int D0 = Bi % 2, D1 = (Bi + 1) % 2;
    bool ok;
    // rows r+2 (transported) and r+1 (fields) are loaded first and arrive during the y-face work
    const int i2 = row_of(r + 2, ok);
    load_q(cell(i2), ok && jvalid, S2);
    const int i1 = row_of(r + 1, ok);
    if (curl) tf_load_f(D, S, cell(i1), ok && jvalid, f[D1]);
    if (r >= nx) return;
    const int i = row_of(r, ok);
    const bool in = i >= M.x.lo && i <= M.x.hi && yint;   // derivatives are zero outside (derivs.cpp:229-233)
    // ---- y-faces of row r: the lower face of this lane's cell, the upper one from the next lane
    const FaceGeom gy = face_geom(M.y, j);
    const double dy = M.y.d[j], rdy = M.y.rd[j];
    double tdiv[TQ], dpy[2], dvy[2], fly[4];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int o = s * 3;
      const double vf = interp(gy, shr1(dv[D0][o + TFD_VY0]), dv[D0][o + TFD_VY0]);
      const double ip = interp(gy, shr1(dv[D0][o + TFD_P0]), dv[D0][o + TFD_P0]);
      dpy[s] = qdiv(shl1(ip) - ip, dy, rdy);
      dvy[s] = qdiv(shl1(vf) - vf, dy, rdy);
      const FaceDir fd = face_dir(vf, gy);
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const double qc = Q(S0, s * 4 + qq);
        const double qm1 = shr1(qc);
        const double qm2 = shr1(qm1);
        const double qp1 = shl1(qc);
        double d2;
        const double Pl = face_flux(qm2, qm1, qc, qp1, gy, fd, d2);
        tdiv[s * 4 + qq] = qdiv(shl1(Pl) - Pl, dy, rdy);
      }
    }
    if (curl) {
      const double a0 = interp(gy, shr1(f[D0][TFF_BZ]), f[D0][TFF_BZ]);
      const double a1 = interp(gy, shr1(f[D0][TFF_BXT]), f[D0][TFF_BXT]);
      const double a2 = interp(gy, shr1(f[D0][TFF_EZ]), f[D0][TFF_EZ]);
      const double a3 = interp(gy, shr1(f[D0][TFF_EX]), f[D0][TFF_EX]);
      fly[0] = qdiv(shl1(a0) - a0, dy, rdy);
      fly[1] = qdiv(shl1(a1) - a1, dy, rdy);
      fly[2] = qdiv(shl1(a2) - a2, dy, rdy);
      fly[3] = qdiv(shl1(a3) - a3, dy, rdy);
    }
    // ---- upper x-face of row r (between r and r+1) from rows r-1 .. r+2
    derive(S1, dv[D1]);
    const double dx = M.x.d[i], rdx = M.x.rd[i];
    double dpx[2], dvx[2], flx[4];
    {
      const FaceGeom gx = face_geom(M.x, r + 1);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int o = s * 3;
        xvf[D1][s] = interp(gx, dv[D0][o + TFD_VX0], dv[D1][o + TFD_VX0]);
        xip[D1][s] = interp(gx, dv[D0][o + TFD_P0], dv[D1][o + TFD_P0]);
        dpx[s] = qdiv(xip[D1][s] - xip[D0][s], dx, rdx);
        dvx[s] = qdiv(xvf[D1][s] - xvf[D0][s], dx, rdx);
        const FaceDir fd = face_dir(xvf[D1][s], gx);
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const int k = s * 4 + qq;
          double d2;
          const double Pn = face_flux(Q(SM, k), Q(S0, k), Q(S1, k), Q(S2, k), gx, fd, d2);
          const double tx = qdiv(Pn - Px[k], dx, rdx);
          tdiv[k] = tx + tdiv[k];
          Px[k] = Pn;
        }
      }
      if (curl) {
        xfl[D1][0] = interp(gx, f[D0][TFF_BZ], f[D1][TFF_BZ]);
        xfl[D1][1] = interp(gx, f[D0][TFF_BYT], f[D1][TFF_BYT]);
        xfl[D1][2] = interp(gx, f[D0][TFF_EZ], f[D1][TFF_EZ]);
        xfl[D1][3] = interp(gx, f[D0][TFF_EY], f[D1][TFF_EY]);
#pragma unroll
        for (int k = 0; k < 4; ++k) flx[k] = qdiv(xfl[D1][k] - xfl[D0][k], dx, rdx);
      }
    }
    if (!out_lane) return;
    // ---- the cell (tf_rhs_cell)
    const size_t c = cell(i);
    double td[TQ], dpxv[2] = {0.0, 0.0}, dpyv[2] = {0.0, 0.0}, dvv[2] = {0.0, 0.0};
    double dbz_x = 0.0, dbz_y = 0.0, dby_x = 0.0, dbx_y = 0.0, dEz_x = 0.0, dEz_y = 0.0, dEy_x = 0.0, dEx_y = 0.0;
#pragma unroll
    for (int k = 0; k < TQ; ++k) td[k] = in ? tdiv[k] : 0.0;
    if (in) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        dpxv[s] = dpx[s];
        dpyv[s] = dpy[s];
        dvv[s] = dvx[s] + dvy[s];                                     // divergence2D(v)
      }
      if (curl) {
        dbz_x = flx[0]; dby_x = flx[1]; dEz_x = flx[2]; dEy_x = flx[3];
        dbz_y = fly[0]; dbx_y = fly[1]; dEz_y = fly[2]; dEx_y = fly[3];
      }
    }
    const double irho = Q(S0, 0), erho = Q(S0, 4);
    const double ivx = dv[D0][TFD_VX0], ivy = dv[D0][TFD_VY0];
    const double evx = dv[D0][TFD_VX1], evy = dv[D0][TFD_VY1];
    const double in_ = qdiv(irho, K.mi, K.rmi), en_ = qdiv(erho, K.me, K.rme);
    const double jx = in_ * E_CH * ivx - en_ * E_CH * evx;          // ideal2F.cpp:128-129
    const double jy = in_ * E_CH * ivy - en_ * E_CH * evy;
    const double bz = S.v[TF_BZ][c];
    const double Ex = S.v[TF_EX][c], Ey = S.v[TF_EY][c];
    const double ivb0 = ivy * bz, ivb1 = (-ivx) * bz, evb0 = evy * bz, evb1 = (-evx) * bz;
    const double iFx = E_CH * in_ * (Ex + qdiv(ivb0, C_L, K.rC));
    const double iFy = E_CH * in_ * (Ey + qdiv(ivb1, C_L, K.rC));
    const double eFx = -E_CH * en_ * (Ex + qdiv(evb0, C_L, K.rC));
    const double eFy = -E_CH * en_ * (Ey + qdiv(evb1, C_L, K.rC));
    const double ipc = dv[D0][TFD_P0], epc = dv[D0][TFD_P1];
    double k[TF_NE];
    k[TF_IRHO] = -td[0];
    k[TF_ERHO] = -td[4];
    k[TF_IMX] = -td[1] - dpxv[0] + irho * D.gx[c] + iFx;
    k[TF_IMY] = -td[2] - dpyv[0] + irho * D.gy[c] + iFy;
    k[TF_EMX] = -td[5] - dpxv[1] + erho * D.gx[c] + eFx;
    k[TF_EMY] = -td[6] - dpyv[1] + erho * D.gy[c] + eFy;
    k[TF_ITE] = -td[3] - ipc * dvv[0];
    k[TF_ETE] = -td[7] - epc * dvv[1];
    if (curl) {                                                    // ideal2F.cpp:75-80
      k[TF_EX] = C_L * dbz_y - 4. * PI * jx;
      k[TF_EY] = -C_L * dbz_x - 4. * PI * jy;
      k[TF_EZ] = C_L * (dby_x - dbx_y);
      k[TF_BX] = -C_L * dEz_y;
      k[TF_BY] = C_L * dEz_x;
      k[TF_BZ] = C_L * (dEx_y - dEy_x);
    } else {                                                       // :83-88
      k[TF_EX] = -4. * PI * jx;
      k[TF_EY] = -4. * PI * jy;
      k[TF_EZ] = 0.0;
      k[TF_BX] = 0.0;
      k[TF_BY] = 0.0;
      k[TF_BZ] = 0.0;
    }
    const double mask = in ? 1.0 : 0.0;                            // :94
#pragma unroll
    for (int v = 0; v < TF_NE; ++v) k[v] *= mask;
    if (mode == TF_STAGE_STORE_K) {
      for (int v = 0; v < TF_NE; ++v) D.k[kslot][v * N + c] = k[v];
    } else if (mode == TF_STAGE_RK4_FINAL) {                       // evolution.cpp:119-121
      for (int v = 0; v < TF_NE; ++v)
        k[v] = qdiv(D.k[0][v * N + c] + k[v], 6.0, K.r6) + qdiv(D.k[1][v * N + c] + D.k[2][v * N + c], 3.0, K.r3);
    }
    const double stp = mult * (*D.step);
    double nv[TF_NE];
#pragma unroll
    for (int v = 0; v < TF_NE; ++v) nv[v] = B.v[v][c] + k[v] * stp;
    // enforceMinimums (ideal2F.cpp:98-105)
    nv[TF_IRHO] = smax(qdiv(nv[TF_IRHO], K.mi, K.rmi), K.nmin) * K.mi;
    nv[TF_ERHO] = smax(qdiv(nv[TF_ERHO], K.me, K.rme), K.nmin) * K.me;
    nv[TF_ITE] = smax(nv[TF_ITE], K.temin);
    nv[TF_ETE] = smax(nv[TF_ETE], K.temin);
#pragma unroll
    for (int v = 0; v < TF_NE; ++v) T.v[v][c] = nv[v];
  };

  for (int r = r0; r < r0 + R; r += 4) {
    step(std::integral_constant<int, 0>{}, r);
    step(std::integral_constant<int, 1>{}, r + 1);
    step(std::integral_constant<int, 2>{}, r + 2);
    step(std::integral_constant<int, 3>{}, r + 3);
  }
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_tf_stage_strip(TfBatch b, int Ssel, int Bsel, int Tsel, double mult, int mode, int kslot) {
  __shared__ TfRows qs;
  tf_stage_strip_tile(b, strip::xcd_tile(blockIdx.x, b.nstrip_tiles), Ssel, Bsel, Tsel, mult, mode, kslot, qs);
}

// ---- persistent form: every phase of nsteps steps of the whole batch in one cooperative launch,
// one 64-lane wavefront per block, phases separated by grid barriers --------------------------

// dt minimum of 64x4 cell tile t (the batch's tile list) by one wavefront: lane = column, 4 rows
__device__ void tf_dt_tile_wave(const TfBatch& b, int t) {
  const int d = b.tile_dom[t];
  const TfDom& D = b.doms[d];
  const int lt = t - D.tile_base;
  const int ty = lt / D.tiles_x, tx = lt - ty * D.tiles_x;
  const int j = tx * 64 + (int)threadIdx.x;
  double v = DBL_MAX;
  for (int r = 0; r < 4; ++r) {
    const int i = ty * 4 + r;
    if (i < D.mesh.nx && j < D.mesh.ny && inside(D.mesh, i, j)) {
      const double dt = tf_dt_cell(D, D.L, i, j);
      v = (dt < v) ? dt : v;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  if (threadIdx.x == 0) D.partial[lt] = v;
}

// the step size of domain d from its tiles' minima (as k_tf_reduce), by one wavefront
__device__ void tf_reduce_wave(const TfBatch& b, int d) {
  const TfDom& D = b.doms[d];
  double v = DBL_MAX;
  const int n = D.tiles_x * D.tiles_y;
  for (int k = threadIdx.x; k < n; k += 64) { const double o = D.partial[k]; v = (o < v) ? o : v; }
  for (int off = 32; off > 0; off >>= 1) { const double o = __shfl_xor(v, off, 64); v = (o < v) ? o : v; }
  if (threadIdx.x == 0) {
    const double s = D.c.eps * v;
    *D.step = s;
    const int idx = *D.counter;
    if (idx < D.hist_cap) D.history[idx] = (v < DBL_MAX) ? s : __builtin_nan("");
    *D.counter = idx + 1;
  }
}

// the copy of set `from` into `to` over 64x4 tile t, by one wavefront
__device__ void tf_copy_tile_wave(const TfBatch& b, int t, int from, int to) {
  const int d = b.tile_dom[t];
  const TfDom& D = b.doms[d];
  const int lt = t - D.tile_base;
  const int ty = lt / D.tiles_x, tx = lt - ty * D.tiles_x;
  const int j = tx * 64 + (int)threadIdx.x;
  const TfSet& F = pick(D, from);
  const TfSet& T = pick(D, to);
  for (int r = 0; r < 4; ++r) {
    const int i = ty * 4 + r;
    if (i >= D.mesh.nx || j >= D.mesh.ny) continue;
    const size_t c = (size_t)i * D.mesh.ny + j;
#pragma unroll
    for (int v = 0; v < TF_NE; ++v) T.v[v][c] = F.v[v][c];
  }
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_tf_persistent(TfBatch b, int nsteps, int integrator, int ucnp_mask, int final_mask, int maxn) {
  __shared__ TfRows qs;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int G = (int)gridDim.x, bid = (int)blockIdx.x;
  // block bid takes strip tiles bid, bid + G, ...: with G a multiple of 8 the block's XCD (bid % 8)
  // is the XCD band of every tile it takes (strip::xcd_tile)
  auto stage = [&](int S, int B, int T, double mult, int mode, int kslot) {
    for (int t = bid; t < b.nstrip_tiles; t += G)
      tf_stage_strip_tile(b, strip::xcd_tile(t, b.nstrip_tiles), S, B, T, mult, mode, kslot, qs);
    grid.sync();
  };
  auto ghost = [&](int Gsel, int mask) {   // x sides, then y sides (launch_tf_ghost)
    if (!mask) return;
    for (int xs = 1; xs >= 0; --xs) {
      for (int w = bid * 64 + (int)threadIdx.x; w < b.ndom * maxn; w += G * 64) {
        const int d = w / maxn, t = w - d * maxn;
        const TfDom& D = b.doms[d];
        const TfSet& Gs = pick(D, Gsel);
        for (int s = xs ? 0 : 2; s < (xs ? 2 : 4); ++s) {
          const int kind = D.bc[s];
          if (kind != 0 && kind != 4 && ((mask >> kind) & 1)) tf_ghost_side(D, Gs, s, kind, t);
        }
      }
      grid.sync();
    }
  };
  // the stages of TfBatchRun::enqueue_step (one call site, so the kernel holds one copy of the
  // stage): Euler L->W; RK2 L->W, W->L; RK4 L->W, W->W2, W2->W, W->L (base L throughout)
  const int nops = integrator == SPRUCE_EULER ? 1 : (integrator == SPRUCE_RK2 ? 2 : 4);
  for (int n = 0; n < nsteps; ++n) {      // TfBatchRun::enqueue_step, phase by phase
    for (int o = 0; o < nops; ++o) {
      const bool last = o == nops - 1;
      const int S = o == 0 ? TF_SEL_L : (integrator == SPRUCE_RK4 && o == 2 ? TF_SEL_W2 : TF_SEL_W);
      const int T = integrator == SPRUCE_EULER ? TF_SEL_W
                  : (last ? TF_SEL_L : (integrator == SPRUCE_RK4 && o == 1 ? TF_SEL_W2 : TF_SEL_W));
      const double mult = (integrator == SPRUCE_EULER || o >= nops / 2) ? 1.0 : 0.5;
      const int mode = integrator != SPRUCE_RK4 ? TF_STAGE_PLAIN : (last ? TF_STAGE_RK4_FINAL : TF_STAGE_STORE_K);
      stage(S, TF_SEL_L, T, mult, mode, last ? 0 : o);
      if (!last) ghost(T, ucnp_mask);
    }
    if (integrator == SPRUCE_EULER) {
      for (int t = bid; t < b.ntiles; t += G) tf_copy_tile_wave(b, t, TF_SEL_W, TF_SEL_L);
      grid.sync();
    }
    ghost(TF_SEL_L, final_mask);
    for (int t = bid; t < b.ntiles; t += G) tf_dt_tile_wave(b, t);
    grid.sync();
    for (int d = bid; d < b.ndom; d += G) tf_reduce_wave(b, d);
    grid.sync();
  }
}
}  // namespace

void launch_tf_stage(const TfBatch& b, int S, int B, int T, double mult, int mode, int kslot, hipStream_t s) {
  if (b.strip_rows > 0)
    hipLaunchKernelGGL(k_tf_stage_strip, dim3(b.nstrip_tiles), dim3(64), 0, s, b, S, B, T, mult, mode, kslot);
  else
    hipLaunchKernelGGL(k_tf_stage, dim3(b.ntiles), dim3(b.bx, b.by), 0, s, b, S, B, T, mult, mode, kslot);
}

void launch_tf_ghost(const TfBatch& b, int G, int kinds_mask, int maxn, hipStream_t s) {
  hipLaunchKernelGGL(k_tf_ghost, dim3((maxn + 127) / 128, b.ndom), dim3(128), 0, s, b, G, kinds_mask, 1);
  hipLaunchKernelGGL(k_tf_ghost, dim3((maxn + 127) / 128, b.ndom), dim3(128), 0, s, b, G, kinds_mask, 0);
}

void launch_tf_dt(const TfBatch& b, hipStream_t s) {
  hipLaunchKernelGGL(k_tf_dt, dim3(b.ntiles), dim3(b.bx, b.by), 0, s, b);
  hipLaunchKernelGGL(k_tf_reduce, dim3(b.ndom), dim3(256), 0, s, b);
}

void launch_tf_populate(const TfBatch& b, const double* const* itemp, const double* const* etemp, hipStream_t s) {
  hipLaunchKernelGGL(k_tf_populate, dim3(b.ntiles), dim3(b.bx, b.by), 0, s, b, itemp, etemp);
}

void launch_tf_derived(const TfBatch& b, int d, int which, double* out, int nx, int ny, hipStream_t s) {
  dim3 blk(64, 4), g((ny + 63) / 64, (nx + 3) / 4);
  hipLaunchKernelGGL(k_tf_derived, g, blk, 0, s, b, d, which, out);
}

void launch_tf_rhs(const TfBatch& b, int d, double* out, int nx, int ny, hipStream_t s) {
  dim3 blk(64, 4), g((ny + 63) / 64, (nx + 3) / 4);
  hipLaunchKernelGGL(k_tf_rhs, g, blk, 0, s, b, d, out);
}

int tf_persistent_blocks() {
  int dev = 0, per_cu = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_tf_persistent),
                                                          64, 0),
            "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  return per_cu * prop.multiProcessorCount;
}

void launch_tf_persistent(const TfBatch& b, int nsteps, int integrator, int ucnp_mask, int final_mask, int maxn,
                          int blocks, hipStream_t s) {
  require(b.strip_rows > 0 && b.bx == 64 && b.by == 4, SPRUCE_E_STATE,
          "the persistent batch step needs a strip launch shape");
  require(blocks > 0, SPRUCE_E_STATE, "no co-resident blocks for the persistent batch step");
  void* args[] = {(void*)&b, (void*)&nsteps, (void*)&integrator, (void*)&ucnp_mask, (void*)&final_mask, (void*)&maxn};
  hip_check(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_tf_persistent), dim3(blocks), dim3(64), args, 0,
                                       s),
            "hipLaunchCooperativeKernel");
}

void launch_tf_reset_counters(const TfBatch& b, hipStream_t s) {
  hipLaunchKernelGGL(k_tf_reset, dim3((b.ndom + 63) / 64), dim3(64), 0, s, b);
}

void launch_tf_copy(const TfBatch& b, int from, int to, hipStream_t s) {
  hipLaunchKernelGGL(k_tf_copy, dim3(b.ntiles), dim3(b.bx, b.by), 0, s, b, from, to);
}

}  // namespace spruce
