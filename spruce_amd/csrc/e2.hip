// IdealMHD2E time-step kernels for gfx950 (source/equationsets/idealmhd2E.cpp), rounded exactly where
// the reference's whole-grid operations round, and the set-up kernels of IdealMHDCons.
// One thread per cell; each cell evaluates its four faces.
#include <cfloat>

#include "arith.hpp"
#include "crpow.hpp"
#include "e2.hpp"

namespace spruce {
namespace {

__device__ __forceinline__ int wr(int k, int n, int per) {
  if (per) {
    if (k < 0) k += n;
    else if (k >= n) k -= n;
  }
  return k;
}
__device__ __forceinline__ bool inside(const Mesh& M, int i, int j) {
  return i >= M.x.lo && i <= M.x.hi && j >= M.y.lo && j <= M.y.hi;
}

// Upwinded quantities in the order idealmhd2E.cpp:32-47 transports them.
enum { Q_RHO = 0, Q_MX, Q_MY, Q_TI, Q_TE, NQ2 };

struct Face2 {
  double P[NQ2];   // surf * face velocity (transportDerivative1D numerator terms, derivs.cpp:155)
  double vf;       // face velocity along the axis
  double ip;       // interpolated total pressure
  double ib;       // interpolated bi_y on x-faces, bi_x on y-faces (curl2D)
  double ize, izi; // interpolated v x b_e and v x b_i (curlZ of CrossProduct2D)
};

__device__ __forceinline__ void cell_vals(const E2Set& S, const E2Static& st, const MhdConst& K, size_t c,
                                          double& vx, double& vy, double& p, double& ze, double& zi) {
  const double r = S.rho[c];
  vx = S.mx[c] / r;
  vy = S.my[c] / r;
  p = S.tei[c] * K.gm1 + S.tee[c] * K.gm1;              // press = i_press + e_press (:88-90)
  ze = vx * st.bey[c] - vy * st.bex[c];                // CrossProduct2D(v, be) (grid.cpp:448-452)
  zi = vx * S.by[c] - vy * S.bx[c];
}

template <int AX>
__device__ __forceinline__ void face2(const E2Set& S, const E2Static& st, const MhdConst& K, const Mesh& M, int i,
                                      int j, Face2& F) {
  const Axis& ax = AX == 0 ? M.x : M.y;
  const int c = AX == 0 ? i : j;
  const int cm2 = wr(c - 2, ax.n, ax.per), cm1 = wr(c - 1, ax.n, ax.per), cp1 = wr(c + 1, ax.n, ax.per);
  auto at = [&](int k) -> size_t { return AX == 0 ? (size_t)k * M.ny + j : (size_t)i * M.ny + k; };
  const size_t i0 = at(cm2), i1 = at(cm1), i2 = at(c), i3 = at(cp1);
  const double hlo = ax.h[cm1], hup = ax.h[c], den = ax.fden[c], rden = ax.rfden[c];
  auto interp = [&](double a, double b) { return qdiv(a * hup + b * hlo, den, rden); };
  double vx1, vy1, p1, ze1, zi1, vx2, vy2, p2, ze2, zi2;
  cell_vals(S, st, K, i1, vx1, vy1, p1, ze1, zi1);
  cell_vals(S, st, K, i2, vx2, vy2, p2, ze2, zi2);
  const double vf = AX == 0 ? interp(vx1, vx2) : interp(vy1, vy2);
  F.vf = vf;
  F.ip = interp(p1, p2);
  F.ib = AX == 0 ? interp(S.by[i1], S.by[i2]) : interp(S.bx[i1], S.bx[i2]);
  F.ize = interp(ze1, ze2);
  F.izi = interp(zi1, zi2);
  const bool pos = vf > 0.0, neg = vf < 0.0;
  const double g = pos ? ax.gup[c] : ax.gdn[c];
  const int ce = pos ? cm1 : cp1;
  const double eden = ax.fden[ce], reden = ax.rfden[ce];
  const double* src[NQ2] = {S.rho, S.mx, S.my, S.tei, S.tee};
#pragma unroll
  for (int q = 0; q < NQ2; ++q) {
    const double q0 = src[q][i0], q1 = src[q][i1], q2 = src[q][i2], q3 = src[q][i3];
    const double d2 = interp(q1, q2);
    const double a = pos ? q0 : q3, b = pos ? q1 : q2;
    const double d1 = a + qdiv((b - a) * g, eden, reden);     // boundaryExtrapolate (derivs.cpp:490-499)
    const bool le = q2 <= q1;
    const bool outer_min = pos ? le : !le;                     // Barton limiter (derivs.cpp:47-68)
    const double inner = outer_min ? smax(d1, d2) : smin(d1, d2);
    const double lim = outer_min ? smin(b, inner) : smax(b, inner);
    F.P[q] = ((pos || neg) ? lim : d2) * vf;
  }
}

// IdealMHD2E::computeTimeDerivativesDerived at one cell, ghost mask applied (idealmhd2E.cpp:22-59).
__device__ void e2_rhs_cell(const E2Set& S, const E2Static& st, const MhdConst& K, const Mesh& M, int i, int j,
                            double k[7]) {
  const size_t c = (size_t)i * M.ny + j;
  const bool in = inside(M, i, j);
  double td[NQ2] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double dpx = 0.0, dpy = 0.0, divv = 0.0, dby_x = 0.0, dbx_y = 0.0;
  double dze_x = 0.0, dze_y = 0.0, dzi_x = 0.0, dzi_y = 0.0;
  if (in) {  // every derivative is 0 outside the interior (derivs.cpp:229-233)
    const double dx = M.x.d[i], rdx = M.x.rd[i], dy = M.y.d[j], rdy = M.y.rd[j];
    double tx[NQ2];
    {
      Face2 lo, up;
      face2<0>(S, st, K, M, i, j, lo);
      face2<0>(S, st, K, M, wr(i + 1, M.x.n, M.x.per), j, up);
#pragma unroll
      for (int q = 0; q < NQ2; ++q) tx[q] = qdiv(up.P[q] - lo.P[q], dx, rdx);
      dpx = qdiv(up.ip - lo.ip, dx, rdx);
      divv = qdiv(up.vf - lo.vf, dx, rdx);
      dby_x = qdiv(up.ib - lo.ib, dx, rdx);
      dze_x = qdiv(up.ize - lo.ize, dx, rdx);
      dzi_x = qdiv(up.izi - lo.izi, dx, rdx);
    }
    {
      Face2 lo, up;
      face2<1>(S, st, K, M, i, j, lo);
      face2<1>(S, st, K, M, i, wr(j + 1, M.y.n, M.y.per), up);
#pragma unroll
      for (int q = 0; q < NQ2; ++q) td[q] = tx[q] + qdiv(up.P[q] - lo.P[q], dy, rdy);
      dpy = qdiv(up.ip - lo.ip, dy, rdy);
      divv = divv + qdiv(up.vf - lo.vf, dy, rdy);
      dbx_y = qdiv(up.ib - lo.ib, dy, rdy);
      dze_y = qdiv(up.ize - lo.ize, dy, rdy);
      dzi_y = qdiv(up.izi - lo.izi, dy, rdy);
    }
  }
  const double mask = in ? 1.0 : 0.0;
  const double rho = S.rho[c];
  const double bix = S.bx[c], biy = S.by[c], bex = st.bex[c], bey = st.bey[c];
  const double curl_db = qdiv(dby_x - dbx_y, 4.0 * PI, K.r4pi);          // :35
  const double ext0 = (-curl_db) * bey, ext1 = curl_db * bex;            // CrossProductZ2D (grid.cpp:455-460)
  const double in0 = (-curl_db) * biy, in1 = curl_db * bix;
  const double ip = S.tei[c] * K.gm1, ep = S.tee[c] * K.gm1;
  k[0] = -td[Q_RHO];
  k[1] = -td[Q_MX] - dpx + mask * (rho * st.gx[c] + ext0 + in0);
  k[2] = -td[Q_MY] - dpy + mask * (rho * st.gy[c] + ext1 + in1);
  k[3] = -td[Q_TI] - ip * divv;
  k[4] = -td[Q_TE] - ep * divv;
  k[5] = dze_y + dzi_y;                                                  // curlZ: {d/dy, -d/dx}
  k[6] = (-dze_x) + (-dzi_x);
#pragma unroll
  for (int v = 0; v < 7; ++v) k[v] *= mask;
}

// EICThermalization::computeTimeDerivativesModule (eic_thermalization.cpp:27-43) at one cell:
// electron-ion energy exchange at the collision rate of the electron coupling parameter, moved
// from the electron to the ion thermal energy (times the ghost-zone mask).  x^(1/3) and x^(3/2) are
// the correctly rounded powers; log is the device logarithm.
__device__ __forceinline__ void eic_apply(const E2Set& S, const MhdConst& K, double nu_c, size_t c, double mask,
                                          double k[7]) {
  const double n = S.n[c];
  const double te_e = smax(S.tee[c] * K.gm1 / (K_B * n), K.tmin);           // e_temp (idealmhd2E.cpp:93)
  const double a = pow_third(3.0 / 4.0 / PI / n);
  const double w_pe = sqrt(4.0 * PI * E_CHARGE * E_CHARGE / M_ELECTRON * n);
  const double gam = E_CHARGE * E_CHARGE / K_B / te_e / a;
  const double g15 = pow_1p5(gam);
  const double lam = 1.0 / sqrt(3.0) / g15;
  const double gam_ei = sqrt(2.0 / 3.0 / PI) * g15 * w_pe * log(lam);
  const double dedt = nu_c * gam_ei * (S.tee[c] - S.tei[c]);
  k[4] = k[4] - dedt * mask;
  k[3] = k[3] + dedt * mask;
}

__global__ void __launch_bounds__(256) k_e2_stage(E2StageArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const Mesh& M = a.mesh;
  if (i >= M.nx || j >= M.ny) return;
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  double k[7];
  e2_rhs_cell(a.S, a.st, a.k, M, i, j, k);
  for (int e = 0; e < a.eic; ++e) eic_apply(a.S, a.k, a.eic_nu, c, inside(M, i, j) ? 1.0 : 0.0, k);
  if (a.mode == STAGE_STORE_K) {
#pragma unroll
    for (int v = 0; v < 7; ++v) a.kbuf[v * N + c] = k[v];
  } else if (a.mode == STAGE_RK4_FINAL) {
#pragma unroll
    for (int v = 0; v < 7; ++v)
      k[v] = qdiv(a.kbuf[v * N + c] + k[v], 6.0, a.k.r6) + qdiv(a.k2buf[v * N + c] + a.k3buf[v * N + c], 3.0, a.k.r3);
  }
  const MhdConst& K = a.k;
  const double h = a.step_mult * (*a.step);
  const double rho_new = a.B.rho[c] + k[0] * h;
  // enforceMinimums (idealmhd2E.cpp:61-67) then recomputeDerivedVarsFromEvolvedVars (:79-90)
  const double rho1 = smax(qdiv(rho_new, K.m, K.rm), K.nmin) * K.m;
  const double n = smax(qdiv(rho1, K.m, K.rm), K.nmin);
  a.T.rho[c] = n * K.m;
  a.T.n[c] = n;
  if (a.write_pre) a.T.rho_pre[c] = rho1;
  a.T.mx[c] = a.B.mx[c] + k[1] * h;
  a.T.my[c] = a.B.my[c] + k[2] * h;
  a.T.tei[c] = smax(smax(a.B.tei[c] + k[3] * h, K.temin), K.temin);
  a.T.tee[c] = smax(smax(a.B.tee[c] + k[4] * h, K.temin), K.temin);
  a.T.bx[c] = a.B.bx[c] + k[5] * h;
  a.T.by[c] = a.B.by[c] + k[6] * h;
}

// populateVariablesFromState up to the ghost fill (equationset.cpp:96-104; idealmhd2E.cpp:61-90)
__global__ void __launch_bounds__(256) k_e2_populate(E2Set L, const double* ti, const double* te, MhdConst K,
                                                     size_t N) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double rho = L.rho[c];
  const double n0 = smax(qdiv(rho, K.m, K.rm), K.nmin);
  const double ipr = n0 * K_B * smax(ti[c], K.tmin), epr = n0 * K_B * smax(te[c], K.tmin);
  const double ite = smax(qdiv(ipr, K.gm1, K.rgm1), K.temin), ete = smax(qdiv(epr, K.gm1, K.rgm1), K.temin);
  const double rho1 = smax(qdiv(rho, K.m, K.rm), K.nmin) * K.m;
  const double n = smax(qdiv(rho1, K.m, K.rm), K.nmin);
  L.rho_pre[c] = rho1;
  L.n[c] = n;
  L.rho[c] = n * K.m;
  L.tei[c] = smax(smax(ite, K.temin), K.temin);
  L.tee[c] = smax(smax(ete, K.temin), K.temin);
}

// Ghost fill of one side (evolution.cpp:126-334) for IdealMHD2E, whose densities() and momenta()
// list the single fluid twice (idealmhd2E.hpp:37-39): the repeated writes are identical, and the
// open boundary takes the larger of the ion and electron sound speeds.  Writes happen within a
// propagate, so the derived update of the written ghost cells follows.
__device__ void e2_ghost_side(E2Set G, const MhdConst& K, const Mesh& M, int side, int kind, int t, double s1,
                              double s2) {
  const bool xside = side < 2, upper = side & 1;
  const Axis& ax = xside ? M.x : M.y;
  const Axis& ot = xside ? M.y : M.x;
  if (kind == 2) { if (t < 0 || t >= ot.n) return; }
  else if (t < ot.lo || t > ot.hi) return;
  const int n = ax.n;
  const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
  auto at = [&](int e) -> size_t { return xside ? (size_t)e * M.ny + t : (size_t)t * M.ny + e; };
  const size_t c1 = at(e1), c2 = at(e2), c3 = at(e3);
  if (kind == 1) {                                                   // open (:158-224)
    const double rho3 = G.rho_pre[c3];
    const double ti3 = G.tei[c3], te3 = G.tee[c3];
    const double rg1 = s1 * rho3, rg2 = s2 * rho3;
    double cs = 0.0;
    const double ci = sqrt(K.gamma * (ti3 * K.gm1) / rho3);
    if (ci > cs) cs = ci;
    const double ce = sqrt(K.gamma * (te3 * K.gm1) / rho3);
    if (ce > cs) cs = ce;
    const double vx = G.mx[c3] / rho3, vy = G.my[c3] / rho3;
    double boost = K.obs * cs;
    if (!upper) boost *= -1.0;
    const double dist = 0.5 * (ax.d[e2] + ax.d[e3]);
    const double vn = xside ? vx : vy;
    const double bv = upper ? smax(0.0, vn + boost) : smin(0.0, vn + boost);
    const double gv = (dist * bv - 0.5 * ax.d[e2] * vn) / (0.5 * ax.d[e3]);
    const double n1 = smax(qdiv(rg1, K.m, K.rm), K.nmin), n2 = smax(qdiv(rg2, K.m, K.rm), K.nmin);
    G.n[c1] = n1; G.n[c2] = n2;
    G.rho[c1] = n1 * K.m; G.rho[c2] = n2 * K.m;
    G.rho_pre[c1] = rg1; G.rho_pre[c2] = rg2;
    G.tei[c1] = smax(s1 * ti3, K.temin); G.tei[c2] = smax(s2 * ti3, K.temin);
    G.tee[c1] = smax(s1 * te3, K.temin); G.tee[c2] = smax(s2 * te3, K.temin);
    G.mx[c1] = rg1 * (xside ? gv : vx); G.mx[c2] = rg2 * (xside ? gv : vx);
    G.my[c1] = rg1 * (xside ? vy : gv); G.my[c2] = rg2 * (xside ? vy : gv);
  } else if (kind == 3) {                                            // reflect (:231-266)
    const double ti3 = G.tei[c3], te3 = G.tee[c3], r3 = G.rho[c3], n3 = G.n[c3], rp3 = G.rho_pre[c3];
    G.tei[c1] = ti3; G.tei[c2] = ti3; G.tee[c1] = te3; G.tee[c2] = te3;
    G.rho[c1] = r3; G.rho[c2] = r3; G.n[c1] = n3; G.n[c2] = n3; G.rho_pre[c1] = rp3; G.rho_pre[c2] = rp3;
    G.mx[c1] = 0.0; G.mx[c2] = 0.0; G.mx[c3] = 0.0;
    G.my[c1] = 0.0; G.my[c2] = 0.0; G.my[c3] = 0.0;
  } else if (kind == 2) {                                            // fixed (:272-282)
    G.mx[c1] = 0.0; G.mx[c2] = 0.0; G.mx[c3] = 0.0;
    G.my[c1] = 0.0; G.my[c2] = 0.0; G.my[c3] = 0.0;
  } else if (kind == 5) {                                            // open_ucnp (:290-334)
    const size_t ci = at(upper ? ax.hi : ax.lo);
    for (int g = 0; g < N_GHOST; ++g) {
      const size_t cb = at(upper ? ax.hi + 1 + g : g);
      G.rho[cb] = G.rho[ci]; G.n[cb] = G.n[ci]; G.rho_pre[cb] = G.rho_pre[ci];
      G.tei[cb] = G.tei[ci]; G.tee[cb] = G.tee[ci];
      G.bx[cb] = G.bx[ci]; G.by[cb] = G.by[ci];
      G.mx[cb] = G.mx[ci]; G.my[cb] = G.my[ci];
    }
  }
}

struct E2GhostArgs {
  E2Set G;
  MhdConst k;
  Mesh mesh;
  int kind[4];
  double s1[4], s2[4];
};

__global__ void __launch_bounds__(256) k_e2_ghost(E2GhostArgs a, int xsides) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s0 = xsides ? 0 : 2;
  for (int s = s0; s < s0 + 2; ++s)
    if (a.kind[s] != 0 && a.kind[s] != 4) e2_ghost_side(a.G, a.k, a.mesh, s, a.kind[s], t, a.s1[s], a.s2[s]);
}

// recomputeDT of IdealMHD2E / IdealMHDCons at one cell (idealmhd2E.cpp:115-137): |v| by abs.
__device__ __forceinline__ double dt_abs(double press, double rho, double bmag, double vx, double vy,
                                         const MhdConst& K, const Mesh& M, int i, int j) {
  const double cs = sqrt(K.gamma * press / rho);
  const double cs2 = cs * cs;
  const double va = bmag / sqrt(4.0 * PI * rho);
  const double va2 = va * va;
  const double sum = cs2 + va2;
  const double delta = sqrt(1.0 - 4.0 * cs2 * va2 / (sum * sum));
  const double vfast = sqrt(0.5 * sum * (1.0 + delta));
  const double vslow = sqrt(0.5 * sum * (1.0 - delta));
  const double cmax = smax(smax(smax(cs, va), vfast), vslow);
  return 1. / (qdiv(fabs(vx) + cmax, M.x.d[i], M.x.rd[i]) + qdiv(fabs(vy) + cmax, M.y.d[j], M.y.rd[j]));
}

__device__ __forceinline__ double e2_dt_cell(const E2Set& L, const E2Static& st, const MhdConst& K, const Mesh& M,
                                             int i, int j) {
  const size_t c = (size_t)i * M.ny + j;
  const double rho = L.rho[c];
  const double press = L.tei[c] * K.gm1 + L.tee[c] * K.gm1;
  const double bx = st.bex[c] + L.bx[c], by = st.bey[c] + L.by[c];
  return dt_abs(press, rho, sqrt(bx * bx + by * by), L.mx[c] / rho, L.my[c] / rho, K, M, i, j);
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

__global__ void __launch_bounds__(256) k_e2_dt(E2Set L, E2Static st, MhdConst K, Mesh M, double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  double v = DBL_MAX;
  if (i < M.nx && j < M.ny && inside(M, i, j)) {
    const double dt = e2_dt_cell(L, st, K, M, i, j);
    v = (dt < v) ? dt : v;
  }
  v = block_min(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = v;
}

__global__ void __launch_bounds__(256) k_e2_derived(E2Set S, E2Static st, MhdConst K, Mesh M, int which,
                                                    double* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double rho = S.rho[c], n = S.n[c];
  const double vx = S.mx[c] / rho, vy = S.my[c] / rho;
  const double ip = S.tei[c] * K.gm1, ep = S.tee[c] * K.gm1;
  const double bx = st.bex[c] + S.bx[c], by = st.bey[c] + S.by[c];
  const double bm = sqrt(bx * bx + by * by);
  double r = 0.0;
  switch (which) {
    case 0: r = smax(ip / (K_B * n), K.tmin); break;         // i_temp
    case 1: r = smax(ep / (K_B * n), K.tmin); break;         // e_temp
    case 2: r = ip; break;
    case 3: r = ep; break;
    case 4: r = ip + ep; break;                               // press
    case 5: r = vx; break;
    case 6: r = vy; break;
    case 7: r = 0.5 * rho * (vx * vx + vy * vy); break;       // kinetic_energy
    case 8: r = bx; break;
    case 9: r = by; break;
    case 10: r = bm; break;
    case 11: r = bm == 0.0 ? 0.0 : bx / bm; break;
    case 12: r = bm == 0.0 ? 0.0 : by / bm; break;
    case 13: r = e2_dt_cell(S, st, K, M, i, j); break;
  }
  out[c] = r;
}

__global__ void __launch_bounds__(256) k_e2_rhs(E2Set S, E2Static st, MhdConst K, Mesh M, double* out, int eic,
                                                double eic_nu) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  double k[7];
  e2_rhs_cell(S, st, K, M, i, j, k);
  for (int e = 0; e < eic; ++e) eic_apply(S, K, eic_nu, c, inside(M, i, j) ? 1.0 : 0.0, k);
  for (int v = 0; v < 7; ++v) out[v * N + c] = k[v];
}

// ---------------------------------------------------------------------------------------------
// IdealMHDCons set-up
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cons_populate(ConsGrids g, MhdConst K, size_t N) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  // recomputeEvolvedVarsFromStateVars (idealmhdcons.cpp:96-109)
  const double rho = g.rho[c];
  const double n = smax(qdiv(rho, K.m, K.rm), K.nmin);
  const double press = 2 * n * K_B * smax(g.temp[c], K.tmin);
  const double te = qdiv(press, K.gm1, K.rgm1);
  const double vx = g.mx[c] / rho, vy = g.my[c] / rho;
  const double ke = 0.5 * rho * (vx * vx + vy * vy);
  const double bx = g.bex[c] + g.bx[c], by = g.bey[c] + g.by[c];
  const double bm = sqrt(bx * bx + by * by);
  const double mag = bm * bm / (8. * PI);
  g.n[c] = n; g.press[c] = press; g.vx[c] = vx; g.vy[c] = vy; g.ke[c] = ke; g.b_x[c] = bx; g.b_y[c] = by;
  g.bmag[c] = bm; g.mag[c] = mag;
  g.energy[c] = te + ke + mag;
  // enforceMinimums (:89-94); rho_pre holds the density the ghost fill and the derived update read
  g.rho_pre[c] = smax(qdiv(rho, K.m, K.rm), K.nmin) * K.m;
  g.te[c] = smax(te, K.temin);
}

__device__ void cons_ghost_side(ConsGrids G, const MhdConst& K, const Mesh& M, int side, int kind, int t,
                                double s1, double s2) {
  const bool xside = side < 2, upper = side & 1;
  const Axis& ax = xside ? M.x : M.y;
  const Axis& ot = xside ? M.y : M.x;
  if (kind == 2) { if (t < 0 || t >= ot.n) return; }
  else if (t < ot.lo || t > ot.hi) return;
  const int n = ax.n;
  const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
  auto at = [&](int e) -> size_t { return xside ? (size_t)e * M.ny + t : (size_t)t * M.ny + e; };
  const size_t c1 = at(e1), c2 = at(e2), c3 = at(e3);
  if (kind == 1) {
    const double rho3 = G.rho_pre[c3], te3 = G.te[c3];
    G.rho_pre[c1] = s1 * rho3; G.rho_pre[c2] = s2 * rho3;
    G.te[c1] = s1 * te3; G.te[c2] = s2 * te3;
    double cs = 0.0;
    const double cn = sqrt(K.gamma * (te3 * K.gm1) / rho3);
    if (cn > cs) cs = cn;
    const double vx = G.mx[c3] / rho3, vy = G.my[c3] / rho3;
    double boost = K.obs * cs;
    if (!upper) boost *= -1.0;
    const double dist = 0.5 * (ax.d[e2] + ax.d[e3]);
    const double vn = xside ? vx : vy;
    const double bv = upper ? smax(0.0, vn + boost) : smin(0.0, vn + boost);
    const double gv = (dist * bv - 0.5 * ax.d[e2] * vn) / (0.5 * ax.d[e3]);
    const double r1 = G.rho_pre[c1], r2 = G.rho_pre[c2];
    G.mx[c1] = r1 * (xside ? gv : vx); G.mx[c2] = r2 * (xside ? gv : vx);
    G.my[c1] = r1 * (xside ? vy : gv); G.my[c2] = r2 * (xside ? vy : gv);
  } else if (kind == 3) {
    const double te3 = G.te[c3], r3 = G.rho_pre[c3];
    G.te[c1] = te3; G.te[c2] = te3;
    G.rho_pre[c1] = r3; G.rho_pre[c2] = r3;
    G.mx[c1] = 0.0; G.mx[c2] = 0.0; G.mx[c3] = 0.0;
    G.my[c1] = 0.0; G.my[c2] = 0.0; G.my[c3] = 0.0;
  } else if (kind == 2) {
    G.mx[c1] = 0.0; G.mx[c2] = 0.0; G.mx[c3] = 0.0;
    G.my[c1] = 0.0; G.my[c2] = 0.0; G.my[c3] = 0.0;
  } else if (kind == 5) {
    const size_t ci = at(upper ? ax.hi : ax.lo);
    for (int g = 0; g < N_GHOST; ++g) {
      const size_t cb = at(upper ? ax.hi + 1 + g : g);
      G.rho_pre[cb] = G.rho_pre[ci]; G.te[cb] = G.te[ci];
      G.bx[cb] = G.bx[ci]; G.by[cb] = G.by[ci];
      G.mx[cb] = G.mx[ci]; G.my[cb] = G.my[ci];
    }
  }
}

struct ConsGhostArgs {
  ConsGrids G;
  MhdConst k;
  Mesh mesh;
  int kind[4];
  double s1[4], s2[4];
};

__global__ void __launch_bounds__(256) k_cons_ghost(ConsGhostArgs a, int xsides) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s0 = xsides ? 0 : 2;
  for (int s = s0; s < s0 + 2; ++s)
    if (a.kind[s] != 0 && a.kind[s] != 4) cons_ghost_side(a.G, a.k, a.mesh, s, a.kind[s], t, a.s1[s], a.s2[s]);
}

// recomputeDerivedVarsFromEvolvedVars (idealmhdcons.cpp:111-128) and recomputeDT (:54-73)
__global__ void __launch_bounds__(256) k_cons_derived(ConsGrids g, MhdConst K, Mesh M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double n = smax(qdiv(g.rho_pre[c], K.m, K.rm), K.nmin);
  const double rho = n * K.m;
  const double vx = g.mx[c] / rho, vy = g.my[c] / rho;
  const double ke = 0.5 * rho * (vx * vx + vy * vy);
  const double bx = g.bex[c] + g.bx[c], by = g.bey[c] + g.by[c];
  const double bm = sqrt(bx * bx + by * by);
  const double mag = bm * bm / (8. * PI);
  const double te = g.energy[c] - mag - smax(ke, K.temin);   // the floor binds to the kinetic term only
  const double press = K.gm1 * te;
  g.n[c] = n; g.rho[c] = rho; g.vx[c] = vx; g.vy[c] = vy; g.ke[c] = ke; g.b_x[c] = bx; g.b_y[c] = by;
  g.bmag[c] = bm; g.mag[c] = mag;
  g.bhx[c] = bm == 0.0 ? 0.0 : bx / bm;
  g.bhy[c] = bm == 0.0 ? 0.0 : by / bm;
  g.te[c] = te;
  g.press[c] = press;
  g.temp[c] = smax(press / (2 * K_B * n), K.tmin);
  g.dt[c] = dt_abs(press, rho, bm, vx, vy, K, M, i, j);
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

void launch_e2_stage(const E2StageArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_e2_stage, g2(a.mesh, dim3(64, 4)), dim3(64, 4), 0, s, a);
}
void launch_e2_populate(E2Set L, const double* ti, const double* te, MhdConst k, int nx, int ny, hipStream_t s) {
  const size_t N = (size_t)nx * ny;
  hipLaunchKernelGGL(k_e2_populate, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, L, ti, te, k, N);
}
void launch_e2_ghost(E2Set G, MhdConst k, Mesh mesh, const int* kinds, const double* s1, const double* s2,
                     hipStream_t s) {
  E2GhostArgs a;
  a.G = G; a.k = k; a.mesh = mesh;
  for (int q = 0; q < 4; ++q) { a.kind[q] = kinds[q]; a.s1[q] = s1[q]; a.s2[q] = s2[q]; }
  if (a.kind[0] || a.kind[1]) hipLaunchKernelGGL(k_e2_ghost, dim3((mesh.ny + 255) / 256), dim3(256), 0, s, a, 1);
  if (a.kind[2] || a.kind[3]) hipLaunchKernelGGL(k_e2_ghost, dim3((mesh.nx + 255) / 256), dim3(256), 0, s, a, 0);
}
void launch_e2_dt(E2Set L, E2Static st, MhdConst k, Mesh mesh, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_e2_dt, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, st, k, mesh, partial);
}
void launch_e2_derived(E2Set S, E2Static st, MhdConst k, Mesh mesh, int which, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_e2_derived, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, S, st, k, mesh, which, out);
}
void launch_e2_rhs(E2Set S, E2Static st, MhdConst k, Mesh mesh, double* out, hipStream_t s, int eic, double eic_nu) {
  hipLaunchKernelGGL(k_e2_rhs, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, S, st, k, mesh, out, eic, eic_nu);
}
void launch_cons_populate(ConsGrids g, MhdConst k, int nx, int ny, hipStream_t s) {
  const size_t N = (size_t)nx * ny;
  hipLaunchKernelGGL(k_cons_populate, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, g, k, N);
}
void launch_cons_ghost(ConsGrids g, MhdConst k, Mesh mesh, const int* kinds, const double* s1, const double* s2,
                       hipStream_t s) {
  ConsGhostArgs a;
  a.G = g; a.k = k; a.mesh = mesh;
  for (int q = 0; q < 4; ++q) { a.kind[q] = kinds[q]; a.s1[q] = s1[q]; a.s2[q] = s2[q]; }
  if (a.kind[0] || a.kind[1]) hipLaunchKernelGGL(k_cons_ghost, dim3((mesh.ny + 255) / 256), dim3(256), 0, s, a, 1);
  if (a.kind[2] || a.kind[3]) hipLaunchKernelGGL(k_cons_ghost, dim3((mesh.nx + 255) / 256), dim3(256), 0, s, a, 0);
}
void launch_cons_derived(ConsGrids g, MhdConst k, Mesh mesh, hipStream_t s) {
  hipLaunchKernelGGL(k_cons_derived, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, g, k, mesh);
}

}  // namespace spruce
