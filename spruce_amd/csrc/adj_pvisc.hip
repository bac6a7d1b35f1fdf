// Discrete adjoint of the physical (Braginskii) viscosity module (physicalviscosity.cpp:47-250;
// forward in pvisc.hip): the subcycle updates of thermal energy and momenta reversed through the
// heating and the viscous force at every cell, the velocity derivatives and T^2.5 they read, the
// derived velocities and temperature, and the frozen b_hat.  The reverse scatters onto the cell and
// its four neighbours with FP64 atomics.
#include <cfloat>

#include "adj_pvisc.hpp"
#include "arith.hpp"
#include "crpow.hpp"

namespace spruce {
namespace {

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
__device__ __forceinline__ double fint(const Axis& A, int up, int lo, double qlo, double qup) {
  return qdiv(qlo * A.h[up] + qup * A.h[lo], A.fden[up], A.rfden[up]);
}
__device__ __forceinline__ double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double part[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  if ((lane & 63) == 0) part[lane >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    for (int k = 0; k < nw; ++k) s += part[k];
  }
  return s;
}

// Weights of derivative1D (D) and secondDerivative1D (S2) along one axis at cell c with respect to
// the values at c-1, c, c+1 (derivs.cpp:229-264, 450-462, 477-487).
struct W3 { double m, c, p; };
__device__ __forceinline__ W3 d_weights(const Axis& A, int c, int cm, int cp) {
  const double up_c = A.h[cp] / A.fden[cp], up_p = A.h[c] / A.fden[cp];
  const double lo_m = A.h[c] / A.fden[c], lo_c = A.h[cm] / A.fden[c];
  return W3{-lo_m / A.d[c], (up_c - lo_c) / A.d[c], up_p / A.d[c]};
}
__device__ __forceinline__ W3 s2_weights(const Axis& A, int c, int cm, int cp) {
  const double up_c = A.h[cp] / A.fden[cp], up_p = A.h[c] / A.fden[cp];
  const double lo_m = A.h[c] / A.fden[c], lo_c = A.h[cm] / A.fden[c];
  return W3{lo_m / A.hsq[c], (up_c - 2.0 + lo_c) / A.hsq[c], up_p / A.hsq[c]};
}

// stencil positions: 0 the cell, 1/2 its lower/upper x neighbours, 3/4 its lower/upper y neighbours
enum { P_C = 0, P_XM, P_XP, P_YM, P_YP, NPOS };
// velocity-derivative components in PvDerivs order
enum { D_XVX = 0, D_XVY, D_XVZ, D_YVX, D_YVY, D_YVZ };

__device__ __forceinline__ double dget(const PvDerivs& d, int k, size_t c) {
  const double* g[6] = {d.xvx, d.xvy, d.xvz, d.yvx, d.yvy, d.yvz};
  return g[k][c];
}
__device__ __forceinline__ double sfo_of(double bx, double by, double bz, const double* dv) {
  return bx * (by * dv[D_YVX]) + by * (bx * dv[D_XVY]) + bz * (bx * dv[D_XVZ] + by * dv[D_YVZ]) -
         (dv[D_XVX] + dv[D_YVY]) / 3.0;
}

__global__ void __launch_bounds__(256) k_adj_pv_apply(PvArgs a, PvLam L) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const Mesh& M = a.mesh;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const bool in = inside(M, i, j);
    const double mask = in ? 1.0 : 0.0;
    const PvDerivs& d = a.d;
    const PvFrozen& f = a.f;
    const MhdConst& K = a.k;
    const bool gc = a.grad_corr != 0;
    const double mult = a.mode == PV_HALF ? 0.5 : 1.0;
    // cell and neighbour indices (the forward takes the cell itself outside the interior)
    int im = i, ip = i, jm = j, jp = j;
    if (in) {
      im = wrp(i - 1, M.nx, M.x.per); ip = wrp(i + 1, M.nx, M.x.per);
      jm = wrp(j - 1, M.ny, M.y.per); jp = wrp(j + 1, M.ny, M.y.per);
    }
    const size_t pc[NPOS] = {c, (size_t)im * M.ny + j, (size_t)ip * M.ny + j, (size_t)i * M.ny + jm,
                             (size_t)i * M.ny + jp};
    const double bx = f.bhx[c], by = f.bhy[c], bz = f.bhz[c];
    double dvc[6];
    for (int k = 0; k < 6; ++k) dvc[k] = dget(d, k, c);
    const double t25 = d.t25[c], coef = f.coef[c];
    // ---- forward: heating (:47-64) ----
    double heat = 0.0, S = 0.0;
    if (a.heating_on && !a.coeff_zero) {
      S = bx * (bx * dvc[D_XVX] + by * dvc[D_YVX]) + by * (bx * dvc[D_XVY] + by * dvc[D_YVY]) +
          bz * (bx * dvc[D_XVZ] + by * dvc[D_YVZ]) - (dvc[D_XVX] + dvc[D_YVY]) / 3.0;
      heat = mask * (3.0 * coef * t25 * (S * S));
    }
    // ---- forward: force (:82-140), keeping what its reverse needs ----
    const double* bh[3] = {f.bhx, f.bhy, f.bhz};
    const double bc[3] = {bx, by, bz};
    W3 wx{0.0, 0.0, 0.0}, wy{0.0, 0.0, 0.0}, sx2{0.0, 0.0, 0.0}, sy2{0.0, 0.0, 0.0};
    double fr[3] = {0.0, 0.0, 0.0};
    double dbx_x = 0.0, dby_x = 0.0, dbx_y = 0.0, dby_y = 0.0, sxx = 0.0, syy = 0.0, x_yvy = 0.0, y_xvx = 0.0;
    double sfd = 0.0, gbd[2] = {0.0, 0.0};
    double sfo[NPOS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double dvn[NPOS][6];
    if (a.force_on && in) {
      wx = d_weights(M.x, i, im, ip);
      wy = d_weights(M.y, j, jm, jp);
      sx2 = s2_weights(M.x, i, im, ip);
      sy2 = s2_weights(M.y, j, jm, jp);
      const double dx = M.x.d[i], rdx = M.x.rd[i], dy = M.y.d[j], rdy = M.y.rd[j];
      auto ddx = [&](double qm, double qc, double qp) {
        return qdiv(fint(M.x, ip, i, qc, qp) - fint(M.x, i, im, qm, qc), dx, rdx);
      };
      auto ddy = [&](double qm, double qc, double qp) {
        return qdiv(fint(M.y, jp, j, qc, qp) - fint(M.y, j, jm, qm, qc), dy, rdy);
      };
      dbx_x = ddx(f.bhx[pc[P_XM]], bx, f.bhx[pc[P_XP]]);
      dby_x = ddx(f.bhy[pc[P_XM]], by, f.bhy[pc[P_XP]]);
      dbx_y = ddy(f.bhx[pc[P_YM]], bx, f.bhx[pc[P_YP]]);
      dby_y = ddy(f.bhy[pc[P_YM]], by, f.bhy[pc[P_YP]]);
      const double vxc = a.eval.vx[c], vyc = a.eval.vy[c];
      sxx = qdiv(fint(M.x, ip, i, vxc, a.eval.vx[pc[P_XP]]) - 2.0 * vxc + fint(M.x, i, im, a.eval.vx[pc[P_XM]], vxc),
                 M.x.hsq[i], M.x.rhsq[i]);
      syy = qdiv(fint(M.y, jp, j, vyc, a.eval.vy[pc[P_YP]]) - 2.0 * vyc + fint(M.y, j, jm, a.eval.vy[pc[P_YM]], vyc),
                 M.y.hsq[j], M.y.rhsq[j]);
      x_yvy = ddx(d.yvy[pc[P_XM]], dvc[D_YVY], d.yvy[pc[P_XP]]);
      y_xvx = ddy(d.xvx[pc[P_YM]], dvc[D_XVX], d.xvx[pc[P_YP]]);
      sfd = bx * (bx * dvc[D_XVX]) + by * (by * dvc[D_YVY]);
      gbd[0] = 2.0 * bx * dbx_x * dvc[D_XVX] + bx * bx * sxx + 2.0 * by * dby_x * dvc[D_YVY] + by * by * x_yvy;
      gbd[1] = 2.0 * bx * dbx_y * dvc[D_XVX] + bx * bx * y_xvx + 2.0 * by * dby_y * dvc[D_YVY] + by * by * syy;
      for (int p = 0; p < NPOS; ++p) {
        for (int k = 0; k < 6; ++k) dvn[p][k] = dget(d, k, pc[p]);
        sfo[p] = sfo_of(f.bhx[pc[p]], f.bhy[pc[p]], f.bhz[pc[p]], dvn[p]);
      }
      for (int jj = 0; jj < 3; ++jj) {
        double r = 0.0;
        for (int ii = 0; ii < 2; ++ii) {
          const double delta = (ii == jj) ? 1.0 / 3.0 : 0.0;
          auto Gat = [&](size_t q) {
            const double t = delta - bh[ii][q] * bh[jj][q];
            return gc ? d.t25[q] * f.coef[q] * t : d.t25[q] * t;
          };
          const double tc = delta - bc[ii] * bc[jj];
          const double Gc = gc ? t25 * coef * tc : t25 * tc;
          const int pm = ii == 0 ? P_XM : P_YM, pp = ii == 0 ? P_XP : P_YP;
          const double Gm = Gat(pc[pm]), Gp = Gat(pc[pp]);
          const double Fm = Gm * sfo[pm], Fc = Gc * sfo[P_C], Fp = Gp * sfo[pp];
          r = r + (ii == 0 ? ddx(Fm, Fc, Fp) : ddy(Fm, Fc, Fp));
          r = r + (gc ? gbd[ii] * t25 * coef * tc : gbd[ii] * t25 * tc);
          r = r + sfd * (ii == 0 ? ddx(Gm, Gc, Gp) : ddy(Gm, Gc, Gp));
        }
        fr[jj] = r * (gc ? -3.0 * mask : -3.0 * coef * mask);
      }
    }
    // ---- the update (:163-245) reversed ----
    double lte = L.lo_te[c], lheat = 0.0;
    if (a.heating_on && a.apply) {
      const double te_pre = a.mode == PV_HALF ? a.base.te[c] + heat * 0.5 * a.h : a.base.te[c] + heat * a.h;
      if (te_pre < K.temin) lte = 0.0;
      lheat = lte * (mult * a.h);
      dot += lte * heat * mult;
    }
    double lF[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 3; ++q) {
      const double lm = L.lo_m[q][c];
      if (a.force_on && a.apply) {
        lF[q] = lm * (mult * a.h);
        dot += lm * fr[q] * mult;
      }
      if (L.base_mode == 1) L.lb_m[q][c] = lm;
      else if (L.base_mode == 2) L.lb_m[q][c] += lm;
    }
    if (L.base_mode == 0 || L.base_mode == 1) L.lb_te[c] = lte;
    else L.lb_te[c] += lte;
    // ---- heating and force reversed onto the stencil ----
    if (in && (lheat != 0.0 || lF[0] != 0.0 || lF[1] != 0.0 || lF[2] != 0.0)) {
      double g_bh[NPOS][3] = {}, g_d[NPOS][6] = {}, g_t25[NPOS] = {}, g_v[NPOS][2] = {};
      if (a.heating_on && !a.coeff_zero && lheat != 0.0) {
        const double g = lheat * 3.0 * coef;
        g_t25[P_C] += g * (S * S);
        const double lS = g * t25 * 2.0 * S;
        g_bh[P_C][0] += lS * ((bx * dvc[D_XVX] + by * dvc[D_YVX]) + bx * dvc[D_XVX] + by * dvc[D_XVY] + bz * dvc[D_XVZ]);
        g_bh[P_C][1] += lS * (bx * dvc[D_YVX] + (bx * dvc[D_XVY] + by * dvc[D_YVY]) + by * dvc[D_YVY] + bz * dvc[D_YVZ]);
        g_bh[P_C][2] += lS * (bx * dvc[D_XVZ] + by * dvc[D_YVZ]);
        g_d[P_C][D_XVX] += lS * (bx * bx - 1.0 / 3.0);
        g_d[P_C][D_YVX] += lS * (bx * by);
        g_d[P_C][D_XVY] += lS * (by * bx);
        g_d[P_C][D_YVY] += lS * (by * by - 1.0 / 3.0);
        g_d[P_C][D_XVZ] += lS * (bz * bx);
        g_d[P_C][D_YVZ] += lS * (bz * by);
      }
      if (a.force_on) {
        const double Kf = gc ? -3.0 : -3.0 * coef;
        double l_sfd = 0.0, l_gbd[2] = {0.0, 0.0}, l_sfo[NPOS] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int jj = 0; jj < 3; ++jj) {
          const double lr = lF[jj] * Kf;
          if (lr == 0.0) continue;
          for (int ii = 0; ii < 2; ++ii) {
            const double delta = (ii == jj) ? 1.0 / 3.0 : 0.0;
            const W3& w = ii == 0 ? wx : wy;
            const int pm = ii == 0 ? P_XM : P_YM, pp = ii == 0 ? P_XP : P_YP;
            const int pos[3] = {pm, P_C, pp};
            double G[3], cf[3], tq[3];
            for (int k = 0; k < 3; ++k) {
              const size_t q = pc[pos[k]];
              tq[k] = delta - bh[ii][q] * bh[jj][q];
              cf[k] = gc ? f.coef[q] : 1.0;
              G[k] = d.t25[q] * cf[k] * tq[k];
            }
            const double wv[3] = {w.m, w.c, w.p};
            const double dG = w.m * G[0] + w.c * G[1] + w.p * G[2];
            l_gbd[ii] += lr * G[1];
            l_sfd += lr * dG;
            double lG[3];
            for (int k = 0; k < 3; ++k) {
              const double lFk = lr * wv[k];                      // d/dF_k of the derivative of F
              lG[k] = lFk * sfo[pos[k]] + lr * sfd * wv[k];        // F = G sfo, and sfd dG
              l_sfo[pos[k]] += lFk * G[k];
            }
            lG[1] += lr * gbd[ii];
            for (int k = 0; k < 3; ++k) {
              const size_t q = pc[pos[k]];
              g_t25[pos[k]] += lG[k] * cf[k] * tq[k];
              g_bh[pos[k]][ii] -= lG[k] * d.t25[q] * cf[k] * bh[jj][q];
              g_bh[pos[k]][jj] -= lG[k] * d.t25[q] * cf[k] * bh[ii][q];
            }
          }
        }
        // sfo at the five positions
        for (int p = 0; p < NPOS; ++p) {
          const double l = l_sfo[p];
          if (l == 0.0) continue;
          const double qx = f.bhx[pc[p]], qy = f.bhy[pc[p]], qz = f.bhz[pc[p]];
          const double* dv = dvn[p];
          g_bh[p][0] += l * (qy * dv[D_YVX] + qy * dv[D_XVY] + qz * dv[D_XVZ]);
          g_bh[p][1] += l * (qx * dv[D_YVX] + qx * dv[D_XVY] + qz * dv[D_YVZ]);
          g_bh[p][2] += l * (qx * dv[D_XVZ] + qy * dv[D_YVZ]);
          g_d[p][D_XVX] -= l / 3.0;
          g_d[p][D_YVY] -= l / 3.0;
          g_d[p][D_YVX] += l * (qx * qy);
          g_d[p][D_XVY] += l * (qy * qx);
          g_d[p][D_XVZ] += l * (qz * qx);
          g_d[p][D_YVZ] += l * (qz * qy);
        }
        // sfd = bx (bx xvx) + by (by yvy)
        g_bh[P_C][0] += l_sfd * 2.0 * bx * dvc[D_XVX];
        g_bh[P_C][1] += l_sfd * 2.0 * by * dvc[D_YVY];
        g_d[P_C][D_XVX] += l_sfd * bx * bx;
        g_d[P_C][D_YVY] += l_sfd * by * by;
        // gbd[0] and gbd[1]
        const double g0 = l_gbd[0], g1 = l_gbd[1];
        g_bh[P_C][0] += g0 * (2.0 * dbx_x * dvc[D_XVX] + 2.0 * bx * sxx) + g1 * (2.0 * dbx_y * dvc[D_XVX] + 2.0 * bx * y_xvx);
        g_bh[P_C][1] += g0 * (2.0 * dby_x * dvc[D_YVY] + 2.0 * by * x_yvy) + g1 * (2.0 * dby_y * dvc[D_YVY] + 2.0 * by * syy);
        g_d[P_C][D_XVX] += g0 * 2.0 * bx * dbx_x + g1 * 2.0 * bx * dbx_y;
        g_d[P_C][D_YVY] += g0 * 2.0 * by * dby_x + g1 * 2.0 * by * dby_y;
        const double l_dbx_x = g0 * 2.0 * bx * dvc[D_XVX], l_dby_x = g0 * 2.0 * by * dvc[D_YVY];
        const double l_dbx_y = g1 * 2.0 * bx * dvc[D_XVX], l_dby_y = g1 * 2.0 * by * dvc[D_YVY];
        const double l_sxx = g0 * bx * bx, l_x_yvy = g0 * by * by, l_y_xvx = g1 * bx * bx, l_syy = g1 * by * by;
        // the derivatives of b_hat, v and of the velocity derivatives they take
        g_bh[P_XM][0] += l_dbx_x * wx.m; g_bh[P_C][0] += l_dbx_x * wx.c; g_bh[P_XP][0] += l_dbx_x * wx.p;
        g_bh[P_XM][1] += l_dby_x * wx.m; g_bh[P_C][1] += l_dby_x * wx.c; g_bh[P_XP][1] += l_dby_x * wx.p;
        g_bh[P_YM][0] += l_dbx_y * wy.m; g_bh[P_C][0] += l_dbx_y * wy.c; g_bh[P_YP][0] += l_dbx_y * wy.p;
        g_bh[P_YM][1] += l_dby_y * wy.m; g_bh[P_C][1] += l_dby_y * wy.c; g_bh[P_YP][1] += l_dby_y * wy.p;
        g_v[P_XM][0] += l_sxx * sx2.m; g_v[P_C][0] += l_sxx * sx2.c; g_v[P_XP][0] += l_sxx * sx2.p;
        g_v[P_YM][1] += l_syy * sy2.m; g_v[P_C][1] += l_syy * sy2.c; g_v[P_YP][1] += l_syy * sy2.p;
        g_d[P_XM][D_YVY] += l_x_yvy * wx.m; g_d[P_C][D_YVY] += l_x_yvy * wx.c; g_d[P_XP][D_YVY] += l_x_yvy * wx.p;
        g_d[P_YM][D_XVX] += l_y_xvx * wy.m; g_d[P_C][D_XVX] += l_y_xvx * wy.c; g_d[P_YP][D_XVX] += l_y_xvx * wy.p;
      }
      for (int p = 0; p < NPOS; ++p) {
        const size_t q = pc[p];
        for (int k = 0; k < 3; ++k)
          if (g_bh[p][k] != 0.0) atomicAdd(&L.l_bh[k][q], g_bh[p][k]);
        for (int k = 0; k < 6; ++k)
          if (g_d[p][k] != 0.0) atomicAdd(&L.l_d[k][q], g_d[p][k]);
        if (g_t25[p] != 0.0) atomicAdd(&L.l_t25[q], g_t25[p]);
        for (int k = 0; k < 2; ++k)
          if (g_v[p][k] != 0.0) atomicAdd(&L.l_v[k][q], g_v[p][k]);
      }
    }
    (void)N;
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    L.partial[b] = L.accumulate ? L.partial[b] + dot : dot;
  }
}

// the velocity derivatives (interior cells) and T^2.5 (every cell) of the evaluation set reversed
__global__ void __launch_bounds__(256) k_adj_pv_derivs(PvSet E, Mesh M, PvLam L) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double lt = L.l_t25[c];
  if (lt != 0.0) {
    L.l_t25[c] = 0.0;
    atomicAdd(&L.l_T[c], lt * 2.5 * pow_1p5(E.T[c]));
  }
  double ld[6];
  bool any = false;
  for (int k = 0; k < 6; ++k) {
    ld[k] = L.l_d[k][c];
    any = any || ld[k] != 0.0;
    L.l_d[k][c] = 0.0;
  }
  if (!any || !inside(M, i, j)) return;   // the derivatives are zero off the interior
  const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
  const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
  const W3 wx = d_weights(M.x, i, im, ip), wy = d_weights(M.y, j, jm, jp);
  const size_t cxm = (size_t)im * M.ny + j, cxp = (size_t)ip * M.ny + j;
  const size_t cym = (size_t)i * M.ny + jm, cyp = (size_t)i * M.ny + jp;
  for (int q = 0; q < 3; ++q) {
    const double lx = ld[q], ly = ld[3 + q];
    if (lx != 0.0) {
      atomicAdd(&L.l_v[q][cxm], lx * wx.m);
      atomicAdd(&L.l_v[q][c], lx * wx.c);
      atomicAdd(&L.l_v[q][cxp], lx * wx.p);
    }
    if (ly != 0.0) {
      atomicAdd(&L.l_v[q][cym], ly * wy.m);
      atomicAdd(&L.l_v[q][c], ly * wy.c);
      atomicAdd(&L.l_v[q][cyp], ly * wy.p);
    }
  }
}

// v = m / rho and T = max((gamma-1) te / (2 K_B n), T_min) of the evaluation set reversed onto its
// thermal energy and momenta (assigned or added) and the frozen density (lam_rho, with n = rho/m_i)
__global__ void __launch_bounds__(256) k_adj_pv_derived(PvSet E, PvFrozen f, MhdConst K, size_t N, PvLam L,
                                                        double* out_te, double* out_mx, double* out_my,
                                                        double* out_mz, int assign, double* lam_rho) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double rho = f.rho[c], n = f.n[c];
  const double m[3] = {E.mx[c], E.my[c], E.mz[c]};
  double* out_m[3] = {out_mx, out_my, out_mz};
  double lr = 0.0;
  for (int q = 0; q < 3; ++q) {
    const double lv = L.l_v[q][c];
    L.l_v[q][c] = 0.0;
    const double lm = lv / rho;
    lr -= lv * m[q] / (rho * rho);
    out_m[q][c] = assign ? lm : out_m[q][c] + lm;
  }
  const double lT = L.l_T[c];
  L.l_T[c] = 0.0;
  double lte = 0.0;
  const double den = 2.0 * K_B * n;
  const double x = K.gm1 * E.te[c] / den;
  if (lT != 0.0 && !(x < K.tmin)) {
    lte = lT * (K.gm1 / den);
    lr += lT * (-x / n) * K.rm;
  }
  out_te[c] = assign ? lte : out_te[c] + lte;
  lam_rho[c] += lr;
}

// b_hat = b / |b| (b = be + bi; zero where |b| = 0) reversed onto bi; l_bh cleared
__global__ void __launch_bounds__(256) k_adj_pv_bhat(MhdSet S, MhdStatic st, size_t N, PvLam L, double* lam) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double g[3] = {L.l_bh[0][c], L.l_bh[1][c], L.l_bh[2][c]};
  L.l_bh[0][c] = 0.0; L.l_bh[1][c] = 0.0; L.l_bh[2][c] = 0.0;
  const double b[3] = {st.bex[c] + S.bx[c], st.bey[c] + S.by[c], st.bez[c] + S.bz[c]};
  const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  if (bm == 0.0) return;
  const double gb = (g[0] * b[0] + g[1] * b[1] + g[2] * b[2]) / (bm * bm * bm);
  for (int k = 0; k < 3; ++k) lam[(5 + k) * N + c] += g[k] / bm - gb * b[k];
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

void launch_adj_pv_apply(const PvArgs& a, const PvLam& L, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_pv_apply, g2(a.mesh, dim3(64, 4)), dim3(64, 4), 0, s, a, L);
}
void launch_adj_pv_derivs(PvSet E, Mesh mesh, const PvLam& L, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_pv_derivs, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, E, mesh, L);
}
void launch_adj_pv_derived(PvSet E, PvFrozen f, MhdConst k, size_t N, const PvLam& L, double* out_te, double* out_mx,
                           double* out_my, double* out_mz, int assign, double* lam_rho, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_pv_derived, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, E, f, k, N, L, out_te,
                     out_mx, out_my, out_mz, assign, lam_rho);
}
void launch_adj_pv_bhat(MhdSet S, MhdStatic st, size_t N, const PvLam& L, double* lam, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_pv_bhat, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, S, st, N, L, lam);
}

}  // namespace spruce
