// Physical (Braginskii) viscosity kernels for gfx950 (source/modules/solar/physicalviscosity.cpp),
// rounded where the reference's whole-grid operations round.  T^2.5 is the correctly rounded power.
#include <cfloat>

#include "arith.hpp"
#include "crpow.hpp"
#include "pvisc.hpp"

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
// boundaryInterpolate on the face `up` (between up-1 and up): (q_lo h_up + q_up h_lo)/(h_up + h_lo)
__device__ __forceinline__ double fint(const Axis& A, int up, int lo, double qlo, double qup) {
  return qdiv(qlo * A.h[up] + qup * A.h[lo], A.fden[up], A.rfden[up]);
}

__global__ void __launch_bounds__(256) k_pv_init(MhdSet L, MhdStatic st, MhdConst K, Mesh M, double* bhx,
                                                 double* bhy, double* bhz, PvSet o) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double bx = st.bex[c] + L.bx[c], by = st.bey[c] + L.by[c], bz = st.bez[c] + L.bz[c];
  const double bm = sqrt(bx * bx + by * by + bz * bz);
  bhx[c] = bm == 0.0 ? 0.0 : bx / bm;                       // idealmhd.cpp:258-276
  bhy[c] = bm == 0.0 ? 0.0 : by / bm;
  bhz[c] = bm == 0.0 ? 0.0 : bz / bm;
  const double rho = L.rho[c], te = L.te[c];
  o.te[c] = te;
  o.T[c] = smax((te * K.gm1) / (L.n[c] * (2.0 * K_B)), K.tmin);
  o.mx[c] = L.mx[c]; o.my[c] = L.my[c]; o.mz[c] = L.mz[c];
  o.vx[c] = L.mx[c] / rho; o.vy[c] = L.my[c] / rho; o.vz[c] = L.mz[c] / rho;
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

// computeViscousSubcycles (physicalviscosity.cpp:67-80): timescale per interior cell, block minima
__global__ void __launch_bounds__(256) k_pv_timescale(PvSet cur, PvFrozen f, double floor, Mesh M,
                                                      double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double v = DBL_MAX;
  if (i < M.nx && j < M.ny && inside(M, i, j)) {
    const size_t c = (size_t)i * M.ny + j;
    const double vx = cur.vx[c], vy = cur.vy[c];
    const double vmag = sqrt(vx * vx + vy * vy);
    const double a = f.bhx[c] * vx / vmag + f.bhy[c] * vy / vmag;
    double df = 3.0 * (a * a) + 1.0;
    if (vmag == 0.0) df = 4.0;
    const double ts = (M.x.d[i] * M.y.d[j]) * f.rho[c] / (df * 3.0 * smax(f.coef[c], floor) * pow_2p5(cur.T[c]));
    v = (ts < v) ? ts : v;
  }
  v = block_min(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = v;
}

// derivative1D of the three velocities along both axes (0 outside the interior), and T^2.5
__global__ void __launch_bounds__(256) k_pv_derivs(PvSet e, Mesh M, PvDerivs d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  d.t25[c] = pow_2p5(e.T[c]);
  double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (inside(M, i, j)) {
    const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
    const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
    const size_t cxm = (size_t)im * M.ny + j, cxp = (size_t)ip * M.ny + j;
    const size_t cym = (size_t)i * M.ny + jm, cyp = (size_t)i * M.ny + jp;
    const double* v[3] = {e.vx, e.vy, e.vz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double vc = v[q][c];
      r[q] = qdiv(fint(M.x, ip, i, vc, v[q][cxp]) - fint(M.x, i, im, v[q][cxm], vc), M.x.d[i], M.x.rd[i]);
      r[3 + q] = qdiv(fint(M.y, jp, j, vc, v[q][cyp]) - fint(M.y, j, jm, v[q][cym], vc), M.y.d[j], M.y.rd[j]);
    }
  }
  d.xvx[c] = r[0]; d.xvy[c] = r[1]; d.xvz[c] = r[2];
  d.yvx[c] = r[3]; d.yvy[c] = r[4]; d.yvz[c] = r[5];
}

__device__ __forceinline__ double sfo_at(const PvDerivs& d, const PvFrozen& f, size_t c) {
  const double bx = f.bhx[c], by = f.bhy[c], bz = f.bhz[c];
  return bx * (by * d.yvx[c]) + by * (bx * d.xvy[c]) + bz * (bx * d.xvz[c] + by * d.yvz[c]) -
         (d.xvx[c] + d.yvy[c]) / 3.0;
}

// computeHeating and computeViscousForce at one cell (:47-64, :82-140), then the subcycle update
// (:163-245).
__global__ void __launch_bounds__(256) k_pv_apply(PvArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  const Mesh& M = a.mesh;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const bool in = inside(M, i, j);
  const double mask = in ? 1.0 : 0.0;
  const PvDerivs& d = a.d;
  const PvFrozen& f = a.f;
  const MhdConst& K = a.k;
  const double bx = f.bhx[c], by = f.bhy[c], bz = f.bhz[c];
  const double xvx = d.xvx[c], xvy = d.xvy[c], xvz = d.xvz[c], yvx = d.yvx[c], yvy = d.yvy[c], yvz = d.yvz[c];
  const double t25 = d.t25[c], coef = f.coef[c];
  double heat = 0.0;
  if (a.heating_on && !a.coeff_zero) {
    const double S = bx * (bx * xvx + by * yvx) + by * (bx * xvy + by * yvy) + bz * (bx * xvz + by * yvz) -
                     (xvx + yvy) / 3.0;
    heat = mask * (3.0 * coef * t25 * (S * S));
  }
  double fr[3] = {0.0, 0.0, 0.0};
  if (a.force_on) {
    int im = i, ip = i, jm = j, jp = j;
    double dbx_x = 0.0, dby_x = 0.0, dbx_y = 0.0, dby_y = 0.0, sxx = 0.0, syy = 0.0, x_yvy = 0.0, y_xvx = 0.0;
    if (in) {
      im = wrp(i - 1, M.nx, M.x.per); ip = wrp(i + 1, M.nx, M.x.per);
      jm = wrp(j - 1, M.ny, M.y.per); jp = wrp(j + 1, M.ny, M.y.per);
    }
    const size_t cxm = (size_t)im * M.ny + j, cxp = (size_t)ip * M.ny + j;
    const size_t cym = (size_t)i * M.ny + jm, cyp = (size_t)i * M.ny + jp;
    const double dx = M.x.d[i], rdx = M.x.rd[i], dy = M.y.d[j], rdy = M.y.rd[j];
    auto ddx = [&](double qm, double qc, double qp) {
      return qdiv(fint(M.x, ip, i, qc, qp) - fint(M.x, i, im, qm, qc), dx, rdx);
    };
    auto ddy = [&](double qm, double qc, double qp) {
      return qdiv(fint(M.y, jp, j, qc, qp) - fint(M.y, j, jm, qm, qc), dy, rdy);
    };
    if (in) {
      dbx_x = ddx(f.bhx[cxm], bx, f.bhx[cxp]);
      dby_x = ddx(f.bhy[cxm], by, f.bhy[cxp]);
      dbx_y = ddy(f.bhx[cym], bx, f.bhx[cyp]);
      dby_y = ddy(f.bhy[cym], by, f.bhy[cyp]);
      const double vxc = a.eval.vx[c], vyc = a.eval.vy[c];
      // secondDerivative1D (derivs.cpp:450)
      sxx = qdiv(fint(M.x, ip, i, vxc, a.eval.vx[cxp]) - 2.0 * vxc + fint(M.x, i, im, a.eval.vx[cxm], vxc),
                 M.x.hsq[i], M.x.rhsq[i]);
      syy = qdiv(fint(M.y, jp, j, vyc, a.eval.vy[cyp]) - 2.0 * vyc + fint(M.y, j, jm, a.eval.vy[cym], vyc),
                 M.y.hsq[j], M.y.rhsq[j]);
      // nested derivatives of derivatives that are zero outside the interior (:98-102)
      x_yvy = ddx(d.yvy[cxm], yvy, d.yvy[cxp]);
      y_xvx = ddy(d.xvx[cym], xvx, d.xvx[cyp]);
    }
    const double sfd = bx * (bx * xvx) + by * (by * yvy);
    const double gbd[2] = {2.0 * bx * dbx_x * xvx + bx * bx * sxx + 2.0 * by * dby_x * yvy + by * by * x_yvy,
                           2.0 * bx * dbx_y * xvx + bx * bx * y_xvx + 2.0 * by * dby_y * yvy + by * by * syy};
    const double* bh[3] = {f.bhx, f.bhy, f.bhz};
    const double bc[3] = {bx, by, bz};
    const size_t nb[2][2] = {{cxm, cxp}, {cym, cyp}};
    double sfo_n[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (in) {
      sfo_n[0][0] = sfo_at(d, f, cxm); sfo_n[0][1] = sfo_at(d, f, cxp);
      sfo_n[1][0] = sfo_at(d, f, cym); sfo_n[1][1] = sfo_at(d, f, cyp);
    }
    const double sfoc = sfo_at(d, f, c);
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
      double r = 0.0;
#pragma unroll
      for (int ii = 0; ii < 2; ++ii) {
        const double delta = (ii == jj) ? 1.0 / 3.0 : 0.0;
        // G = T^2.5 [* coef] * (delta - b_i b_j), F = G * sfo, at c and its two neighbours along ii
        auto Gat = [&](size_t q) {
          const double t = delta - bh[ii][q] * bh[jj][q];
          return a.grad_corr ? d.t25[q] * f.coef[q] * t : d.t25[q] * t;
        };
        const double tc = delta - bc[ii] * bc[jj];
        const double Gc = a.grad_corr ? t25 * coef * tc : t25 * tc;
        double dF = 0.0, dG = 0.0;
        if (in) {
          const double Gm = Gat(nb[ii][0]), Gp = Gat(nb[ii][1]);
          const double Fm = Gm * sfo_n[ii][0], Fc = Gc * sfoc, Fp = Gp * sfo_n[ii][1];
          dF = ii == 0 ? ddx(Fm, Fc, Fp) : ddy(Fm, Fc, Fp);
          dG = ii == 0 ? ddx(Gm, Gc, Gp) : ddy(Gm, Gc, Gp);
        }
        r = r + dF;
        r = r + (a.grad_corr ? gbd[ii] * t25 * coef * tc : gbd[ii] * t25 * tc);
        r = r + sfd * dG;
      }
      fr[jj] = r * (a.grad_corr ? -3.0 * mask : -3.0 * coef * mask);
    }
  }
  if (a.mode == PV_FULL) {
    if (a.avg_heat && a.heating_on) a.avg_heat[c] = a.avg_heat[c] + heat / a.nsd;
    if (a.avg_f[0] && a.force_on) {
#pragma unroll
      for (int q = 0; q < 3; ++q) a.avg_f[q][c] = a.avg_f[q][c] + fr[q] / a.nsd;
    }
    if (a.ms_ion && a.heating_on && a.apply) {
      if (a.msf < 1.0) a.ms_ion[c] = a.ms_ion[c] + ((1.0 - a.msf) * heat) * a.h;
      if (a.msf > 0.0) a.ms_el[c] = a.ms_el[c] + (a.msf * heat) * a.h;
    }
  }
  // subcycle update (euler :163-188; rk2 half step :194-210 and full step :220-245)
  double te = a.base.te[c], T = a.base.T[c];
  double m[3] = {a.base.mx[c], a.base.my[c], a.base.mz[c]};
  double v[3] = {a.base.vx[c], a.base.vy[c], a.base.vz[c]};
  if (a.heating_on && a.apply) {
    te = a.mode == PV_HALF ? te + heat * 0.5 * a.h : te + heat * a.h;
    te = smax(te, K.temin);
    T = smax(K.gm1 * te / (2 * K_B * f.n[c]), K.tmin);
  }
  if (a.force_on && a.apply) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      m[q] = a.mode == PV_HALF ? m[q] + fr[q] * 0.5 * a.h : m[q] + fr[q] * a.h;
      v[q] = m[q] / f.rho[c];
    }
  }
  a.out.te[c] = te; a.out.T[c] = T;
  a.out.mx[c] = m[0]; a.out.my[c] = m[1]; a.out.mz[c] = m[2];
  a.out.vx[c] = v[0]; a.out.vy[c] = v[1]; a.out.vz[c] = v[2];
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

void launch_pv_init(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double* bhx, double* bhy, double* bhz, PvSet out,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_pv_init, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, st, k, mesh, bhx, bhy, bhz, out);
}
int pv_blocks(const Mesh& m) {
  const dim3 g = g2(m, dim3(64, 4));
  return (int)(g.x * g.y);
}
void launch_pv_timescale(PvSet cur, PvFrozen f, double coeff_floor, Mesh mesh, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_pv_timescale, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, cur, f, coeff_floor, mesh, partial);
}
void launch_pv_derivs(PvSet e, Mesh mesh, PvDerivs d, hipStream_t s) {
  hipLaunchKernelGGL(k_pv_derivs, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, e, mesh, d);
}
void launch_pv_apply(const PvArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pv_apply, g2(a.mesh, dim3(64, 4)), dim3(64, 4), 0, s, a);
}

}  // namespace spruce
