// Artificial-viscosity kernels for gfx950 (source/modules/viscosity.cpp), operation for operation:
// nu = ((s*1) / (1/d_x^2 + 1/d_y^2)) / 2 / dt, dqdt = nu * lap(q) * scale [+ gradient correction].
#include "arith.hpp"
#include "visc.hpp"

namespace spruce {
namespace {

__device__ __forceinline__ int wrapv(int k, int n, int per) {
  if (per) {
    if (k < 0) k += n;
    else if (k >= n) k -= n;
  }
  return k;
}
__device__ __forceinline__ bool inside(const Mesh& M, int i, int j) {
  return i >= M.x.lo && i <= M.x.hi && j >= M.y.lo && j <= M.y.hi;
}

// A variable of the set at one cell, computed as recomputeDerivedVarsFromEvolvedVars does
// (idealmhd.cpp:241-277).
__device__ double cell_var(const MhdSet& S, const MhdStatic& st, const MhdConst& K, int var, size_t c) {
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

// nu * scale at one cell; nu = ((s * 1) / (1/d_x^2 + 1/d_y^2)) / 2 / dt (viscosity.cpp:212)
__device__ __forceinline__ double visc_coeff(const ViscTermArgs& a, int i, int j, size_t c) {
  const Mesh& M = a.mesh;
  const double dx = M.x.d[i], dy = M.y.d[j];
  const double s = a.strength ? a.strength[c] : a.strength_c;
  const double dt = a.dtgrid ? a.dtgrid[c] : *a.dtmin;
  return s / (1.0 / (dx * dx) + 1.0 / (dy * dy)) / 2. / dt;
}
__device__ __forceinline__ double scale_of(const ViscTermArgs& a, size_t c) {
  if (a.scale == 0) return 1.0;
  return a.S.n[c] * a.scale_c;
}

__global__ void __launch_bounds__(256) k_visc_term(ViscTermArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const Mesh& M = a.mesh;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const bool in = inside(M, i, j);
  const double nu = visc_coeff(a, i, j, c);
  const double sc = scale_of(a, c);
  double lap = 0.0, gx = 0.0, gy = 0.0, fx = 0.0, fy = 0.0;
  if (in) {
    const int im = wrapv(i - 1, M.x.n, M.x.per), ip = wrapv(i + 1, M.x.n, M.x.per);
    const int jm = wrapv(j - 1, M.y.n, M.y.per), jp = wrapv(j + 1, M.y.n, M.y.per);
    const size_t cxm = (size_t)im * M.ny + j, cxp = (size_t)ip * M.ny + j;
    const size_t cym = (size_t)i * M.ny + jm, cyp = (size_t)i * M.ny + jp;
    const double q = cell_var(a.S, a.st, a.k, a.var, c);
    const double qxm = cell_var(a.S, a.st, a.k, a.var, cxm), qxp = cell_var(a.S, a.st, a.k, a.var, cxp);
    const double qym = cell_var(a.S, a.st, a.k, a.var, cym), qyp = cell_var(a.S, a.st, a.k, a.var, cyp);
    // boundaryInterpolate across the faces of this cell (derivs.cpp:477-487)
    const Axis& X = M.x;
    const Axis& Y = M.y;
    auto ixu = [&](double qc, double qp) { return qdiv(qc * X.h[ip] + qp * X.h[i], X.fden[ip], X.rfden[ip]); };
    auto ixl = [&](double qm, double qc) { return qdiv(qm * X.h[i] + qc * X.h[im], X.fden[i], X.rfden[i]); };
    auto iyu = [&](double qc, double qp) { return qdiv(qc * Y.h[jp] + qp * Y.h[j], Y.fden[jp], Y.rfden[jp]); };
    auto iyl = [&](double qm, double qc) { return qdiv(qm * Y.h[j] + qc * Y.h[jm], Y.fden[j], Y.rfden[j]); };
    const double qxu = ixu(q, qxp), qxl = ixl(qxm, q), qyu = iyu(q, qyp), qyl = iyl(qym, q);
    // laplacian = secondDerivative1D(q,0) + secondDerivative1D(q,1) (derivs.cpp:417-462)
    const double sx = qdiv(qxu - 2.0 * q + qxl, X.hsq[i], X.rhsq[i]);
    const double sy = qdiv(qyu - 2.0 * q + qyl, Y.hsq[j], Y.rhsq[j]);
    lap = sx + sy;
    if (a.grad) {
      // derivative1D of nu*scale and of q (derivs.cpp:223-264)
      const double f = nu * sc;
      const double fxm = visc_coeff(a, im, j, cxm) * scale_of(a, cxm);
      const double fxp = visc_coeff(a, ip, j, cxp) * scale_of(a, cxp);
      const double fym = visc_coeff(a, i, jm, cym) * scale_of(a, cym);
      const double fyp = visc_coeff(a, i, jp, cyp) * scale_of(a, cyp);
      fx = qdiv(ixu(f, fxp) - ixl(fxm, f), X.d[i], X.rd[i]);
      fy = qdiv(iyu(f, fyp) - iyl(fym, f), Y.d[j], Y.rd[j]);
      gx = qdiv(qxu - qxl, X.d[i], X.rd[i]);
      gy = qdiv(qyu - qyl, Y.d[j], Y.rd[j]);
    }
  }
  double r = nu * lap * sc;
  if (a.grad) r = r + fx * gx + fy * gy;
  a.out[c] = a.masked ? r * (in ? 1.0 : 0.0) : r;
  if (a.lap_out) a.lap_out[c] = lap;   // zero outside the interior, as laplacian() leaves it
  if (a.str_out) a.str_out[c] = a.strength ? a.strength[c] : a.strength_c;
  if (a.dt_out) a.dt_out[c] = a.dtgrid ? a.dtgrid[c] : *a.dtmin;
}

__global__ void __launch_bounds__(256) k_visc_update(double* target, const double* init, const double* d, double h,
                                                     Mesh M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double mask = inside(M, i, j) ? 1.0 : 0.0;
  target[c] = init[c] + mask * h * d[c];
}

__global__ void __launch_bounds__(256) k_visc_rk4(const double* d1, const double* d2, const double* d3,
                                                  const double* d4, double* out, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = (d1[c] + 2.0 * d2[c] + 2.0 * d3[c] + d4[c]) / 6.0;
}

__global__ void __launch_bounds__(256) k_add(double* out, const double* add, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = out[c] + add[c];
}

__global__ void __launch_bounds__(256) k_add_step_times(double* y, const double* x, const double* step, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) y[c] = y[c] + (*step) * x[c];
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

void launch_visc_term(const ViscTermArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_visc_term, g2(a.mesh, dim3(64, 4)), dim3(64, 4), 0, s, a);
}

void launch_visc_update(double* target, const double* init, const double* d, double h, Mesh mesh, hipStream_t s) {
  hipLaunchKernelGGL(k_visc_update, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, target, init, d, h, mesh);
}

void launch_visc_rk4(const double* d1, const double* d2, const double* d3, const double* d4, double* out, size_t n,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_visc_rk4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d1, d2, d3, d4, out, n);
}

void launch_add(double* out, const double* add, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, add, n);
}

void launch_add_step_times(double* y, const double* x, const double* step, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_add_step_times, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, x, step, n);
}

}  // namespace spruce
