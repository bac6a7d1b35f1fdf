// Source-term module kernels for gfx950 (reference: source/modules/solar/localizedheating.cpp,
// massinjection.cpp, momentuminjection.cpp, ambientheatingsink.cpp, fieldheating.cpp,
// divcleaning.cpp).  Each whole-grid expression of the reference is evaluated per cell with the
// same operation order; derivatives are the reference's derivative1D / secondDerivative1D
// (derivs.cpp:376-455): zero off the interior, face values by boundaryInterpolate.
#include "arith.hpp"
#include "sources.hpp"

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
// boundaryInterpolate across the face whose upper cell is c on axis ax (derivs.cpp:477-487)
__device__ __forceinline__ double fint(const Axis& ax, int c, int cm1, double qlo, double qup) {
  return qdiv(qlo * ax.h[c] + qup * ax.h[cm1], ax.fden[c], ax.rfden[c]);
}

// Neighbour offsets of an interior cell (periodic axes wrap).
struct Nbr {
  int im, ip, jm, jp;
  size_t c, xm, xp, ym, yp;
};
__device__ __forceinline__ Nbr nbr(const Mesh& M, int i, int j) {
  Nbr n;
  n.im = wrp(i - 1, M.nx, M.x.per);
  n.ip = wrp(i + 1, M.nx, M.x.per);
  n.jm = wrp(j - 1, M.ny, M.y.per);
  n.jp = wrp(j + 1, M.ny, M.y.per);
  n.c = (size_t)i * M.ny + j;
  n.xm = (size_t)n.im * M.ny + j;
  n.xp = (size_t)n.ip * M.ny + j;
  n.ym = (size_t)i * M.ny + n.jm;
  n.yp = (size_t)i * M.ny + n.jp;
  return n;
}
// derivative1D(q, 0) and derivative1D(q, 1) at an interior cell
__device__ __forceinline__ double ddx(const Mesh& M, int i, const Nbr& n, const double* q) {
  const double qc = q[n.c];
  return qdiv(fint(M.x, n.ip, i, qc, q[n.xp]) - fint(M.x, i, n.im, q[n.xm], qc), M.x.d[i], M.x.rd[i]);
}
__device__ __forceinline__ double ddy(const Mesh& M, int j, const Nbr& n, const double* q) {
  const double qc = q[n.c];
  return qdiv(fint(M.y, n.jp, j, qc, q[n.yp]) - fint(M.y, j, n.jm, q[n.ym], qc), M.y.d[j], M.y.rd[j]);
}
// secondDerivative1D(q, 0) and (q, 1) at an interior cell: the denominator is (0.5 d)^2
__device__ __forceinline__ double d2x(const Mesh& M, int i, const Nbr& n, const double* q) {
  const double qc = q[n.c];
  return qdiv(fint(M.x, n.ip, i, qc, q[n.xp]) - 2.0 * qc + fint(M.x, i, n.im, q[n.xm], qc), M.x.hsq[i], M.x.rhsq[i]);
}
__device__ __forceinline__ double d2y(const Mesh& M, int j, const Nbr& n, const double* q) {
  const double qc = q[n.c];
  return qdiv(fint(M.y, n.jp, j, qc, q[n.yp]) - 2.0 * qc + fint(M.y, j, n.jm, q[n.ym], qc), M.y.hsq[j], M.y.rhsq[j]);
}

__global__ void __launch_bounds__(256) k_source(double* dst, const double* t, const double* w, double a, double b,
                                                const double* a_dev, int mode, Mesh M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double mask = inside(M, i, j) ? 1.0 : 0.0;
  const double av = a_dev ? *a_dev : a;
  const double d = dst[c];
  double r;
  if (mode == SRC_ADD) r = d + mask * (av * t[c]);
  else if (mode == SRC_ADD_SCALED) r = d + mask * ((av * t[c]) * b);
  else if (mode == SRC_ADD_WEIGHTED) r = d + mask * ((av * t[c]) * w[c]);
  else r = d - av * t[c];
  dst[c] = r;
}

__global__ void __launch_bounds__(256) k_fheat_pre(MhdSet L, MhdStatic st, const double* bhx, const double* bhy,
                                                   FheatParams p, Mesh M, double* heat) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  if (p.coeff == 0.0) {
    heat[c] = 0.0;
    return;
  }
  const bool in = inside(M, i, j);
  double h = p.coeff;
  if (p.cpow != 0.0) {
    // curl2D(be_x + bi_x, be_y + bi_y) = derivative1D(Y, 0) - derivative1D(X, 1) (derivs.cpp:472-474)
    double curl = 0.0;
    if (in) {
      const Nbr n = nbr(M, i, j);
      const double yc = st.bey[n.c] + L.by[n.c];
      const double yxm = st.bey[n.xm] + L.by[n.xm], yxp = st.bey[n.xp] + L.by[n.xp];
      const double xc = st.bex[n.c] + L.bx[n.c];
      const double xym = st.bex[n.ym] + L.bx[n.ym], xyp = st.bex[n.yp] + L.bx[n.yp];
      const double dY = qdiv(fint(M.x, n.ip, i, yc, yxp) - fint(M.x, i, n.im, yxm, yc), M.x.d[i], M.x.rd[i]);
      const double dX = qdiv(fint(M.y, n.jp, j, xc, xyp) - fint(M.y, j, n.jm, xym, xc), M.y.d[j], M.y.rd[j]);
      curl = dY - dX;
    }
    h = h * pow(p.cc * fabs(curl), p.cpow);
  }
  if (p.bpow != 0.0) {
    const double bx = st.bex[c] + L.bx[c], by = st.bey[c] + L.by[c], bz = st.bez[c] + L.bz[c];
    h = h * pow(sqrt(bx * bx + by * by + bz * bz), p.bpow);
  }
  if (p.npow != 0.0) h = h * pow(L.n[c], p.npow);
  if (p.rpow != 0.0) {
    // radius of curvature of the field lines (fieldheating.cpp:42-48)
    double kx = 0.0, ky = 0.0;
    const double bx = bhx[c], by = bhy[c];
    if (in) {
      const Nbr n = nbr(M, i, j);
      kx = bx * ddx(M, i, n, bhx) + by * ddy(M, j, n, bhx);
      ky = bx * ddx(M, i, n, bhy) + by * ddy(M, j, n, bhy);
    } else {
      kx = bx * 0.0 + by * 0.0;
      ky = kx;
    }
    const double roc = 1.0 / smax(sqrt(kx * kx + ky * ky), 1.0e-16);
    h = h / pow(roc, p.rpow);
  }
  heat[c] = h;
}

__global__ void __launch_bounds__(256) k_fheat_apply(double* te, double* heat, const double* step, Mesh M, int apply) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double mask = inside(M, i, j) ? 1.0 : 0.0;
  const double v = mask * (*step * heat[c]);
  heat[c] = v;
  if (apply) te[c] = te[c] + v;
}

__global__ void __launch_bounds__(256) k_field_sum(const double* bix, const double* biy, const double* bex,
                                                   const double* bey, size_t n, double* bx, double* by) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  bx[c] = bix[c] + bex[c];
  by[c] = biy[c] + bey[c];
}

__global__ void __launch_bounds__(256) k_dc_inner(const double* bx, const double* by, Mesh M, double* dx_bx,
                                                  double* dy_by) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  double a = 0.0, b = 0.0;
  if (inside(M, i, j)) {
    const Nbr n = nbr(M, i, j);
    a = ddx(M, i, n, bx);
    b = ddy(M, j, n, by);
  }
  dx_bx[c] = a;
  dy_by[c] = b;
}

__global__ void __launch_bounds__(256) k_dc_update(const double* bx, const double* by, const double* dx_bx,
                                                   const double* dy_by, double* bix, double* biy, const double* bex,
                                                   const double* bey, const double* coeff, double h, Mesh M,
                                                   double* nbx, double* nby) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const bool in = inside(M, i, j);
  double sx = 0.0, sy = 0.0;
  if (in) {
    const Nbr n = nbr(M, i, j);
    sx = ddx(M, i, n, dy_by) + d2x(M, i, n, bx);
    sy = ddy(M, j, n, dx_bx) + d2y(M, j, n, by);
  }
  const double mh = (in ? 1.0 : 0.0) * h;
  const double x = bix[c] + (mh * coeff[c]) * sx;
  const double y = biy[c] + (mh * coeff[c]) * sy;
  bix[c] = x;
  biy[c] = y;
  nbx[c] = x + bex[c];
  nby[c] = y + bey[c];
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

void launch_source(double* dst, const double* t, const double* w, double a, double b, const double* a_dev, int mode,
                   Mesh mesh, hipStream_t s) {
  hipLaunchKernelGGL(k_source, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, dst, t, w, a, b, a_dev, mode, mesh);
}

void launch_fheat_pre(MhdSet L, MhdStatic st, const double* bhx, const double* bhy, FheatParams p, Mesh mesh,
                      double* heat, hipStream_t s) {
  hipLaunchKernelGGL(k_fheat_pre, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, st, bhx, bhy, p, mesh, heat);
}

void launch_fheat_apply(double* te, double* heat, const double* step, Mesh mesh, int apply, hipStream_t s) {
  hipLaunchKernelGGL(k_fheat_apply, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, te, heat, step, mesh, apply);
}

void launch_field_sum(const double* bix, const double* biy, const double* bex, const double* bey, size_t n,
                      double* bx, double* by, hipStream_t s) {
  hipLaunchKernelGGL(k_field_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bix, biy, bex, bey, n, bx, by);
}

void launch_dc_inner(const double* bx, const double* by, Mesh mesh, double* dx_bx, double* dy_by, hipStream_t s) {
  hipLaunchKernelGGL(k_dc_inner, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, bx, by, mesh, dx_bx, dy_by);
}

void launch_dc_update(const double* bx, const double* by, const double* dx_bx, const double* dy_by, double* bix,
                      double* biy, const double* bex, const double* bey, const double* coeff, double h, Mesh mesh,
                      double* nbx, double* nby, hipStream_t s) {
  hipLaunchKernelGGL(k_dc_update, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, bx, by, dx_bx, dy_by, bix, biy, bex, bey,
                     coeff, h, mesh, nbx, nby);
}

}  // namespace spruce
