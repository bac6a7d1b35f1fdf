// Source-term module kernels for gfx950 (reference: source/modules/solar/localizedheating.cpp,
// massinjection.cpp, momentuminjection.cpp, ambientheatingsink.cpp, fieldheating.cpp,
// divcleaning.cpp).  Each whole-grid expression of the reference is evaluated per cell with the
// same operation order; derivatives are the reference's derivative1D / secondDerivative1D
// (derivs.cpp:376-455): zero off the interior, face values by boundaryInterpolate.
#include <algorithm>
#include <cfloat>

#include "arith.hpp"
#include "crpow.hpp"
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

// ---- anomalous resistivity (anomalousresistivity.cpp) -------------------------------------------

// C/(4 pi) |curl B_i| (computeCurrentDensity, :272-276); 0 off the interior, as the derivatives
__global__ void __launch_bounds__(256) k_ar_current(const double* bix, const double* biy, const double* biz, double cc,
                                                    Mesh M, double* J) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  double curl = 0.0, cz0 = 0.0, cz1 = 0.0;
  if (inside(M, i, j)) {
    const Nbr n = nbr(M, i, j);
    curl = ddx(M, i, n, biy) - ddy(M, j, n, bix);
    cz0 = ddy(M, j, n, biz);
    cz1 = -ddx(M, i, n, biz);
  }
  J[c] = cc * sqrt(curl * curl + cz0 * cz0 + cz1 * cz1);
}

// b = be + bi and |b| in the plane
__global__ void __launch_bounds__(256) k_ar_field(const double* bex, const double* bey, const double* bix,
                                                  const double* biy, size_t n, double* bx, double* by, double* bm) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double x = bex[c] + bix[c], y = bey[c] + biy[c];
  bx[c] = x;
  by[c] = y;
  bm[c] = sqrt(x * x + y * y);
}

// frobenius template, first pass: |grad b|_F / |b| on every cell
__global__ void __launch_bounds__(256) k_ar_frob_ratio(const double* bx, const double* by, const double* bm, Mesh M,
                                                       double* ratio) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  double a = 0.0, b = 0.0, e = 0.0, f = 0.0;
  if (inside(M, i, j)) {
    const Nbr n = nbr(M, i, j);
    a = ddx(M, i, n, bx);
    b = ddy(M, j, n, bx);
    e = ddx(M, i, n, by);
    f = ddy(M, j, n, by);
  }
  ratio[c] = sqrt(a * a + b * b + e * e + f * f) / bm[c];
}

// frobenius template, second pass: min(coeff * laplacian(ratio)^2, 1)
__global__ void __launch_bounds__(256) k_ar_frob_tmpl(const double* ratio, double coeff, Mesh M, double* tmpl) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  double lap = 0.0;
  if (inside(M, i, j)) {
    const Nbr n = nbr(M, i, j);
    lap = d2x(M, i, n, ratio) + d2y(M, j, n, ratio);
  }
  tmpl[c] = smin(coeff * (lap * lap), 1.0);
}

// the localized argmin (argminLocalized, :202-225) over the box [il, iu] x [jl, ju]: cells within
// radius of the centre, smallest value, lowest index among equals; one block
__global__ void __launch_bounds__(256) k_ar_argmin_box(const double* q, const double* px, const double* py, int ci,
                                                       int cj, double radius, int il, int iu, int jl, int ju, int ny,
                                                       double* out) {
  double v = DBL_MAX;
  long long idx = -1;
  const double cx = px[(size_t)ci * ny + cj], cy = py[(size_t)ci * ny + cj];
  const int w = ju - jl + 1, cnt = (iu - il + 1) * w;
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int i = il + t / w, j = jl + t % w;
    const size_t c = (size_t)i * ny + j;
    const double dx = cx - px[c], dy = cy - py[c];
    if (sqrt(dx * dx + dy * dy) <= radius && (q[c] < v || (q[c] == v && (long long)c < idx))) {
      v = q[c];
      idx = (long long)c;
    }
  }
  __shared__ double sv[256];
  __shared__ long long si[256];
  sv[threadIdx.x] = v;
  si[threadIdx.x] = idx;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double ov = sv[threadIdx.x + off];
      const long long oi = si[threadIdx.x + off];
      if (oi >= 0 && (ov < sv[threadIdx.x] || (ov == sv[threadIdx.x] && (si[threadIdx.x] < 0 || oi < si[threadIdx.x])))) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sv[0];
    out[1] = (double)si[0];
    out[2] = q[(size_t)ci * ny + cj];
  }
}

__global__ void __launch_bounds__(256) k_ar_flood_seed(const double* q, size_t n, size_t start, double thr, double* res) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  res[c] = (c == start && q[c] < thr) ? 1.0 : 0.0;
}

// one sweep of the flood fill (Grid::floodFill, grid.cpp:487-509): a cell below the threshold joins
// the region when a 4-neighbour is in it.  Writes only ever set 1, so reading neighbours that the
// same sweep updated is harmless; *changed is set when any cell joined.
__global__ void __launch_bounds__(256) k_ar_flood_sweep(const double* q, double thr, int nx, int ny, double* res,
                                                        int* changed) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const size_t c = (size_t)i * ny + j;
  if (res[c] >= 1.0 || !(q[c] < thr)) return;
  const bool nb = (i > 0 && res[c - ny] >= 1.0) || (i + 1 < nx && res[c + ny] >= 1.0) ||
                  (j > 0 && res[c - 1] >= 1.0) || (j + 1 < ny && res[c + 1] >= 1.0);
  if (nb) {
    res[c] = 1.0;
    *changed = 1;
  }
}

// circularMask and currentThresholdMask (:227-256) applied to the template; both loop columns over
// the row count, so columns j >= xdim get 0
__global__ void __launch_bounds__(256) k_ar_masks(double* tmpl, const double* J, const double* px, const double* py,
                                                  size_t cnull, double max_radius, double min_current, double ramp,
                                                  Mesh M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  double t = tmpl[c];
  if (max_radius > 0.0) {
    double m = 0.0;
    if (j < M.nx) {
      const double dx = px[c] - px[cnull], dy = py[c] - py[cnull];
      if (sqrt(dx * dx + dy * dy) <= max_radius) m = 1.0;
    }
    t = m * t;
  }
  if (min_current > 0.0) {
    double m = 0.0;
    if (j < M.nx) {
      double v = (J[c] - min_current) / ramp + 0.5;
      v = smax(v, 0.0);
      m = smin(v, 1.0);
    }
    t = m * t;
  }
  tmpl[c] = t;
}

// metric smoothing (:178-190): sum of kernel(k, l) * t(i+k, j+l) over the kernel, k outer, l inner,
// skipping cells outside the grid
__global__ void __launch_bounds__(256) k_ar_smooth(const double* t, const double* kern, int kr, Mesh M, double* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const int ks = 2 * kr + 1;
  double s = 0.0;
  for (int k = -kr; k <= kr; ++k) {
    if (i + k < 0 || i + k >= M.nx) continue;
    for (int l = -kr; l <= kr; ++l) {
      if (j + l < 0 || j + l >= M.ny) continue;
      s += kern[(k + kr) * ks + (l + kr)] * t[(size_t)(i + k) * M.ny + (j + l)];
    }
  }
  out[(size_t)i * M.ny + j] = s;
}

// frobenius final shaping: min(10 t, 1)^1.5 (the reference's libm pow, here correctly rounded)
__global__ void __launch_bounds__(256) k_ar_frob_final(double* t, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  t[c] = pow_1p5(smin(10.0 * t[c], 1.0));
}

// computeDiffusion (:124-156) from the current density of the given field
__global__ void __launch_bounds__(256) k_ar_diffusion(const double* J, const double* n, ArParams p, Mesh M,
                                                      double* diff) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  double d;
  if (p.model == AR_TIME_SCALE) {
    const double dx = M.x.d[i], dy = M.y.d[j];
    d = 1.0 / (1.0 / (dx * dx) + 1.0 / (dy * dy)) / 2. / p.ts;
  } else if (p.model == AR_SYNTELIS) {
    double s = J[c] / p.p2;
    if (s < 1.0) s = 0.0;
    d = p.p1 * s + p.p0;
  } else {
    const double vr = J[c] / n[c] / E_CHARGE / p.p0;
    const double w = vr - 1.0;
    d = smin(p.p1 * (w * w), p.p2);
    if (vr <= 1.0) d = 0.0;
  }
  diff[c] = d;
}

// coeff = mask * template * diffusivity, and template * diffusivity for the subcycle count
__global__ void __launch_bounds__(256) k_ar_coeff(const double* tmpl, const double* diff, Mesh M, double* coeff,
                                                  double* td) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double mask = inside(M, i, j) ? 1.0 : 0.0;
  coeff[c] = mask * tmpl[c] * diff[c];
  td[c] = tmpl[c] * diff[c];
}

// computeTimeDerivatives (:79-102): d0..d2 for bi_x, bi_y, bi_z and d3 = Joule heating
__global__ void __launch_bounds__(256) k_ar_rhs(ArRhsArgs a) {
  const Mesh& M = a.mesh;
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double coeff = a.coeff[c];
  const double eta = a.eta_factor * coeff;
  const double Jc = a.J[c];
  double l0 = 0.0, l1 = 0.0, l2 = 0.0, x0 = 0.0, x1 = 0.0, x2 = 0.0;
  if (inside(M, i, j)) {
    const Nbr n = nbr(M, i, j);
    l0 = d2x(M, i, n, a.bix) + d2y(M, j, n, a.bix);
    l1 = d2x(M, i, n, a.biy) + d2y(M, j, n, a.biy);
    l2 = d2x(M, i, n, a.biz) + d2y(M, j, n, a.biz);
    if (a.grad_corr) {
      const double ge0 = ddx(M, i, n, a.coeff), ge1 = ddy(M, j, n, a.coeff);
      const double cb = ddx(M, i, n, a.bix) - ddy(M, j, n, a.bix);   // curl2D(bi_x, bi_x) (:87)
      const double mz = -1.0 * cb;
      x0 = (-mz) * ge1;
      x1 = mz * ge0;
      const double cz0 = ddy(M, j, n, a.biz), cz1 = -ddx(M, i, n, a.biz);
      x2 = ge0 * cz1 - ge1 * cz0;
    }
  } else if (a.grad_corr) {
    // every derivative is 0 off the interior: (-(-1*0))*0 and the like
    const double mz = -1.0 * (0.0 - 0.0);
    x0 = (-mz) * 0.0;
    x1 = mz * 0.0;
    x2 = 0.0 * (-0.0) - 0.0 * 0.0;
  }
  const double t0 = coeff * (a.lbx[c] + l0), t1 = coeff * (a.lby[c] + l1), t2 = coeff * (a.lbz[c] + l2);
  if (a.grad_corr) {
    a.d0[c] = x0 + t0;
    a.d1[c] = x1 + t1;
    a.d2[c] = x2 + t2;
  } else {
    a.d0[c] = t0;
    a.d1[c] = t1;
    a.d2[c] = t2;
  }
  a.d3[c] = eta * Jc * Jc;
}

// the laplacian of a static field (be), 0 off the interior
__global__ void __launch_bounds__(256) k_ar_laplacian(const double* q, Mesh M, double* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  double l = 0.0;
  if (inside(M, i, j)) {
    const Nbr n = nbr(M, i, j);
    l = d2x(M, i, n, q) + d2y(M, j, n, q);
  }
  out[(size_t)i * M.ny + j] = l;
}

// out_k = base_k + s * d_k for the four fields (bi_x, bi_y, bi_z, te); rk4: s * (k1 + 2 k2 + 2 k3 + k4) / 6
__global__ void __launch_bounds__(256) k_ar_axpy(ArAxpyArgs a) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n) return;
  for (int q = 0; q < a.nf; ++q) {
    double inc;
    if (a.rk4) inc = a.s * (a.k1[q][c] + 2.0 * a.k2[q][c] + 2.0 * a.k3[q][c] + a.k4[q][c]) / 6.0;
    else inc = a.s * a.k1[q][c];
    a.out[q][c] = a.base[q][c] + inc;
  }
}

// whole-grid max (std::max) and the minimum of the diffusive time scale over [xl..xu] x [yl..yu]
// (computeNumSubcycles, :105-122): out[0] = max(td), out[1] = min((1/(1/dx^2+1/dy^2))/2/td)
__global__ void __launch_bounds__(256) k_ar_scales(const double* td, Mesh M, int xl, int xu, int yl, int yu,
                                                   double* part_max, double* part_min) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double mx = -DBL_MAX, mn = DBL_MAX;
  if (i < M.nx && j < M.ny) {
    const double v = td[(size_t)i * M.ny + j];
    mx = (mx < v) ? v : mx;
    if (i >= xl && i <= xu && j >= yl && j <= yu) {
      const double dx = M.x.d[i], dy = M.y.d[j];
      const double t = 1.0 / (1.0 / (dx * dx) + 1.0 / (dy * dy)) / 2. / v;
      mn = (t < mn) ? t : mn;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double a = __shfl_xor(mx, off, 64), b = __shfl_xor(mn, off, 64);
    mx = (mx < a) ? a : mx;
    mn = (b < mn) ? b : mn;
  }
  __shared__ double pa[16], pb[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  if ((lane & 63) == 0) { pa[lane >> 6] = mx; pb[lane >> 6] = mn; }
  __syncthreads();
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    for (int k = 1; k < nw; ++k) { mx = (mx < pa[k]) ? pa[k] : mx; mn = (pb[k] < mn) ? pb[k] : mn; }
    part_max[blockIdx.y * gridDim.x + blockIdx.x] = mx;
    part_min[blockIdx.y * gridDim.x + blockIdx.x] = mn;
  }
}

__global__ void __launch_bounds__(256) k_ar_scales_final(const double* pmax, const double* pmin, int nb, double* out) {
  double mx = -DBL_MAX, mn = DBL_MAX;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    mx = (mx < pmax[k]) ? pmax[k] : mx;
    mn = (pmin[k] < mn) ? pmin[k] : mn;
  }
  __shared__ double sa[256], sb[256];
  sa[threadIdx.x] = mx;
  sb[threadIdx.x] = mn;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      sa[threadIdx.x] = (sa[threadIdx.x] < sa[threadIdx.x + off]) ? sa[threadIdx.x + off] : sa[threadIdx.x];
      sb[threadIdx.x] = (sb[threadIdx.x + off] < sb[threadIdx.x]) ? sb[threadIdx.x + off] : sb[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = sa[0]; out[1] = sb[0]; }
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }


// b_hat of one cell as IdealMHD::recomputeDerivedVarsFromEvolvedVars builds it
// (idealmhd.cpp:255-262): b = be + bi, b_hat = b / |b|, and 0 where |b| is 0
__device__ inline void of_bhat(const MhdSet& L, const MhdStatic& st, size_t c, double& hx, double& hy) {
  const double bx = st.bex[c] + L.bx[c], by = st.bey[c] + L.by[c], bz = st.bez[c] + L.bz[c];
  const double bm = sqrt(bx * bx + by * by + bz * bz);
  hx = bm == 0.0 ? 0.0 : bx / bm;
  hy = bm == 0.0 ? 0.0 : by / bm;
}

// the better of two (value, row-order index) pairs for a left fold with std::max: the larger
// value, and of equal values the earlier one; NaN never wins
__device__ inline void of_pick(double& v, long long& k, double ov, long long ok) {
  if (ok >= 0 && (ov > v || (ov == v && (k < 0 || ok < k)))) {
    v = ov;
    k = ok;
  }
}

__global__ void __launch_bounds__(256) k_outflow_max(MhdSet L, MhdStatic st, OutflowParams p, int ny,
                                                     double* pv, double* pk) {
  const int w = p.wyu - p.wyl + 1;
  const long long cnt = (long long)(p.wxu - p.wxl + 1) * w;
  double v = -INFINITY;
  long long k = -1;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += (long long)gridDim.x * blockDim.x) {
    const int i = p.wxl + (int)(t / w), j = p.wyl + (int)(t % w);
    const size_t c = (size_t)i * ny + j;
    double hx, hy;
    of_bhat(L, st, c, hx, hy);
    const double vx = L.mx[c] / L.rho[c], vy = L.my[c] / L.rho[c];
    double cur = hx * vx + hy * vy;
    // orient towards the boundary (boundaryoutflow.cpp:225-229)
    if (p.side == OF_X1 && hx > 0.0) cur *= -1.0;
    else if (p.side == OF_X2 && hx < 0.0) cur *= -1.0;
    else if (p.side == OF_Y2 && hy < 0.0) cur *= -1.0;
    else if (p.side == OF_Y1 && hy > 0.0) cur *= -1.0;
    of_pick(v, k, cur, (long long)c);
  }
  __shared__ double sv[256];
  __shared__ long long sk[256];
  sv[threadIdx.x] = v;
  sk[threadIdx.x] = k;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) of_pick(sv[threadIdx.x], sk[threadIdx.x], sv[threadIdx.x + off], sk[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pk[blockIdx.x] = (double)sk[0];
  }
}

__global__ void __launch_bounds__(256) k_outflow_final(const double* pv, const double* pk, int nb, OutflowParams p,
                                                       int set_acc, double* sc) {
  double v = -INFINITY;
  long long k = -1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) of_pick(v, k, pv[b], (long long)pk[b]);
  __shared__ double sv[256];
  __shared__ long long sk[256];
  sv[threadIdx.x] = v;
  sk[threadIdx.x] = k;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) of_pick(sv[threadIdx.x], sk[threadIdx.x], sv[threadIdx.x + off], sk[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // the fold starts from -dynamic_target_speed (:217) and takes a cell only when it is larger
    const double init = -1.0 * p.target;
    const double m = (sk[0] >= 0 && init < sv[0]) ? sv[0] : init;
    sc[0] = m;
    if (set_acc) {
      double a = p.amax;
      if (p.dynamic) {  // :41-45
        a = (p.target - m) / p.dtime;
        a = (p.amax < a) ? p.amax : a;
        a = (a < 0.0) ? 0.0 : a;
      }
      sc[1] = a;
    }
  }
}

__global__ void __launch_bounds__(256) k_outflow_apply(MhdSet L, MhdStatic st, const double* tmpl, const double* sc,
                                                       const double* step, OutflowParams p, Mesh M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double a = *step * (sc[1] * tmpl[c]);
  const double rho = L.rho[c];
  if (p.aligned) {
    double hx, hy;
    of_bhat(L, st, c, hx, hy);
    L.mx[c] = L.mx[c] + (a * hx) * rho;
    L.my[c] = L.my[c] + (a * hy) * rho;
  } else if (p.side == OF_X1 || p.side == OF_X2) {
    L.mx[c] = L.mx[c] + a * rho;
  } else {
    L.my[c] = L.my[c] + a * rho;
  }
}

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

static inline dim3 ar_blocks(const Mesh& m) { return dim3((m.ny + 63) / 64, (m.nx + 3) / 4); }
static inline dim3 ar_lin(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_ar_current(const double* bix, const double* biy, const double* biz, double cc, Mesh mesh, double* J,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_ar_current, ar_blocks(mesh), dim3(64, 4), 0, s, bix, biy, biz, cc, mesh, J);
}
void launch_ar_field(const double* bex, const double* bey, const double* bix, const double* biy, size_t n, double* bx,
                     double* by, double* bm, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_field, ar_lin(n), dim3(256), 0, s, bex, bey, bix, biy, n, bx, by, bm);
}
void launch_ar_frob(const double* bx, const double* by, const double* bm, double coeff, Mesh mesh, double* ratio,
                    double* tmpl, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_frob_ratio, ar_blocks(mesh), dim3(64, 4), 0, s, bx, by, bm, mesh, ratio);
  hipLaunchKernelGGL(k_ar_frob_tmpl, ar_blocks(mesh), dim3(64, 4), 0, s, (const double*)ratio, coeff, mesh, tmpl);
}
void launch_ar_argmin_box(const double* q, const double* px, const double* py, int ci, int cj, double radius, int il,
                          int iu, int jl, int ju, int ny, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_argmin_box, dim3(1), dim3(256), 0, s, q, px, py, ci, cj, radius, il, iu, jl, ju, ny, out);
}
void launch_ar_flood_seed(const double* q, size_t n, size_t start, double thr, double* res, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_flood_seed, ar_lin(n), dim3(256), 0, s, q, n, start, thr, res);
}
void launch_ar_flood_sweep(const double* q, double thr, Mesh mesh, double* res, int* changed, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_flood_sweep, ar_blocks(mesh), dim3(64, 4), 0, s, q, thr, mesh.nx, mesh.ny, res, changed);
}
void launch_ar_masks(double* tmpl, const double* J, const double* px, const double* py, size_t cnull, double max_radius,
                     double min_current, double ramp, Mesh mesh, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_masks, ar_blocks(mesh), dim3(64, 4), 0, s, tmpl, J, px, py, cnull, max_radius, min_current,
                     ramp, mesh);
}
void launch_ar_smooth(const double* t, const double* kern, int kr, Mesh mesh, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_smooth, ar_blocks(mesh), dim3(64, 4), 0, s, t, kern, kr, mesh, out);
}
void launch_ar_frob_final(double* t, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_frob_final, ar_lin(n), dim3(256), 0, s, t, n);
}
void launch_ar_diffusion(const double* J, const double* n, ArParams p, Mesh mesh, double* diff, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_diffusion, ar_blocks(mesh), dim3(64, 4), 0, s, J, n, p, mesh, diff);
}
void launch_ar_coeff(const double* tmpl, const double* diff, Mesh mesh, double* coeff, double* td, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_coeff, ar_blocks(mesh), dim3(64, 4), 0, s, tmpl, diff, mesh, coeff, td);
}
void launch_ar_rhs(const ArRhsArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_rhs, ar_blocks(a.mesh), dim3(64, 4), 0, s, a);
}
void launch_ar_laplacian(const double* q, Mesh mesh, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_laplacian, ar_blocks(mesh), dim3(64, 4), 0, s, q, mesh, out);
}
void launch_ar_axpy(const ArAxpyArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_ar_axpy, ar_lin(a.n), dim3(256), 0, s, a);
}
int ar_scale_blocks(const Mesh& m) {
  const dim3 g = ar_blocks(m);
  return (int)(g.x * g.y);
}
void launch_ar_scales(const double* td, Mesh mesh, int xl, int xu, int yl, int yu, double* partial, double* out,
                      hipStream_t s) {
  const int nb = ar_scale_blocks(mesh);
  hipLaunchKernelGGL(k_ar_scales, ar_blocks(mesh), dim3(64, 4), 0, s, td, mesh, xl, xu, yl, yu, partial, partial + nb);
  hipLaunchKernelGGL(k_ar_scales_final, dim3(1), dim3(256), 0, s, (const double*)partial, (const double*)(partial + nb),
                     nb, out);
}


int outflow_blocks(const OutflowParams& p) {
  const long long cnt = (long long)(p.wxu - p.wxl + 1) * (p.wyu - p.wyl + 1);
  return (int)std::max(1LL, std::min(256LL, (cnt + 255) / 256));
}
void launch_outflow_max(MhdSet L, MhdStatic st, OutflowParams p, Mesh mesh, double* partial, double* sc, int set_acc,
                        hipStream_t s) {
  const int nb = outflow_blocks(p);
  hipLaunchKernelGGL(k_outflow_max, dim3(nb), dim3(256), 0, s, L, st, p, mesh.ny, partial, partial + nb);
  hipLaunchKernelGGL(k_outflow_final, dim3(1), dim3(256), 0, s, (const double*)partial,
                     (const double*)(partial + nb), nb, p, set_acc, sc);
}
void launch_outflow_apply(MhdSet L, MhdStatic st, const double* tmpl, const double* sc, const double* step,
                          OutflowParams p, Mesh mesh, hipStream_t s) {
  hipLaunchKernelGGL(k_outflow_apply, ar_blocks(mesh), dim3(64, 4), 0, s, L, st, tmpl, sc, step, p, mesh);
}

}  // namespace spruce
