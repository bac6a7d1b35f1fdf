// Thermal conduction and radiative-loss kernels for gfx950 (reference:
// source/modules/solar/thermalconduction.cpp, radiativelosses.cpp).  Operation order and rounding
// follow the reference whole-grid expressions; powers are correctly rounded (crpow.hpp).
#include <cfloat>

#include "arith.hpp"
#include "crpow.hpp"
#include "solar.hpp"
#include "strip.hpp"
#include "tc_cell.hpp"
#include "tc_strip.hpp"

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
// temp = max((gamma-1) te / (2 K_B n), temp_min) (thermalconduction.cpp:109; idealmhd.cpp:254)
__device__ __forceinline__ double temp_of(double te, double n, const MhdConst& K) {
  return smax(te * K.gm1 / (n * (2.0 * K_B)), K.tmin);
}
// boundaryInterpolate across the face whose upper cell is c on axis ax (derivs.cpp:477-487)
__device__ __forceinline__ double fint(const Axis& ax, int c, int cm1, double qlo, double qup) {
  return qdiv(qlo * ax.h[c] + qup * ax.h[cm1], ax.fden[c], ax.rfden[c]);
}

__device__ __forceinline__ double block_reduce(double v, bool is_min) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = is_min ? ((o < v) ? o : v) : ((v < o) ? o : v);
  }
  __shared__ double part[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  __syncthreads();
  if ((lane & 63) == 0) part[lane >> 6] = v;
  __syncthreads();
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    double m = part[0];
    for (int k = 1; k < nw; ++k) m = is_min ? ((part[k] < m) ? part[k] : m) : ((m < part[k]) ? part[k] : m);
    v = m;
  }
  return v;
}

__global__ void __launch_bounds__(256) k_bhat(MhdSet L, MhdStatic st, Mesh M, double* bhx, double* bhy) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double bx = st.bex[c] + L.bx[c], by = st.bey[c] + L.by[c], bz = st.bez[c] + L.bz[c];
  const double bm = sqrt(bx * bx + by * by + bz * bz);
  bhx[c] = bm == 0.0 ? 0.0 : bx / bm;
  bhy[c] = bm == 0.0 ? 0.0 : by / bm;
}

__global__ void __launch_bounds__(256) k_bhat_derivs(const double* bhx, const double* bhy, Mesh M, double* dbxx,
                                                     double* dbxy, double* dbyx, double* dbyy) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (inside(M, i, j)) {
    const int ny = M.ny;
    const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
    const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
    const size_t cxm = (size_t)im * ny + j, cxp = (size_t)ip * ny + j, cym = (size_t)i * ny + jm, cyp = (size_t)i * ny + jp;
    const double dx = M.x.d[i], rdx = M.x.rd[i], dy = M.y.d[j], rdy = M.y.rd[j];
    const double bx = bhx[c], by = bhy[c];
    r[0] = qdiv(fint(M.x, ip, i, bx, bhx[cxp]) - fint(M.x, i, im, bhx[cxm], bx), dx, rdx);
    r[1] = qdiv(fint(M.y, jp, j, bx, bhx[cyp]) - fint(M.y, j, jm, bhx[cym], bx), dy, rdy);
    r[2] = qdiv(fint(M.x, ip, i, by, bhy[cxp]) - fint(M.x, i, im, bhy[cxm], by), dx, rdx);
    r[3] = qdiv(fint(M.y, jp, j, by, bhy[cyp]) - fint(M.y, j, jm, bhy[cym], by), dy, rdy);
  }
  dbxx[c] = r[0]; dbxy[c] = r[1]; dbyx[c] = r[2]; dbyy[c] = r[3];
}

// Field-aligned conductive flux at an interior cell and its saturated magnitude
// (thermalconduction.cpp:198-232).
struct SatFlux { double fx, fy, fm, sfm; };
__device__ __forceinline__ SatFlux sat_flux(double T, double gx, double gy, double rho, double bx, double by,
                                            double dx, double dy, const TcParams& p, const MhdConst& K) {
  const double kmax = qdiv(dx * dy * K_B * qdiv(rho, K.m, K.rm), p.dsm, 1.0 / p.dsm);
  const double kap = smin(pow_2p5(T) * p.k0, kmax);
  const double cfx = (-kap) * gx, cfy = (-kap) * gy;
  const double fmag = cfx * bx + cfy * by;
  SatFlux s;
  s.fx = fmag * bx;
  s.fy = fmag * by;
  const double sat = qdiv(rho, K.m, K.rm) * p.c_sat * pow_1p5(T * K_B) / p.sqrt_me;
  s.fm = sqrt(s.fx * s.fx + s.fy * s.fy);
  const double scale = sat / sqrt(sat * sat + s.fm * s.fm);
  const double sx = s.fx * scale, sy = s.fy * scale;
  s.sfm = sqrt(sx * sx + sy * sy);
  return s;
}

// Central gradient of T at (i,j) from the temperatures of its four neighbours.
__device__ __forceinline__ void grad_at(const Mesh& M, int i, int j, double Tc, double Txm, double Txp, double Tym,
                                        double Typ, double& gx, double& gy) {
  const int ip = wrp(i + 1, M.nx, M.x.per), jp = wrp(j + 1, M.ny, M.y.per);
  const int im = wrp(i - 1, M.nx, M.x.per), jm = wrp(j - 1, M.ny, M.y.per);
  gx = qdiv(fint(M.x, ip, i, Tc, Txp) - fint(M.x, i, im, Txm, Tc), M.x.d[i], M.x.rd[i]);
  gy = qdiv(fint(M.y, jp, j, Tc, Typ) - fint(M.y, j, jm, Tym, Tc), M.y.d[j], M.y.rd[j]);
}

__global__ void __launch_bounds__(256) k_tc_pre(MhdSet L, const double* bhx, const double* bhy, TcParams p,
                                                MhdConst K, Mesh M, double* pmin, double* pmax) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double vmin = DBL_MAX, vmax = -DBL_MAX;
  if (i < M.nx && j < M.ny && inside(M, i, j)) {
    const size_t c = (size_t)i * M.ny + j;
    auto T = [&](int a, int b) {
      const size_t q = (size_t)wrp(a, M.nx, M.x.per) * M.ny + wrp(b, M.ny, M.y.per);
      return temp_of(L.te[q], L.n[q], K);
    };
    const double Tc = T(i, j);
    const double rho = L.rho[c];
    double dts;
    if (!p.saturation) {
      // K_B/(w kappa_0) * (rho/m) * d_x * d_y / T^2.5 (thermalconduction.cpp:183)
      dts = qdiv(rho, K.m, K.rm) * p.c1 * M.x.d[i] * M.y.d[j] / pow_2p5(Tc);
    } else {
      double gx, gy;
      grad_at(M, i, j, Tc, T(i - 1, j), T(i + 1, j), T(i, j - 1), T(i, j + 1), gx, gy);
      const double bx = bhx[c], by = bhy[c];
      const double ftg = gx * bx + gy * by;
      vmax = fabs(ftg);
      const SatFlux s = sat_flux(Tc, gx, gy, rho, bx, by, M.x.d[i], M.y.d[j], p, K);
      const double kap = fabs(s.sfm / ftg);                  // saturatedKappa (:236-248)
      dts = K_B / kap * qdiv(rho, K.m, K.rm) * M.x.d[i] * M.y.d[j];
    }
    vmin = (dts < vmin) ? dts : vmin;
  } else if (i < M.nx && j < M.ny) {
    vmax = 0.0;  // |b_hat . grad T| is zero outside the interior
  }
  const double a = block_reduce(vmin, true);
  const double b = block_reduce(vmax, false);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    pmin[blk] = a;
    pmax[blk] = b;
  }
}

__global__ void __launch_bounds__(1024) k_finish_minmax(const double* pmin, const double* pmax, int n, double* out) {
  double a = DBL_MAX, b = -DBL_MAX;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    a = (pmin[k] < a) ? pmin[k] : a;
    b = (b < pmax[k]) ? pmax[k] : b;
  }
  a = block_reduce(a, true);
  b = block_reduce(b, false);
  if (threadIdx.x == 0) { out[0] = a; out[1] = b; }
}

__global__ void __launch_bounds__(256) k_tc_sat(const double* te, const double* n, const double* rho, const double* bhx,
                                                const double* bhy, TcParams p, MhdConst K, Mesh M, double* rx,
                                                double* ry, double* coef) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  if (!inside(M, i, j)) { rx[c] = 0.0; ry[c] = 0.0; coef[c] = 1.0; return; }
  auto T = [&](int a, int b) {
    const size_t q = (size_t)wrp(a, M.nx, M.x.per) * M.ny + wrp(b, M.ny, M.y.per);
    return temp_of(te[q], n[q], K);
  };
  const double Tc = T(i, j);
  double gx, gy;
  grad_at(M, i, j, Tc, T(i - 1, j), T(i + 1, j), T(i, j - 1), T(i, j + 1), gx, gy);
  const SatFlux s = sat_flux(Tc, gx, gy, rho[c], bhx[c], bhy[c], M.x.d[i], M.y.d[j], p, K);
  rx[c] = s.fx;
  ry[c] = s.fy;
  coef[c] = s.fm != 0.0 ? s.sfm / s.fm : 1.0;                  // thermalconduction.cpp:264-268
}

// ThermalConduction::thermalEnergyDerivative at an interior cell (thermalconduction.cpp:158-176),
// from the temperatures of its 3x3 neighbourhood: the unsaturated part (tc_cell.hpp), then the
// saturation factor and correction.
__device__ double tc_energy_derivative(const TcStepArgs& a, int i, int j, const double T[3][3]) {
  const Mesh& M = a.mesh;
  const int ny = M.ny;
  const int im = wrp(i - 1, M.nx, M.x.per), ip = wrp(i + 1, M.nx, M.x.per);
  const int jm = wrp(j - 1, M.ny, M.y.per), jp = wrp(j + 1, M.ny, M.y.per);
  const double dx = M.x.d[i], rdx = M.x.rd[i], dy = M.y.d[j], rdy = M.y.rd[j];
  // b_hat and its derivatives (frozen for the step, computed once by k_bhat_derivs)
  const size_t c = (size_t)i * ny + j;
  const size_t cxm = (size_t)im * ny + j, cxp = (size_t)ip * ny + j, cym = (size_t)i * ny + jm, cyp = (size_t)i * ny + jp;
  double result = tcc::tc_core<double>(M, i, j, T, a.bhx[c], a.bhy[c], a.dbxx[c], a.dbxy[c], a.dbyx[c], a.dbyy[c],
                                       a.p.k0);
  if (a.p.saturation) {                                         // thermalconduction.cpp:171-174, 255-271
    const double cc = a.coef[c];
    const double dcx = qdiv(fint(M.x, ip, i, cc, a.coef[cxp]) - fint(M.x, i, im, a.coef[cxm], cc), dx, rdx);
    const double dcy = qdiv(fint(M.y, jp, j, cc, a.coef[cyp]) - fint(M.y, j, jm, a.coef[cym], cc), dy, rdy);
    const double add = -1.0 * (dcx * a.rx[c] + dcy * a.ry[c]);
    result = cc * result + add;
  }
  return result;
}

// One conduction update.  The block's temperatures and a one-cell halo are computed once into LDS
// (each temperature is a division), then every cell reads its 3x3 neighbourhood from there.
__global__ void __launch_bounds__(256) k_tc_step(TcStepArgs a) {
  constexpr int BX = 64, BY = 4;
  __shared__ double Ts[BY + 2][BX + 2];
  const Mesh& M = a.mesh;
  const int j0 = blockIdx.x * BX, i0 = blockIdx.y * BY;
  for (int t = threadIdx.y * BX + threadIdx.x; t < (BY + 2) * (BX + 2); t += BX * BY) {
    const int r = t / (BX + 2), q = t - r * (BX + 2);
    const int ii = wrp(i0 - 1 + r, M.nx, M.x.per), jj = wrp(j0 - 1 + q, M.ny, M.y.per);
    double v = 0.0;
    if (ii >= 0 && ii < M.nx && jj >= 0 && jj < M.ny) {
      const size_t cc = (size_t)ii * M.ny + jj;
      v = temp_of(a.te_src[cc], a.n[cc], a.k);
    }
    Ts[r][q] = v;
  }
  __syncthreads();
  const int j = j0 + threadIdx.x, i = i0 + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double base = a.te_base[c];
  if (!inside(M, i, j)) {
    // mask = 0: te + 0*dE; every derivative is zero there, so dE = 0
    a.te_out[c] = smax(base + 0.0, a.k.temin);
    if (a.mode == TC_STORE_K) a.k_out[c] = 0.0;
    return;
  }
  double T[3][3];
#pragma unroll
  for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
    for (int b1 = 0; b1 < 3; ++b1) T[a1][b1] = Ts[threadIdx.y + a1][threadIdx.x + b1];
  double dE = tc_energy_derivative(a, i, j, T);
  double v;
  if (a.mode == TC_RK4_FINAL) {
    const double sum = a.k1[c] + 2.0 * a.k2[c] + 2.0 * a.k3[c] + dE;
    v = base + qdiv(1.0 * a.h * sum, 6.0, a.k.r6);
  } else {
    if (a.mode == TC_STORE_K) a.k_out[c] = dE;
    v = base + 1.0 * (a.mult * a.h) * dE;
  }
  a.te_out[c] = smax(v, a.k.temin);
}

// Strip form of one conduction update (the production path of launch_tc_step).  One 64-lane
// wavefront owns 64 consecutive columns (y) and marches down R rows (x): lane l holds column
// 62*tile + l - 1, lanes 1..62 write outputs and lanes 0, 63 are the y-halo.  Every face value the
// energy derivative takes -- the interpolants of T, of dT/dx along y, of dT/dy along x and of b_hat
// along both axes -- is computed once and shared by the two cells of the face: along y through DPP
// lane shifts, along x by keeping the previous row's faces in registers.  The derivatives of the
// frozen b_hat are formed here from b_hat, so a subcycle reads te, n, b_hat_x, b_hat_y and writes
// te.  Each value is the same sequence of IEEE operations as tc_core's (tc_cell.hpp), so the result
// is bit-identical to the one-thread-per-cell form.

// Four wavefronts per SIMD (at most 128 VGPRs; the compiler spills about 16 bytes per lane, outside
// the row loop): 79.9 us per subcycle on the 2048x2048 arcade against 84.3 us at 152 VGPRs and three
// wavefronts per SIMD.  Reading the four b_hat derivatives from grids written once per step instead
// of forming them here was slower (102 us): the kernel is bound by latency and issue, not by the
// VALU work those derivatives take.
template <int MODE, bool SAT, bool SEP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_tc_strip(TcStepArgs a, int R, int nstrips, int total) {
  using namespace tcs;
  const Mesh& M = a.mesh;
  const int t = strip::xcd_tile(blockIdx.x, total);
  const int tile = t / nstrips, sidx = t - tile * nstrips;
  const int lane = threadIdx.x;
  Lane L;
  {
    const int jc = tile * 62 + lane - 1;
    L.j = row_index(jc, M.ny, M.y.per, L.val);
    L.iny = L.val && L.j >= M.y.lo && L.j <= M.y.hi;
    L.out = lane >= 1 && lane <= 62 && jc < M.ny;
    bool vp;
    const int jp = row_index(L.j + 1, M.ny, M.y.per, vp);
    L.hc = M.y.h[jp];
    L.hn = M.y.h[L.j];
    L.fd = M.y.fden[jp];
    L.rfd = M.y.rfden[jp];
    L.dy = M.y.d[L.j];
    L.rdy = M.y.rd[L.j];
    L.hsq = M.y.hsq[L.j];
    L.rhsq = M.y.rhsq[L.j];
  }
  const int r0 = sidx * R, r1 = min(r0 + R, M.nx);
  Row cur = make_row<SAT>(a, L, r0, load_raw<SAT>(a, L, r0));
  XFaces lo;
  {
    const Row prev = make_row<SAT>(a, L, r0 - 1, load_raw<SAT>(a, L, r0 - 1));
    lo = xfaces<SAT>(M, prev, cur);
  }
  Raw ahead = load_raw<SAT>(a, L, r0 + 1);
  for (int r = r0; r < r1; ++r) {
    const Raw raw = ahead;
    if (r + 1 < r1) ahead = load_raw<SAT>(a, L, r + 2);
    const Row nxt = make_row<SAT>(a, L, r + 1, raw);
    const XFaces hi = xfaces<SAT>(M, cur, nxt);
    // cell values at row r along x
    const int i = cur.i;
    const double dx = M.x.d[i], rdx = M.x.rd[i];
    const bool in = cur.inx && L.iny;
    const double gx = qdiv(hi.T - lo.T, dx, rdx);
    // dtemp_dx is zero outside the interior; its y-derivative takes the neighbouring lanes' values
    const double dgx_y = yderiv(L, yface(L, in ? gx : 0.0));
    if (L.out && cur.val) {
      const size_t c = (size_t)i * M.ny + L.j;
      const double base = SEP ? a.te_base[c] : cur.te;
      if (!in) {
        // mask = 0: te + 0*dE; every derivative is zero there, so dE = 0
        a.te_out[c] = smax(base + 0.0, a.k.temin);
        if (MODE == TC_STORE_K) a.k_out[c] = 0.0;
      } else {
        const double s0 = qdiv(hi.T - 2.0 * cur.T + lo.T, M.x.hsq[i], M.x.rhsq[i]);
        const double dgy_x = qdiv(hi.gy - lo.gy, dx, rdx);
        const double dbxx = qdiv(hi.bx - lo.bx, dx, rdx);
        const double dbyx = qdiv(hi.by - lo.by, dx, rdx);
        double dE = tcc::tc_combine<double>(cur.T, gx, cur.gy, s0, cur.s1, dgx_y, dgy_x, cur.bx, cur.by, dbxx,
                                            cur.dbxy, dbyx, cur.dbyy, a.p.k0);
        if (SAT) {                                              // thermalconduction.cpp:171-174, 255-271
          const double dcx = qdiv(hi.coef - lo.coef, dx, rdx);
          const double add = -1.0 * (dcx * a.rx[c] + cur.dcy * a.ry[c]);
          dE = cur.coef * dE + add;
        }
        double v;
        if (MODE == TC_RK4_FINAL) {
          const double sum = a.k1[c] + 2.0 * a.k2[c] + 2.0 * a.k3[c] + dE;
          v = base + qdiv(1.0 * a.h * sum, 6.0, a.k.r6);
        } else {
          if (MODE == TC_STORE_K) a.k_out[c] = dE;
          v = base + 1.0 * (a.mult * a.h) * dE;
        }
        a.te_out[c] = smax(v, a.k.temin);
      }
    }
    cur = nxt;
    lo = hi;
  }
}

// RadiativeLosses::computeLosses at one interior cell (radiativelosses.cpp:149-197).
__device__ __forceinline__ double rl_loss(double t, double n, double te_live, double dt_live, const RlParams& p) {
  if (t < p.cut) return 0.0;
  // log10(t) <= 4.97 etc. as the host's log10 decides it (p.tb)
  double chi, pw;
  if (t <= p.tb[0]) { chi = 1.09e-31; pw = t * t; }
  else if (t <= p.tb[1]) { chi = 8.87e-17; pw = 1.0 / t; }
  else if (t <= p.tb[2]) { chi = 1.90e-22; pw = 1.0; }
  else if (t <= p.tb[3]) { chi = 3.53e-13; pw = pow_m1p5(t); }
  else if (t <= p.tb[4]) { chi = 3.46e-25; pw = pow_third(t); }
  else if (t <= p.tb[5]) { chi = 5.49e-16; pw = 1.0 / t; }
  else { chi = 1.96e-27; pw = sqrt(t); }
  double v = n * n * chi * pw;
  if (t < p.cut_plus_ramp) {
    const double ramp = (t - p.cut) / p.ramp;
    v *= ramp;
  }
  if (p.prevent) {
    if (0.1 * p.eps * (te_live / v) < p.pd_eps * (dt_live)) v = 0.1 * p.eps * (te_live) / (p.pd_eps * (dt_live));
  }
  return v;
}

__global__ void __launch_bounds__(256) k_rl_pre(MhdSet L, const double* dtg, RlParams p, MhdConst K, Mesh M,
                                                double* pmin, double* pmax) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double vmin = DBL_MAX, vmax = -DBL_MAX;
  if (i < M.nx && j < M.ny) {
    if (inside(M, i, j)) {
      const size_t c = (size_t)i * M.ny + j;
      const double te = L.te[c], n = L.n[c];
      const double l = rl_loss(temp_of(te, n, K), n, te, p.prevent ? dtg[c] : 0.0, p);
      vmax = l;
      const double r = fabs(te / l);
      vmin = (r < vmin) ? r : vmin;
    } else {
      vmax = 0.0;
    }
  }
  const double a = block_reduce(vmin, true);
  const double b = block_reduce(vmax, false);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    pmin[blk] = a;
    pmax[blk] = b;
  }
}

__global__ void __launch_bounds__(256) k_rl_iterate(MhdSet L, const double* dtg, RlParams p, MhdConst K, Mesh M,
                                                    int nsub, double h, int integ, double* te_out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double te_live = L.te[c];
  if (!inside(M, i, j)) { te_out[c] = te_live; return; }
  const double n = L.n[c];
  const double dtl = p.prevent ? dtg[c] : 0.0;
  auto loss = [&](double te) { return rl_loss(temp_of(te, n, K), n, te_live, dtl, p); };
  double te = te_live;
  const double hh = 1.0 * h, half = 1.0 * (0.5 * h);
  for (int s = 0; s < nsub; ++s) {
    if (integ == 0) {                                             // radiativelosses.cpp:92-96
      te = smax(te - hh * loss(te), K.temin);
    } else if (integ == 1) {                                      // :97-108
      const double hs = smax(te - half * loss(te), K.temin);
      te = smax(te - hh * loss(hs), K.temin);
    } else {                                                      // :109-130
      const double k1 = -loss(te);
      double im = smax(te + half * k1, K.temin);
      const double k2 = -loss(im);
      im = smax(te + half * k2, K.temin);
      const double k3 = -loss(im);
      im = smax(te + hh * k3, K.temin);
      const double k4 = -loss(im);
      te = smax(te + qdiv(hh * (k1 + 2.0 * k2 + 2.0 * k3 + k4), 6.0, K.r6), K.temin);
    }
  }
  te_out[c] = te;
}

__global__ void __launch_bounds__(256) k_propagate_live(MhdSet L, MhdConst K, size_t N, int write_pre) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  // enforceMinimums then recomputeDerivedVarsFromEvolvedVars (idealmhd.cpp:234-252)
  const double rho1 = smax(qdiv(L.rho[c], K.m, K.rm), K.nmin) * K.m;
  const double te1 = smax(L.te[c], K.temin);
  const double n = smax(qdiv(rho1, K.m, K.rm), K.nmin);
  if (write_pre) L.rho_pre[c] = rho1;
  L.n[c] = n;
  L.rho[c] = n * K.m;
  L.te[c] = smax(te1, K.temin);
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

int solar_blocks(const Mesh& m) {
  dim3 g = g2(m, dim3(64, 4));
  return (int)(g.x * g.y);
}

void launch_bhat(MhdSet L, MhdStatic st, Mesh mesh, double* bhx, double* bhy, hipStream_t s) {
  hipLaunchKernelGGL(k_bhat, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, st, mesh, bhx, bhy);
}

void launch_bhat_derivs(const double* bhx, const double* bhy, Mesh mesh, double* dbxx, double* dbxy, double* dbyx,
                        double* dbyy, hipStream_t s) {
  hipLaunchKernelGGL(k_bhat_derivs, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, bhx, bhy, mesh, dbxx, dbxy, dbyx, dbyy);
}

void launch_tc_pre(MhdSet L, const double* bhx, const double* bhy, TcParams p, MhdConst k, Mesh mesh, double* partial,
                   int nblocks, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_tc_pre, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, bhx, bhy, p, k, mesh, partial,
                     partial + nblocks);
  hipLaunchKernelGGL(k_finish_minmax, dim3(1), dim3(1024), 0, s, partial, partial + nblocks, nblocks, out);
}

void launch_tc_sat(const double* te_src, const double* n, const double* rho, const double* bhx, const double* bhy,
                   TcParams p, MhdConst k, Mesh mesh, double* rx, double* ry, double* coef, hipStream_t s) {
  hipLaunchKernelGGL(k_tc_sat, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, te_src, n, rho, bhx, bhy, p, k, mesh, rx, ry,
                     coef);
}

// Measured on one MI355X: on the 2048x2048 arcade strips of 8 rows (8704 wavefronts) take 80 us
// per subcycle, against 85 us with 4 rows, 92 us with 16 and 124 us with one thread per cell; on the
// 512x1024 arcade the strips give too few wavefronts (2176 with 4 rows) and one thread per cell is
// faster (20.3 us against 22.0 us).
int tc_strip_rows_default(int nx, int ny) {
  const int tiles = (ny + 61) / 62;
  for (int R : {8, 4})
    if (tiles * ((nx + R - 1) / R) >= 4096) return R;
  return 2;
}

int tc_kernel_default(int nx, int ny) {
  const int tiles = (ny + 61) / 62;
  return tiles * ((nx + 7) / 8) >= 4096 ? 0 : 1;
}

template <int MODE, bool SAT>
static void launch_tc_strip_m(const TcStepArgs& a, int R, hipStream_t s) {
  const int tiles = (a.mesh.ny + 61) / 62, nstrips = (a.mesh.nx + R - 1) / R, total = tiles * nstrips;
  if (a.te_base != a.te_src)
    hipLaunchKernelGGL((k_tc_strip<MODE, SAT, true>), dim3(total), dim3(64), 0, s, a, R, nstrips, total);
  else
    hipLaunchKernelGGL((k_tc_strip<MODE, SAT, false>), dim3(total), dim3(64), 0, s, a, R, nstrips, total);
}

void launch_tc_step(const TcStepArgs& a, hipStream_t s) {
  if (a.rows <= 0) {
    hipLaunchKernelGGL(k_tc_step, g2(a.mesh, dim3(64, 4)), dim3(64, 4), 0, s, a);
    return;
  }
  if (a.p.saturation) {
    if (a.mode == TC_STORE_K) launch_tc_strip_m<TC_STORE_K, true>(a, a.rows, s);
    else if (a.mode == TC_RK4_FINAL) launch_tc_strip_m<TC_RK4_FINAL, true>(a, a.rows, s);
    else launch_tc_strip_m<TC_UPDATE, true>(a, a.rows, s);
  } else {
    if (a.mode == TC_STORE_K) launch_tc_strip_m<TC_STORE_K, false>(a, a.rows, s);
    else if (a.mode == TC_RK4_FINAL) launch_tc_strip_m<TC_RK4_FINAL, false>(a, a.rows, s);
    else launch_tc_strip_m<TC_UPDATE, false>(a, a.rows, s);
  }
}

void launch_rl_pre(MhdSet L, const double* dt_grid, RlParams p, MhdConst k, Mesh mesh, double* partial, int nblocks,
                   double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_rl_pre, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, dt_grid, p, k, mesh, partial,
                     partial + nblocks);
  hipLaunchKernelGGL(k_finish_minmax, dim3(1), dim3(1024), 0, s, partial, partial + nblocks, nblocks, out);
}

void launch_rl_iterate(MhdSet L, const double* dt_grid, RlParams p, MhdConst k, Mesh mesh, int nsub, double h,
                       int integrator, double* te_out, hipStream_t s) {
  hipLaunchKernelGGL(k_rl_iterate, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, L, dt_grid, p, k, mesh, nsub, h,
                     integrator, te_out);
}

void launch_propagate_live(MhdSet L, MhdConst k, int nx, int ny, int write_pre, hipStream_t s) {
  const size_t N = (size_t)nx * ny;
  hipLaunchKernelGGL(k_propagate_live, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, L, k, N, write_pre);
}

}  // namespace spruce
