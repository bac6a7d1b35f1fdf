// IdealMHD stage kernel, wavefront-strip form (the production path of launch_mhd_stage).
//
// Each 64-lane wavefront owns a strip of 64 columns (y, contiguous in memory) and marches down R
// rows (x).  Lanes 2..61 produce outputs; lanes 0, 1, 62, 63 only supply the y-halo.  Along x the
// four rows a face needs (c-2 .. c+1) stay in a rotating register window, so every x-face is
// computed once and reused by the two cells that share it; along y the neighbouring columns are
// read from the neighbouring lanes with DPP wave shifts, so every y-face is computed once too.
// Every cell-centred derived value (v = mom/rho, press) is computed once per row.  The arithmetic
// of each value is the reference's (derivs.cpp, idealmhd.cpp), operation for operation.
#include <cfloat>
#include <type_traits>

#include "arith.hpp"
#include "mhd.hpp"
#include "strip.hpp"

namespace spruce {
namespace {

constexpr int NQ = 11;  // transported quantities, in idealmhd.cpp:52-86 order
enum { Q_RHO = 0, Q_MX, Q_MY, Q_MZ, Q_TE, Q_BIX, Q_BEX, Q_BIY, Q_BEY, Q_BIZ, Q_BEZ };
using namespace strip;

// qdiv for a data-dependent divisor: the reciprocal is exact-rounded only if it is a normal number.
__device__ __forceinline__ double qdiv_d(double a, double b, double rb) {
  const double ab = __builtin_fabs(b);
  if (__builtin_expect(ab > 0x1p-900 && ab < 0x1p+900, 1)) return qdiv(a, b, rb);
  return a / b;
}

// Face velocity of one face: interpolated v and press, and the upwind side (derivs.cpp:10-73, 155).
template <int AX>
__device__ __forceinline__ FaceDir face_velocity(const double* vlo, const double* vup, double plo, double pup,
                                                 const FaceGeom& g, double* iv, double& ip) {
#pragma unroll
  for (int k = 0; k < 3; ++k) iv[k] = interp(g, vlo[k], vup[k]);
  ip = interp(g, plo, pup);
  return face_dir(iv[AX], g);
}

__device__ __forceinline__ void load_row(const StageArgs& a, int i, bool ivalid, int j, bool jvalid, double* q) {
  if (ivalid && jvalid) {
    const size_t c = (size_t)i * a.mesh.ny + j;
    q[Q_RHO] = a.S.rho[c];
    q[Q_MX] = a.S.mx[c];
    q[Q_MY] = a.S.my[c];
    q[Q_MZ] = a.S.mz[c];
    q[Q_TE] = a.S.te[c];
    q[Q_BIX] = a.S.bx[c];
    q[Q_BEX] = a.st.bex[c];
    q[Q_BIY] = a.S.by[c];
    q[Q_BEY] = a.st.bey[c];
    q[Q_BIZ] = a.S.bz[c];
    q[Q_BEZ] = a.st.bez[c];
  } else {
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    q[Q_RHO] = 1.0;
  }
}

// v = mom/rho (idealmhd.cpp:248-250) and press = (gamma-1) te (idealmhd.cpp:253)
__device__ __forceinline__ void derive(const double* q, const MhdConst& K, double* v, double& p) {
  const double rho = q[Q_RHO];
  const double rr = 1.0 / rho;
  v[0] = qdiv_d(q[Q_MX], rho, rr);
  v[1] = qdiv_d(q[Q_MY], rho, rr);
  v[2] = qdiv_d(q[Q_MZ], rho, rr);
  p = q[Q_TE] * K.gm1;
}

// Strip of a workgroup: launch index L mapped XCD-aware (xcd_tile) onto the strips in column-major
// order, so strips that run together on one XCD are x-neighbours and share their halo rows in L2.
__device__ __forceinline__ void strip_of(int L, int nstrips, int total, int& col, int& strip) {
  const int t = xcd_tile(L, total);
  col = t / nstrips;
  strip = t - col * nstrips;
}

template <int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_mhd_stage_strip(StageArgs a, int R, int nstrips) {
  const Mesh& M = a.mesh;
  const MhdConst& K = a.k;
  const int nx = M.nx, ny = M.ny;
  const int lane = threadIdx.x;
  int col, strip;
  strip_of(blockIdx.x, nstrips, gridDim.x, col, strip);
  const int jraw = col * OUT_COLS + lane - 2;
  int j;
  bool jvalid;
  if (M.y.per) { j = wrapn(jraw, ny); jvalid = true; }
  else { jvalid = jraw >= 0 && jraw < ny; j = clampn(jraw, ny); }
  const bool out_lane = lane >= 2 && lane < 62 && jraw >= 0 && jraw < ny;
  const bool yint = jvalid && j >= M.y.lo && j <= M.y.hi;
  const double stepv = a.step_mult * (*a.step);
  const size_t N = (size_t)nx * ny;
  const int r0 = strip * R;

  auto row_of = [&](int r, bool& ok) -> int {
    if (M.x.per) { ok = true; return wrapn(r, nx); }
    ok = r >= 0 && r < nx;
    return clampn(r, nx);
  };

  double q[4][NQ];      // rotating window of rows r-1 .. r+2
  double v[2][3], p[2];  // v and press of rows r and r+1
  double Px[NQ];         // flux of row r's lower x-face (the upper face of row r-1)
  double xiv[2][3], xip[2], xib0[2], xib1[2];   // interpolated values of the lower/upper x-face

  {  // prologue: rows r0-2 .. r0+1 and the lower face of row r0
    bool ok;
    int i;
    i = row_of(r0 - 2, ok); load_row(a, i, ok, j, jvalid, q[3]);
    i = row_of(r0 - 1, ok); load_row(a, i, ok, j, jvalid, q[0]);
    i = row_of(r0, ok);     load_row(a, i, ok, j, jvalid, q[1]);
    i = row_of(r0 + 1, ok); load_row(a, i, ok, j, jvalid, q[2]);
    double vt[3], pt;
    derive(q[0], K, vt, pt);
    derive(q[1], K, v[0], p[0]);
    const FaceGeom gx = face_geom(M.x, r0);
    const FaceDir fd = face_velocity<0>(vt, v[0], pt, p[0], gx, xiv[0], xip[0]);
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      double d2;
      Px[k] = face_flux(q[3][k], q[0][k], q[1][k], q[2][k], gx, fd, d2);
      if (k == Q_BIY) xib0[0] = d2;
      if (k == Q_BIZ) xib1[0] = d2;
    }
  }

  auto step = [&](auto BC, int r) {
    constexpr int B = decltype(BC)::value;
    constexpr int SM = B % 4, S0 = (B + 1) % 4, S1 = (B + 2) % 4, S2 = (B + 3) % 4;
    constexpr int D0 = B % 2, D1 = (B + 1) % 2;
    bool ok;
    // row r+2 is loaded first and arrives while row r's y-faces are computed
    const int i2 = row_of(r + 2, ok);
    load_row(a, i2, ok, j, jvalid, q[S2]);
    if (r >= nx) return;
    const int i = row_of(r, ok);
    // lower y-face of this lane's column in row r (cells j-2, j-1, j, j+1 from neighbouring lanes),
    // one transported quantity at a time: the upper y-face of the column is the next lane's lower one
    double fyiv[3], fyip, fyib0 = 0.0, fyib1 = 0.0;
    double vm1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) vm1[k] = shr1(v[D0][k]);
    // geometry of this lane's lower y-face (between j-1 and j), re-read per row (cached) rather
    // than held in registers for the whole strip
    const FaceGeom gy = face_geom(M.y, j);
    const double dy = M.y.d[j], rdy = M.y.rd[j];
    const FaceDir fdy = face_velocity<1>(vm1, v[D0], shr1(p[D0]), p[D0], gy, fyiv, fyip);
    double tdiv[NQ];       // y part first; the x part is added when the upper x-face is known
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const double qc_k = q[S0][k];
      const double qm1 = shr1(qc_k);
      const double qm2 = shr1(qm1);
      const double qp1 = shl1(qc_k);
      double d2;
      const double Pl = face_flux(qm2, qm1, qc_k, qp1, gy, fdy, d2);
      // y-faces feed derivative1D(bi_x, 1) and derivative1D(bi_z, 1)
      if (k == Q_BIX) fyib0 = d2;
      if (k == Q_BIZ) fyib1 = d2;
      tdiv[k] = qdiv(shl1(Pl) - Pl, dy, rdy);
    }
    const double ivu0 = shl1(fyiv[0]), ivu1 = shl1(fyiv[1]), ivu2 = shl1(fyiv[2]);
    const double ipu = shl1(fyip), ibu0 = shl1(fyib0), ibu1 = shl1(fyib1);
    // upper x-face of row r from rows r-1 .. r+2; divergence term (up - lo)/dx + y term
    derive(q[S1], K, v[D1], p[D1]);
    const double dx = M.x.d[i], rdx = M.x.rd[i];
    {
      const FaceGeom gx = face_geom(M.x, r + 1);
      const FaceDir fd = face_velocity<0>(v[D0], v[D1], p[D0], p[D1], gx, xiv[D1], xip[D1]);
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        double d2;
        const double Pn = face_flux(q[SM][k], q[S0][k], q[S1][k], q[S2][k], gx, fd, d2);
        // x-faces feed derivative1D(bi_y, 0) and derivative1D(bi_z, 0)
        if (k == Q_BIY) xib0[D1] = d2;
        if (k == Q_BIZ) xib1[D1] = d2;
        const double tx = qdiv(Pn - Px[k], dx, rdx);
        tdiv[k] = tx + tdiv[k];
        Px[k] = Pn;
      }
    }
    if (!out_lane) return;
    const size_t c = (size_t)i * ny + j;
    double k8[8];
    if (i >= M.x.lo && i <= M.x.hi && yint) {
      auto ddx = [&](double up, double lo) { return qdiv(up - lo, dx, rdx); };
      auto ddy = [&](double up, double lo) { return qdiv(up - lo, dy, rdy); };
      const double dpx = ddx(xip[D1], xip[D0]), dpy = ddy(ipu, fyip);
      const double dvx_x = ddx(xiv[D1][0], xiv[D0][0]), dvx_y = ddy(ivu0, fyiv[0]);
      const double dvy_x = ddx(xiv[D1][1], xiv[D0][1]), dvy_y = ddy(ivu1, fyiv[1]);
      const double dvz_x = ddx(xiv[D1][2], xiv[D0][2]), dvz_y = ddy(ivu2, fyiv[2]);
      const double dbiy_x = ddx(xib0[D1], xib0[D0]), dbix_y = ddy(ibu0, fyib0);
      const double dbiz_x = ddx(xib1[D1], xib1[D0]), dbiz_y = ddy(ibu1, fyib1);
      const double* qc = q[S0];
      const double rho = qc[Q_RHO];
      const double press = p[D0];
      const double vx = v[D0][0], vy = v[D0][1], vz = v[D0][2];
      const double bix = qc[Q_BIX], biy = qc[Q_BIY], biz = qc[Q_BIZ];
      const double bex = qc[Q_BEX], bey = qc[Q_BEY], bez = qc[Q_BEZ];
      const double curl_db = qdiv(dbiy_x - dbix_y, 4.0 * PI, K.r4pi);
      const double ext0 = (-curl_db) * bey, ext1 = curl_db * bex;
      const double in0 = (-curl_db) * biy, in1 = curl_db * bix;
      const double cz0 = dbiz_y, cz1 = -dbiz_x;
      const double bzi = qdiv(biz, 4.0 * PI, K.r4pi), bze = qdiv(bez, 4.0 * PI, K.r4pi);
      const double inz0 = bzi * cz1, inz1 = (-bzi) * cz0;
      const double extz0 = bze * cz1, extz1 = (-bze) * cz0;
      const double fz_ext = qdiv(cz0 * bey - cz1 * bex, 4.0 * PI, K.r4pi);
      const double fz_in = qdiv(cz0 * biy - cz1 * bix, 4.0 * PI, K.r4pi);
      k8[0] = -tdiv[Q_RHO] + 0.0;
      k8[1] = (-tdiv[Q_MX] - dpx + rho * a.st.gx[c] + ext0 + in0 + inz0 + extz0) + (rho * 0.0 + vx * 0.0);
      k8[2] = (-tdiv[Q_MY] - dpy + rho * a.st.gy[c] + ext1 + in1 + inz1 + extz1) + (rho * 0.0 + vy * 0.0);
      k8[3] = (-tdiv[Q_MZ] + fz_ext + fz_in) + (rho * 0.0 + vz * 0.0);
      k8[4] = (-tdiv[Q_TE] - press * (dvx_x + dvy_y)) + 0.0;
      k8[5] = (-tdiv[Q_BIX] - tdiv[Q_BEX] + (bix + bex) * dvx_x + (biy + bey) * dvx_y) + 0.0;
      k8[6] = (-tdiv[Q_BIY] - tdiv[Q_BEY] + (bix + bex) * dvy_x + (biy + bey) * dvy_y) + 0.0;
      k8[7] = (-tdiv[Q_BIZ] - tdiv[Q_BEZ] + (bix + bex) * dvz_x + (biy + bey) * dvz_y) + 0.0;
    } else if (a.moc) {  // open_moc ghost cells: masked RHS (0) + characteristic terms (idealmhd.cpp:99-103)
#pragma unroll
      for (int w = 0; w < 8; ++w) k8[w] = 0.0 + a.moc[w * N + c];
    } else {
#pragma unroll
      for (int w = 0; w < 8; ++w) k8[w] = 0.0;
    }
    for (int e = 0; e < a.n_extra; ++e) {
#pragma unroll
      for (int w = 0; w < 8; ++w)
        if (w == a.extra_var[e]) k8[w] = k8[w] + a.extra[e][c];
    }
    if (MODE == STAGE_STORE_K) {
#pragma unroll
      for (int w = 0; w < 8; ++w) a.kbuf[w * N + c] = k8[w];
    } else if (MODE == STAGE_RK4_FINAL) {
#pragma unroll
      for (int w = 0; w < 8; ++w) {
        const double k1 = a.kbuf[w * N + c], k2 = a.k2buf[w * N + c], k3 = a.k3buf[w * N + c];
        k8[w] = qdiv(k1 + k8[w], 6.0, K.r6) + qdiv(k2 + k3, 3.0, K.r3);
      }
      if (a.kf_out) {
#pragma unroll
        for (int w = 0; w < 8; ++w) a.kf_out[w * N + c] = k8[w];
      }
    }
    double b[8];
    if (a.B.rho == a.S.rho) {
      b[0] = q[S0][Q_RHO]; b[1] = q[S0][Q_MX]; b[2] = q[S0][Q_MY]; b[3] = q[S0][Q_MZ];
      b[4] = q[S0][Q_TE]; b[5] = q[S0][Q_BIX]; b[6] = q[S0][Q_BIY]; b[7] = q[S0][Q_BIZ];
    } else {
      b[0] = a.B.rho[c]; b[1] = a.B.mx[c]; b[2] = a.B.my[c]; b[3] = a.B.mz[c];
      b[4] = a.B.te[c]; b[5] = a.B.bx[c]; b[6] = a.B.by[c]; b[7] = a.B.bz[c];
    }
    const double rho_new = b[0] + k8[0] * stepv;
    const double rho1 = smax(qdiv(rho_new, K.m, K.rm), K.nmin) * K.m;
    const double te1 = smax(b[4] + k8[4] * stepv, K.temin);
    const double n = smax(qdiv(rho1, K.m, K.rm), K.nmin);
    a.T.rho[c] = n * K.m;
    if (a.write_n) a.T.n[c] = n;
    if (a.write_pre) a.T.rho_pre[c] = rho1;
    a.T.mx[c] = b[1] + k8[1] * stepv;
    a.T.my[c] = b[2] + k8[2] * stepv;
    a.T.mz[c] = b[3] + k8[3] * stepv;
    a.T.te[c] = smax(te1, K.temin);
    a.T.bx[c] = b[5] + k8[5] * stepv;
    a.T.by[c] = b[6] + k8[6] * stepv;
    a.T.bz[c] = b[7] + k8[7] * stepv;
  };

  for (int r = r0; r < r0 + R; r += 4) {
    step(std::integral_constant<int, 0>{}, r);
    step(std::integral_constant<int, 1>{}, r + 1);
    step(std::integral_constant<int, 2>{}, r + 2);
    step(std::integral_constant<int, 3>{}, r + 3);
  }
}

}  // namespace

int mhd_strip_rows_default(int nx, int ny) {
  // the tallest strip that still leaves about 32 wavefronts per CU (256 CUs); shorter strips re-read
  // more halo rows, but on 2048x2048 and 512x1024 the extra waves hide more latency than that costs
  const long tiles = (ny + OUT_COLS - 1) / OUT_COLS;
  for (int R : {32, 16, 8}) {
    const long waves = tiles * ((nx + R - 1) / R);
    if (waves >= 8192) return R;
  }
  return 4;
}

int mhd_stage_kernel_default(int nx, int ny) {
  // with 4-row strips a 256x256 grid gives 320 wavefronts, too few to fill 256 CUs (strips 49 us,
  // one thread per cell 30 us, 8x8 tiles 21 us per stage); from about four wavefronts per CU on
  // (512x512: strips 127 us per step, tiles 139 us), the strips are faster
  const long waves = (long)((ny + OUT_COLS - 1) / OUT_COLS) * ((nx + 3) / 4);
  return waves >= 1024 ? 0 : 2;
}

void launch_mhd_stage_strip(const StageArgs& a, int R, hipStream_t s) {
  const int cols = (a.mesh.ny + OUT_COLS - 1) / OUT_COLS, nstrips = (a.mesh.nx + R - 1) / R;
  const dim3 grid(cols * nstrips);
  if (a.mode == STAGE_STORE_K)
    hipLaunchKernelGGL(k_mhd_stage_strip<STAGE_STORE_K>, grid, dim3(64), 0, s, a, R, nstrips);
  else if (a.mode == STAGE_RK4_FINAL)
    hipLaunchKernelGGL(k_mhd_stage_strip<STAGE_RK4_FINAL>, grid, dim3(64), 0, s, a, R, nstrips);
  else
    hipLaunchKernelGGL(k_mhd_stage_strip<STAGE_PLAIN>, grid, dim3(64), 0, s, a, R, nstrips);
}

}  // namespace spruce
