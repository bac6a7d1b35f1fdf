// IdealMHD time-step kernels for gfx950.
//
// Arithmetic follows the reference operation by operation (source/equationsets/idealmhd.cpp,
// source/mhd/derivs.cpp, source/mhd/evolution.cpp); every quantity is rounded exactly where the
// reference's whole-grid operations round it, so results are bit-identical to the CPU build.
// Memory layout: one contiguous FP64 array per variable, x-major with y contiguous, so a wavefront
// of 64 lanes reads 512 contiguous bytes per variable.
#include <cfloat>

#include "arith.hpp"
#include "mhd.hpp"
#include "mhd_cell.hpp"
#include "strip.hpp"

namespace spruce {

namespace {

// The right-hand side of a ghost cell: the masked derivative (0) plus, with open_moc sides, the
// characteristic terms (idealmhd.cpp:99-103).
__device__ __forceinline__ void ghost_rhs(const StageArgs& a, size_t N, size_t c, double k[8]) {
  if (a.moc) {
#pragma unroll
    for (int v = 0; v < 8; ++v) k[v] = 0.0 + a.moc[v * N + c];
  } else {
#pragma unroll
    for (int v = 0; v < 8; ++v) k[v] = 0.0;
  }
}

// The values a stage writes that the step size reads.
struct StageOut {
  double rho, mx, my, te, bx, by, bz;
};

// The rest of a stage at cell c once the equation set's k is known: module terms, the RK4
// combination, the update, enforceMinimums and the derived density.
__device__ __forceinline__ StageOut stage_finish(const StageArgs& a, size_t N, size_t c, double k[8]) {
  for (int e = 0; e < a.n_extra; ++e) {
#pragma unroll
    for (int v = 0; v < 8; ++v)
      if (v == a.extra_var[e]) k[v] = k[v] + a.extra[e][c];
  }
  if (a.mode == STAGE_STORE_K) {
#pragma unroll
    for (int v = 0; v < 8; ++v) a.kbuf[v * N + c] = k[v];
  } else if (a.mode == STAGE_RK4_FINAL) {
    // (k1 + k4)/6 + (k2 + k3)/3 (evolution.cpp:119-121)
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const double k1 = a.kbuf[v * N + c], k2 = a.k2buf[v * N + c], k3 = a.k3buf[v * N + c];
      k[v] = qdiv(k1 + k[v], 6.0, a.k.r6) + qdiv(k2 + k3, 3.0, a.k.r3);
    }
    if (a.kf_out) {
#pragma unroll
      for (int v = 0; v < 8; ++v) a.kf_out[v * N + c] = k[v];
    }
  }
  const double step = a.step_mult * (*a.step);
  const MhdConst& K = a.k;
  const double rho_new = a.B.rho[c] + k[0] * step;
  const double mx = a.B.mx[c] + k[1] * step;
  const double my = a.B.my[c] + k[2] * step;
  const double mz = a.B.mz[c] + k[3] * step;
  const double te_new = a.B.te[c] + k[4] * step;
  const double bx = a.B.bx[c] + k[5] * step;
  const double by = a.B.by[c] + k[6] * step;
  const double bz = a.B.bz[c] + k[7] * step;
  // enforceMinimums (idealmhd.cpp:237-238)
  const double rho1 = smax(qdiv(rho_new, K.m, K.rm), K.nmin) * K.m;
  const double te1 = smax(te_new, K.temin);
  // recomputeDerivedVarsFromEvolvedVars (idealmhd.cpp:246-252)
  const double n = smax(qdiv(rho1, K.m, K.rm), K.nmin);
  const StageOut o{n * K.m, mx, my, smax(te1, K.temin), bx, by, bz};
  a.T.rho[c] = o.rho;
  if (a.write_n) a.T.n[c] = n;
  if (a.write_pre) a.T.rho_pre[c] = rho1;
  a.T.mx[c] = mx;
  a.T.my[c] = my;
  a.T.mz[c] = mz;
  a.T.te[c] = o.te;
  a.T.bx[c] = bx;
  a.T.by[c] = by;
  a.T.bz[c] = bz;
  return o;
}

// Stage: T = propagate-minimums-derived(B + step * k(S)) (equationset.cpp:222-230, 212-220;
// idealmhd.cpp:234-263).  Ghost cells have k = +0.
__global__ void __launch_bounds__(256) k_mhd_stage(StageArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const Mesh& M = a.mesh;
  if (i >= M.nx || j >= M.ny) return;
  const size_t N = (size_t)M.nx * M.ny;
  const size_t c = (size_t)i * M.ny + j;
  double k[8];
  if (interior(M, i, j)) {
    mhd_rhs_cell(a.S, a.st, a.k, M, i, j, k);
  } else {
    ghost_rhs(a, N, c, k);
  }
  (void)stage_finish(a, N, c, k);
}

// populateVariablesFromState up to the ghost fill (equationset.cpp:96-104; idealmhd.cpp:226-252).
__global__ void __launch_bounds__(256) k_mhd_populate(MhdSet L, const double* temp, MhdConst K, int nx, int ny) {
  const size_t N = (size_t)nx * ny;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double rho = L.rho[c];
  const double n0 = smax(qdiv(rho, K.m, K.rm), K.nmin);
  const double press = n0 * 2.0 * K_B * smax(temp[c], K.tmin);
  const double te = qdiv(press, K.gm1, K.rgm1);
  const double rho1 = smax(qdiv(rho, K.m, K.rm), K.nmin) * K.m;
  const double te1 = smax(te, K.temin);
  const double n = smax(qdiv(rho1, K.m, K.rm), K.nmin);
  L.rho_pre[c] = rho1;
  L.n[c] = n;
  L.rho[c] = n * K.m;
  L.te[c] = smax(te1, K.temin);
}

// Ghost-cell fill of one side for one boundary column/row (evolution.cpp:126-334).
// side: 0 x lower, 1 x upper, 2 y lower, 3 y upper; t indexes along the boundary.
__device__ void ghost_side(MhdSet G, const MhdConst& K, const Mesh& M, int side, int kind, int t,
                           double s1, double s2, int mode) {
  const bool xside = side < 2, upper = side & 1;
  const Axis& ax = xside ? M.x : M.y;
  const Axis& ot = xside ? M.y : M.x;
  const int n = ax.n;
  if (kind == 2 /*fixed*/) {
    if (t < 0 || t >= ot.n) return;
  } else {
    if (t < ot.lo || t > ot.hi) return;
  }
  const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
  auto at = [&](int e) -> size_t { return xside ? (size_t)e * M.ny + t : (size_t)t * M.ny + e; };
  const size_t c1 = at(e1), c2 = at(e2), c3 = at(e3);
  if (kind == 1 /*open*/) {                                           // evolution.cpp:158-224
    const double rho3 = (mode == GHOST_NORMAL) ? G.rho_pre[c3] : G.rho[c3];
    const double te3 = G.te[c3];
    const double rg1 = s1 * rho3, rg2 = s2 * rho3;
    const double tg1 = s1 * te3, tg2 = s2 * te3;
    double cs = 0.0;
    const double csn = sqrt(K.gamma * (te3 * K.gm1) / rho3);
    if (csn > cs) cs = csn;
    const double vx = G.mx[c3] / rho3, vy = G.my[c3] / rho3;
    double boost = K.obs * cs;
    if (!upper) boost *= -1.0;
    const double dist = 0.5 * (ax.d[e2] + ax.d[e3]);
    const double vn = xside ? vx : vy;
    const double bv = upper ? smax(0.0, vn + boost) : smin(0.0, vn + boost);
    const double gv = (dist * bv - 0.5 * ax.d[e2] * vn) / (0.5 * ax.d[e3]);
    const double mxg1 = rg1 * (xside ? gv : vx), mxg2 = rg2 * (xside ? gv : vx);
    const double myg1 = rg1 * (xside ? vy : gv), myg2 = rg2 * (xside ? vy : gv);
    if (mode == GHOST_NORMAL) {
      const double n1 = smax(qdiv(rg1, K.m, K.rm), K.nmin), n2 = smax(qdiv(rg2, K.m, K.rm), K.nmin);
      G.n[c1] = n1; G.n[c2] = n2;
      G.rho[c1] = n1 * K.m; G.rho[c2] = n2 * K.m;
      G.te[c1] = smax(tg1, K.temin); G.te[c2] = smax(tg2, K.temin);
    } else {
      G.rho[c1] = rg1; G.rho[c2] = rg2;
      G.te[c1] = tg1; G.te[c2] = tg2;
    }
    G.mx[c1] = mxg1; G.mx[c2] = mxg2;
    G.my[c1] = myg1; G.my[c2] = myg2;
  } else if (kind == 3 /*reflect*/) {                                 // evolution.cpp:231-266
    const double te3 = G.te[c3], rho3 = G.rho[c3];
    G.te[c1] = te3; G.te[c2] = te3;
    G.rho[c1] = rho3; G.rho[c2] = rho3;
    if (mode == GHOST_NORMAL) { const double n3 = G.n[c3]; G.n[c1] = n3; G.n[c2] = n3; }
    G.mx[c1] = 0.0; G.mx[c2] = 0.0; G.mx[c3] = 0.0;
    G.my[c1] = 0.0; G.my[c2] = 0.0; G.my[c3] = 0.0;
    G.mz[c1] = 0.0; G.mz[c2] = 0.0; G.mz[c3] = 0.0;
  } else if (kind == 2 /*fixed*/) {                                   // evolution.cpp:272-282
    G.mx[c1] = 0.0; G.mx[c2] = 0.0; G.mx[c3] = 0.0;
    G.my[c1] = 0.0; G.my[c2] = 0.0; G.my[c3] = 0.0;
    G.mz[c1] = 0.0; G.mz[c2] = 0.0; G.mz[c3] = 0.0;
  } else if (kind == 5 /*open_ucnp*/) {                               // evolution.cpp:290-334
    const int ein = upper ? ax.hi : ax.lo;
    const size_t ci = at(ein);
    for (int g = 0; g < N_GHOST; ++g) {
      const size_t cb = at(upper ? ax.hi + 1 + g : g);
      // the reference copies rho (densities(), idealmhd.hpp:37); a propagate's derived step then
      // gives the ghost the interior's n, which the normal mode copies directly
      G.rho[cb] = G.rho[ci]; G.te[cb] = G.te[ci];
      if (mode == GHOST_NORMAL) G.n[cb] = G.n[ci];
      G.bx[cb] = G.bx[ci]; G.by[cb] = G.by[ci]; G.bz[cb] = G.bz[ci];
      G.mx[cb] = G.mx[ci]; G.my[cb] = G.my[ci]; G.mz[cb] = G.mz[ci];
    }
  }
}

struct GhostArgs {
  MhdSet G;
  MhdConst k;
  Mesh mesh;
  int kind[4];
  double s1[4], s2[4];
  int mode;
};

// With lines > 0 each thread fills one line of one of the two sides (the sides touch disjoint cells
// and read disjoint cells once the axis has at least 6 cells); otherwise one thread fills a line of
// both sides in order.
__global__ void __launch_bounds__(256) k_mhd_ghost(GhostArgs a, int xsides, int lines) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s0 = xsides ? 0 : 2;
  if (lines > 0) {
    const int s = s0 + (t >= lines ? 1 : 0), line = t >= lines ? t - lines : t;
    if (t < 2 * lines && a.kind[s] != 0 && a.kind[s] != 4)
      ghost_side(a.G, a.k, a.mesh, s, a.kind[s], line, a.s1[s], a.s2[s], a.mode);
    return;
  }
  for (int s = s0; s < s0 + 2; ++s)
    if (a.kind[s] != 0 && a.kind[s] != 4) ghost_side(a.G, a.k, a.mesh, s, a.kind[s], t, a.s1[s], a.s2[s], a.mode);
}

// IdealMHD::recomputeDT at one cell (idealmhd.cpp:279-304).
// recomputeDT at one cell from its density, thermal energy, momenta and total field
__device__ __forceinline__ double mhd_dt_vals(double rho, double te, double mxc, double myc, double bx, double by,
                                              double bz, const MhdConst& K, const Mesh& M, int i, int j) {
  const double press = te * K.gm1;
  const double bmag = sqrt(bx * bx + by * by + bz * bz);
  const double vx = mxc / rho, vy = myc / rho;
  const double cs = sqrt(press * K.gamma / rho);
  const double cs2 = cs * cs;
  const double va = bmag / sqrt(rho * (4.0 * PI));
  const double va2 = va * va;
  const double sum = cs2 + va2;
  const double delta = sqrt(1.0 - cs2 * 4.0 * va2 / (sum * sum));
  const double vfast = sqrt(sum * 0.5 * (1.0 + delta));
  const double vslow = sqrt(sum * 0.5 * (1.0 - delta));
  const double vmx = sqrt(vx * vx), vmy = sqrt(vy * vy);
  const double cmax = smax(smax(smax(cs, va), vfast), vslow);
  return 1.0 / (qdiv(vmx + cmax, M.x.d[i], M.x.rd[i]) + qdiv(vmy + cmax, M.y.d[j], M.y.rd[j]));
}

__device__ __forceinline__ double mhd_dt_cell(const MhdSet& L, const MhdStatic& st, const MhdConst& K,
                                              const Mesh& M, int i, int j) {
  const size_t c = (size_t)i * M.ny + j;
  return mhd_dt_vals(L.rho[c], L.te[c], L.mx[c], L.my[c], st.bex[c] + L.bx[c], st.bey[c] + L.by[c],
                     st.bez[c] + L.bz[c], K, M, i, j);
}

__device__ __forceinline__ double block_min(double v) {
  // wavefront reduction with std::min semantics (grid.cpp:74-79); inputs are never NaN here
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  __shared__ double part[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  const int w = lane >> 6;
  if ((lane & 63) == 0) part[w] = v;
  __syncthreads();
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
  if (lane == 0) {
    double m = part[0];
    for (int k = 1; k < nw; ++k) m = (part[k] < m) ? part[k] : m;
    v = m;
  }
  return v;
}

__global__ void __launch_bounds__(256) k_mhd_dt(MhdSet L, MhdStatic st, MhdConst K, Mesh M, double* dt_grid,
                                                double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  double v = DBL_MAX;
  if (i < M.nx && j < M.ny) {
    const double dt = mhd_dt_cell(L, st, K, M, i, j);
    if (dt_grid) dt_grid[(size_t)i * M.ny + j] = dt;
    if (i >= M.x.dlo && i <= M.x.dhi && j >= M.y.dlo && j <= M.y.dhi) v = (dt < v) ? dt : v;
  }
  v = block_min(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = v;
}

// Tile form of the stage.  A block owns cells i0..i0+TX-1 x j0..j0+TY-1 (tiles row-major, handed
// to the XCDs in contiguous bands).  Face f of the x
// list lies between rows i0+f/TY-1 and i0+f/TY (column j0+f%TY); face g of the y list between
// columns j0+g%(TY+1)-1 and j0+g%(TY+1) (row i0+g/(TY+1)).  A face is computed when a cell on
// either side of it is interior, exactly as the one-thread-per-cell form computes it for that
// cell, so every value is bit-identical to that form; the faces shared by two cells are computed
// once instead of twice.
constexpr int TILE_X = 8, TILE_Y = 8, TILE_THREADS = 192;

__global__ void __launch_bounds__(TILE_THREADS) k_mhd_stage_tile(StageArgs a) {
  constexpr int NFX = (TILE_X + 1) * TILE_Y, NFY = TILE_X * (TILE_Y + 1);
  static_assert(NFX + NFY <= 2 * TILE_THREADS && TILE_X * TILE_Y == 64, "tile shape");
  __shared__ Face fx[NFX];
  __shared__ Face fy[NFY];
  const Mesh& M = a.mesh;
  const int ntj = (M.ny + TILE_Y - 1) / TILE_Y, nti = (M.nx + TILE_X - 1) / TILE_X;
  const int tile = strip::xcd_tile(blockIdx.x, ntj * nti);
  const int i0 = (tile / ntj) * TILE_X, j0 = (tile % ntj) * TILE_Y;
  const int t = threadIdx.x;
  for (int f = t; f < NFX + NFY; f += TILE_THREADS) {
    if (f < NFX) {
      const int i = i0 + f / TILE_Y, j = j0 + f % TILE_Y;
      if (j < M.ny && i <= M.nx) {
        const bool lower_of = i < M.nx && interior(M, i, j);   // the lower x-face of cell i
        const bool upper_of = i >= 1 && interior(M, i - 1, j);  // the upper x-face of cell i-1
        if (lower_of || upper_of) compute_face<0>(a.S, a.st, a.k, M, wrapi(i, M.x.n, M.x.per), j, fx[f]);
      }
    } else {
      const int g = f - NFX;
      const int i = i0 + g / (TILE_Y + 1), j = j0 + g % (TILE_Y + 1);
      if (i < M.nx && j <= M.ny) {
        const bool lower_of = j < M.ny && interior(M, i, j);
        const bool upper_of = j >= 1 && interior(M, i, j - 1);
        if (lower_of || upper_of) compute_face<1>(a.S, a.st, a.k, M, i, wrapi(j, M.y.n, M.y.per), fy[g]);
      }
    }
  }
  __syncthreads();
  if (t >= TILE_X * TILE_Y) return;   // one wavefront combines
  const int ti = t / TILE_Y, tj = t % TILE_Y;
  const int i = i0 + ti, j = j0 + tj;
  double v = DBL_MAX;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny;
    const size_t c = (size_t)i * M.ny + j;
    double k[8];
    if (interior(M, i, j)) {
      mhd_rhs_faces(a.S, a.st, a.k, M, i, j, fx[ti * TILE_Y + tj], fx[(ti + 1) * TILE_Y + tj],
                    fy[ti * (TILE_Y + 1) + tj], fy[ti * (TILE_Y + 1) + tj + 1], k);
    } else {
      ghost_rhs(a, N, c, k);
    }
    const StageOut o = stage_finish(a, N, c, k);
    if (a.dt_partial && i >= M.x.dlo && i <= M.x.dhi && j >= M.y.dlo && j <= M.y.dhi) {
      const double dt = mhd_dt_vals(o.rho, o.te, o.mx, o.my, a.st.bex[c] + o.bx, a.st.bey[c] + o.by,
                                    a.st.bez[c] + o.bz, a.k, M, i, j);
      v = (dt < v) ? dt : v;
    }
  }
  if (a.dt_partial) {
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off, 64);
      v = (o < v) ? o : v;
    }
    if (t == 0) a.dt_partial[blockIdx.x] = v;
  }
}

__global__ void __launch_bounds__(256) k_moc_visc_part(MhdSet S, MhdStatic st, MhdConst K, Mesh M,
                                                       double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  double v = DBL_MAX;
  if (i < M.nx && j < M.ny && i >= M.x.dlo && i <= M.x.dhi && j >= M.y.dlo && j <= M.y.dhi) {
    const double dx = M.x.d[i], dy = M.y.d[j];
    const double t = 1.0 / (1.0 / (dx * dx) + 1.0 / (dy * dy)) / mhd_dt_cell(S, st, K, M, i, j);
    v = (t < v) ? t : v;
  }
  v = block_min(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = v;
}

__global__ void __launch_bounds__(1024) k_moc_visc_final(const double* partial, int n, double gv, double* vc,
                                                         int* err) {
  double v = DBL_MAX;
  for (int k = threadIdx.x; k < n; k += blockDim.x) { const double o = partial[k]; v = (o < v) ? o : v; }
  for (int off = 32; off > 0; off >>= 1) { const double o = __shfl_xor(v, off, 64); v = (o < v) ? o : v; }
  __shared__ double part[16];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = (part[k] < m) ? part[k] : m;
    if (!(m < DBL_MAX)) *err = 1;   // Grid::min asserts (grid.cpp:80)
    *vc = gv * 0.5 * m;
  }
}

__global__ void __launch_bounds__(1024) k_reduce_step(const double* partial, int n, double eps, double* step,
                                                      double* history, int idx, int* err) {
  double v = DBL_MAX;
  for (int k = threadIdx.x; k < n; k += blockDim.x) { const double o = partial[k]; v = (o < v) ? o : v; }
  for (int off = 32; off > 0; off >>= 1) { const double o = __shfl_xor(v, off, 64); v = (o < v) ? o : v; }
  __shared__ double part[16];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = (part[k] < m) ? part[k] : m;
    if (!(m < DBL_MAX)) *err = 1;   // reference asserts (grid.cpp:80)
    const double s = eps * m;
    *step = s;
    if (history) history[idx] = s;
  }
}

__global__ void __launch_bounds__(256) k_mhd_rhs_full(MhdSet S, MhdStatic st, MhdConst K, Mesh M, double* out,
                                                     const double* moc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  double k[8];
  if (interior(M, i, j)) mhd_rhs_cell(S, st, K, M, i, j, k);
  else if (moc) for (int v = 0; v < 8; ++v) k[v] = 0.0 + moc[v * N + c];
  else for (int v = 0; v < 8; ++v) k[v] = 0.0;
  for (int v = 0; v < 8; ++v) out[v * N + c] = k[v];
}

// Derived variables of idealmhd.cpp:241-277 for download.
__global__ void __launch_bounds__(256) k_mhd_derived(MhdSet S, MhdStatic st, MhdConst K, Mesh M, int which,
                                                     double* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double rho = S.rho[c];
  const double press = S.te[c] * K.gm1;
  const double vx = S.mx[c] / rho, vy = S.my[c] / rho, vz = S.mz[c] / rho;
  const double bx = st.bex[c] + S.bx[c], by = st.bey[c] + S.by[c], bz = st.bez[c] + S.bz[c];
  const double bmag = sqrt(bx * bx + by * by + bz * bz);
  double r = 0.0;
  switch (which) {
    case 0: r = smax(press / (S.n[c] * (2.0 * K_B)), K.tmin); break;   // temp
    case 1: r = press; break;
    case 2: r = vx; break;
    case 3: r = vy; break;
    case 4: r = vz; break;
    case 5: r = rho * 0.5 * (vx * vx + vy * vy); break;                 // kinetic_energy
    case 6: r = bx; break;
    case 7: r = by; break;
    case 8: r = bz; break;
    case 9: r = bmag; break;
    case 10: r = bmag == 0.0 ? 0.0 : bx / bmag; break;
    case 11: r = bmag == 0.0 ? 0.0 : by / bmag; break;
    case 12: r = bmag == 0.0 ? 0.0 : bz / bmag; break;
    case 13: r = mhd_dt_cell(S, st, K, M, i, j); break;
  }
  out[c] = r;
}

inline dim3 grid2d(int nx, int ny, dim3 b) { return dim3((ny + b.x - 1) / b.x, (nx + b.y - 1) / b.y); }

// n module scalars, the step size and the error flag into host-visible memory in one launch, so the
// host reads everything a step's subcycle counts need after a single synchronisation
__global__ void k_gather_scalars(const double* scal, int n, const double* step, const int* err, double* out) {
  const int t = threadIdx.x;
  if (t < n) out[t] = scal[t];
  if (t == n) out[n] = *step;
  if (t == n + 1) out[n + 1] = (double)*err;
}

}  // namespace

void launch_mhd_stage(const StageArgs& a, hipStream_t s) {
  dim3 b(64, 4);
  hipLaunchKernelGGL(k_mhd_stage, grid2d(a.mesh.nx, a.mesh.ny, b), b, 0, s, a);
}

void launch_mhd_stage_tile(const StageArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mhd_stage_tile, dim3(mhd_tile_blocks(a.mesh)), dim3(TILE_THREADS), 0, s, a);
}

int mhd_tile_blocks(const Mesh& m) {
  return ((m.ny + TILE_Y - 1) / TILE_Y) * ((m.nx + TILE_X - 1) / TILE_X);
}

void launch_mhd_populate(MhdSet L, const double* temp, MhdConst k, int nx, int ny, hipStream_t s) {
  const size_t N = (size_t)nx * ny;
  hipLaunchKernelGGL(k_mhd_populate, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, L, temp, k, nx, ny);
}

void launch_mhd_ghost(MhdSet G, MhdConst k, Mesh mesh, const int* kinds, const double* s1, const double* s2,
                      int mode, hipStream_t s) {
  GhostArgs a;
  a.G = G; a.k = k; a.mesh = mesh; a.mode = mode;
  for (int q = 0; q < 4; ++q) { a.kind[q] = kinds[q]; a.s1[q] = s1[q]; a.s2[q] = s2[q]; }
  // x sides fill rows {0,1,2} and {nx-3..nx-1} from rows 2 and nx-3: independent when nx >= 6
  if (a.kind[0] || a.kind[1]) {
    const int lines = mesh.nx >= 6 ? mesh.ny : 0;
    hipLaunchKernelGGL(k_mhd_ghost, dim3((2 * mesh.ny + 255) / 256), dim3(256), 0, s, a, 1, lines);
  }
  if (a.kind[2] || a.kind[3]) {
    const int lines = mesh.ny >= 6 ? mesh.nx : 0;
    hipLaunchKernelGGL(k_mhd_ghost, dim3((2 * mesh.nx + 255) / 256), dim3(256), 0, s, a, 0, lines);
  }
}

int mhd_dt_blocks(const Mesh& m) {
  dim3 b(64, 4);
  dim3 g = grid2d(m.nx, m.ny, b);
  return (int)(g.x * g.y);
}

void launch_mhd_dt(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double* dt_grid, double* partial, hipStream_t s) {
  dim3 b(64, 4);
  hipLaunchKernelGGL(k_mhd_dt, grid2d(mesh.nx, mesh.ny, b), b, 0, s, L, st, k, mesh, dt_grid, partial);
}

void launch_moc_visc(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, double gv, double* partial, double* vc, int* err,
                     hipStream_t s) {
  dim3 b(64, 4);
  const dim3 g = grid2d(mesh.nx, mesh.ny, b);
  hipLaunchKernelGGL(k_moc_visc_part, g, b, 0, s, S, st, k, mesh, partial);
  hipLaunchKernelGGL(k_moc_visc_final, dim3(1), dim3(1024), 0, s, (const double*)partial, (int)(g.x * g.y), gv, vc, err);
}

void launch_gather_scalars(const double* scal, int n, const double* step, const int* err, double* host_out,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_gather_scalars, dim3(1), dim3(((n + 2 + 63) / 64) * 64), 0, s, scal, n, step, err, host_out);
}

void launch_reduce_step(const double* partial, int n, double eps, double* step, double* history, int idx, int* err,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_step, dim3(1), dim3(1024), 0, s, partial, n, eps, step, history, idx, err);
}

void launch_mhd_rhs_full(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, double* out, hipStream_t s,
                         const double* moc) {
  dim3 b(64, 4);
  hipLaunchKernelGGL(k_mhd_rhs_full, grid2d(mesh.nx, mesh.ny, b), b, 0, s, S, st, k, mesh, out, moc);
}

void launch_mhd_derived(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, int which, double* out, hipStream_t s) {
  dim3 b(64, 4);
  hipLaunchKernelGGL(k_mhd_derived, grid2d(mesh.nx, mesh.ny, b), b, 0, s, S, st, k, mesh, which, out);
}

}  // namespace spruce
