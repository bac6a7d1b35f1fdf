// Discrete-adjoint kernels of the IdealMHD time step for gfx950 (reverse of mhd.hip's stage, dt and
// populate kernels).  The per-cell algebra is adj_cell.hpp; these kernels sequence it over the
// grid.  Face contributions land on four cells each and are accumulated with FP64 atomics, so the
// gradient is reproducible to rounding, not bit for bit.
#include <cfloat>

#include "adj_cell.hpp"
#include "adjoint.hpp"
#include "arith.hpp"

namespace spruce {
namespace {

__device__ __forceinline__ bool interior_cell(const Mesh& M, int i, int j) {
  return i >= M.x.lo && i <= M.x.hi && j >= M.y.lo && j <= M.y.hi;
}

__device__ __forceinline__ double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double part[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  if ((lane & 63) == 0) part[lane >> 6] = v;
  __syncthreads();
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    double m = 0.0;
    for (int k = 0; k < nw; ++k) m += part[k];
    v = m;
  }
  return v;
}

// Reverse of one stage for one cell: T = minimums/derived(B + mult*h*kv) reversed (lam_B gets mu,
// lamk = mult*h*mu, the block's partial of mult*mu.kv), then the cell's own part of the transposed
// time derivative at S applied to lamk (its face coefficients to cx, cy; its own part to lam_S).
// assign_B / assign_S: this launch is the first writer of lam_B / lam_S in the step (no clear).
__global__ void __launch_bounds__(256) k_adj_stage_cells(MhdSet B, const double* kv, const double* step, double mult,
                                                         MhdSet S, MhdStatic st, MhdConst K, Mesh M,
                                                         const double* lam_T, double* lam_B, int assign_B,
                                                         double* lam_S, int assign_S, double* cx, double* cy,
                                                         double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double h = mult * (*step);
    // derivative of the minimums/derived maps (idealmhd.cpp:237-238, 246-252): rho passes when
    // neither floor is active, te when its floor is not
    const double rho_new = B.rho[c] + kv[c] * h;
    const double x = rho_new / K.m;
    const double rho1 = smax(x, K.nmin) * K.m;
    const double drho = (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
    const double te_new = B.te[c] + kv[4 * N + c] * h;
    const double dte = te_new < K.temin ? 0.0 : 1.0;
    double lamk[8];
    for (int v = 0; v < 8; ++v) {
      const double d = v == 0 ? drho : (v == 4 ? dte : 1.0);
      const double mu = lam_T[v * N + c] * d;
      lam_B[v * N + c] = assign_B ? mu : lam_B[v * N + c] + mu;
      lamk[v] = h * mu;
      dot += mu * kv[v * N + c];
    }
    dot *= mult;
    double fx[ADJ_NCOEF], fy[ADJ_NCOEF];
    if (interior_cell(M, i, j)) {
      double own[8];
      adj_cell_coeffs(S, st, K, M, i, j, lamk, own, fx, fy);
      for (int v = 0; v < 8; ++v) lam_S[v * N + c] = assign_S ? own[v] : lam_S[v * N + c] + own[v];
    } else {
      for (int q = 0; q < ADJ_NCOEF; ++q) { fx[q] = 0.0; fy[q] = 0.0; }
      if (assign_S) for (int v = 0; v < 8; ++v) lam_S[v * N + c] = 0.0;
    }
    for (int q = 0; q < ADJ_NCOEF; ++q) { cx[q * N + c] = fx[q]; cy[q * N + c] = fy[q]; }
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
}

// Reverse of the pointwise part of one stage, T = minimums/derived(B + mult*h*kv), alone:
// lam_B (+)= mu = dT/dpre^T lam_T, lamk = mult*h*mu (written), partial: block sums of mult*mu.kv.
__global__ void __launch_bounds__(256) k_adj_update(MhdSet B, const double* kv, const double* step, double mult,
                                                    MhdConst K, Mesh M, const double* lam_T, double* lam_B,
                                                    int assign_B, double* lamk, double* partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double dot = 0.0;
  if (i < M.nx && j < M.ny) {
    const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
    const double h = mult * (*step);
    const double rho_new = B.rho[c] + kv[c] * h;
    const double x = rho_new / K.m;
    const double rho1 = smax(x, K.nmin) * K.m;
    const double drho = (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
    const double te_new = B.te[c] + kv[4 * N + c] * h;
    const double dte = te_new < K.temin ? 0.0 : 1.0;
    for (int v = 0; v < 8; ++v) {
      const double d = v == 0 ? drho : (v == 4 ? dte : 1.0);
      const double mu = lam_T[v * N + c] * d;
      lam_B[v * N + c] = assign_B ? mu : lam_B[v * N + c] + mu;
      lamk[v * N + c] = h * mu;
      dot += mu * kv[v * N + c];
    }
    dot *= mult;
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = dot;
}

// The per-cell part of the transposed time derivative at S applied to lamk = sa*ka + kb (kb may be
// null): its own part to lam_S, its face coefficients to cx, cy (then k_adj_faces).
__global__ void __launch_bounds__(256) k_adj_rhs_cells(MhdSet S, MhdStatic st, MhdConst K, Mesh M, const double* ka,
                                                       double sa, const double* kb, double* lam_S, int assign_S,
                                                       double* cx, double* cy) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t N = (size_t)M.nx * M.ny, c = (size_t)i * M.ny + j;
  double fx[ADJ_NCOEF], fy[ADJ_NCOEF];
  if (interior_cell(M, i, j)) {
    double lamk[8], own[8];
    for (int v = 0; v < 8; ++v) lamk[v] = kb ? sa * ka[v * N + c] + kb[v * N + c] : sa * ka[v * N + c];
    adj_cell_coeffs(S, st, K, M, i, j, lamk, own, fx, fy);
    for (int v = 0; v < 8; ++v) lam_S[v * N + c] = assign_S ? own[v] : lam_S[v * N + c] + own[v];
  } else {
    for (int q = 0; q < ADJ_NCOEF; ++q) { fx[q] = 0.0; fy[q] = 0.0; }
    if (assign_S) for (int v = 0; v < 8; ++v) lam_S[v * N + c] = 0.0;
  }
  for (int q = 0; q < ADJ_NCOEF; ++q) { cx[q * N + c] = fx[q]; cy[q * N + c] = fy[q]; }
}

// The face scatter without a global atomic per contribution.  x faces: one thread per column
// marches a strip of FX_ROWS faces down the rows; a face adds to the rows f-2 .. f+1 of its column,
// so the thread keeps those four rows' adjoint in registers and writes a row once no later face of
// the strip can reach it: plainly where no other strip reaches the row either (rows fs+1 ..
// fe-3 of the strip [fs, fe)), atomically otherwise.  y faces: a 64 x 4 block sums its faces'
// contributions per row in LDS and writes each cell once, plainly for the cells only its own faces
// reach (columns j0+1 .. j0+61), atomically for the rest.  The x and y launches run one after the
// other, so no cell is written plainly by two threads at once.
constexpr int FX_ROWS = 16;

__global__ void __launch_bounds__(64) k_adj_faces_x(MhdSet S, MhdStatic st, MhdConst K, Mesh M, const double* cx,
                                                    double* lam_S) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= M.ny) return;
  const Axis& ax = M.x;
  const size_t N = (size_t)M.nx * M.ny;
  const int fs = blockIdx.y * FX_ROWS, fe = min(fs + FX_ROWS, M.nx);
  double acc[4][8];
  for (int a = 0; a < 4; ++a)
    for (int v = 0; v < 8; ++v) acc[a][v] = 0.0;
  auto flush = [&](int row, const double* a8, bool plain) {
    if (!ax.per && (row < 0 || row >= ax.n)) return;
    const size_t c = (size_t)wrapi(row, ax.n, ax.per) * M.ny + j;
    for (int v = 0; v < 8; ++v) {
      if (a8[v] == 0.0) continue;
      if (plain) lam_S[v * N + c] += a8[v];
      else atomicAdd(&lam_S[v * N + c], a8[v]);
    }
  };
  for (int f = fs; f < fe; ++f) {
    if (ax.per || (f >= ax.lo && f <= ax.hi + 1)) {     // faces of interior cells only
      const int fl = wrapi(f - 1, ax.n, ax.per);
      const size_t cu = (size_t)f * M.ny + j, cl = (size_t)fl * M.ny + j;
      double fa[ADJ_NCOEF];
      bool any = false;
      for (int q = 0; q < ADJ_NCOEF; ++q) {
        fa[q] = cx[q * N + cl] - cx[q * N + cu];
        any = any || fa[q] != 0.0;
      }
      if (any) adj_face_slots<0>(S, st, K, M, f, j, fa, [&](int slot, int v, double val) { acc[slot][v] += val; });
    }
    // row f-2 is complete: later faces of the strip reach rows f-1 and above
    flush(f - 2, acc[0], f - 3 >= fs);
    for (int a = 0; a < 3; ++a)
      for (int v = 0; v < 8; ++v) acc[a][v] = acc[a + 1][v];
    for (int v = 0; v < 8; ++v) acc[3][v] = 0.0;
  }
  for (int a = 0; a < 3; ++a) flush(fe - 2 + a, acc[a], false);   // rows the next strip reaches too
}

constexpr int FY_W = 64 + 3;   // cells j0-2 .. j0+64 of a row

__global__ void __launch_bounds__(256) k_adj_faces_y(MhdSet S, MhdStatic st, MhdConst K, Mesh M, const double* cy,
                                                     double* lam_S) {
  __shared__ double acc[4][8][FY_W];
  const int li = threadIdx.y, lj = threadIdx.x;
  for (int t = li * 64 + lj; t < 4 * 8 * FY_W; t += 256) (&acc[0][0][0])[t] = 0.0;
  __syncthreads();
  const Axis& ax = M.y;
  const size_t N = (size_t)M.nx * M.ny;
  const int i = blockIdx.y * 4 + li, j0 = blockIdx.x * 64, j = j0 + lj;
  if (i < M.nx && j < M.ny && (ax.per || (j >= ax.lo && j <= ax.hi + 1))) {
    const int fl = wrapi(j - 1, ax.n, ax.per);
    const size_t cu = (size_t)i * M.ny + j, cl = (size_t)i * M.ny + fl;
    double fa[ADJ_NCOEF];
    bool any = false;
    for (int q = 0; q < ADJ_NCOEF; ++q) {
      fa[q] = cy[q * N + cl] - cy[q * N + cu];
      any = any || fa[q] != 0.0;
    }
    if (any)
      adj_face_slots<1>(S, st, K, M, i, j, fa,
                        [&](int slot, int v, double val) { atomicAdd(&acc[li][v][lj + slot], val); });
  }
  __syncthreads();
  for (int t = li * 64 + lj; t < 4 * 8 * FY_W; t += 256) {
    const int r = t / (8 * FY_W), rem = t - r * (8 * FY_W), v = rem / FY_W, col = rem - v * FY_W;
    const double val = acc[r][v][col];
    const int ii = blockIdx.y * 4 + r;
    if (val == 0.0 || ii >= M.nx) continue;
    const int jj = j0 - 2 + col;
    if (!ax.per && (jj < 0 || jj >= ax.n)) continue;
    const size_t c = (size_t)ii * M.ny + wrapi(jj, ax.n, ax.per);
    // owned: the faces jj-1 .. jj+2 are all lanes of this block (none wraps round a periodic axis)
    if (col >= 3 && col <= 63 && jj + 2 < ax.n) lam_S[v * N + c] += val;
    else atomicAdd(&lam_S[v * N + c], val);
  }
}

// per-block (min dt, index) over the interior, std::min semantics (NaN never selected)
__global__ void __launch_bounds__(256) k_adj_dt_blocks(MhdSet L, MhdStatic st, MhdConst K, Mesh M, double* pmin,
                                                       int* pidx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  double v = DBL_MAX;
  int idx = -1;
  if (i < M.nx && j < M.ny && interior_cell(M, i, j)) {
    const double dt = dt_cell_dual(L, st, K, M, i, j).v;
    if (dt < v) { v = dt; idx = i * M.ny + j; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov < v || (ov == v && oi >= 0 && (idx < 0 || oi < idx))) { v = ov; idx = oi; }
  }
  __shared__ double pv[16];
  __shared__ int pi[16];
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  if ((lane & 63) == 0) { pv[lane >> 6] = v; pi[lane >> 6] = idx; }
  __syncthreads();
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    for (int k = 1; k < nw; ++k)
      if (pv[k] < v || (pv[k] == v && pi[k] >= 0 && (idx < 0 || pi[k] < idx))) { v = pv[k]; idx = pi[k]; }
    pmin[blockIdx.y * gridDim.x + blockIdx.x] = v;
    pidx[blockIdx.y * gridDim.x + blockIdx.x] = idx;
  }
}

__global__ void __launch_bounds__(1024) k_adj_dt_apply(MhdSet L, MhdStatic st, MhdConst K, Mesh M, double eps,
                                                       const double* p1, const double* p2, const double* gh,
                                                       const double* pmin, const int* pidx, int nb, double* lam_L) {
  // the adjoint of the step size: the sum of the stage kernels' block partials, plus the modules';
  // and the (min, lowest index) over the block results, a rule independent of the visiting order.
  // Sixteen wavefronts, each reduced by shuffles, then combined by the first thread.
  __shared__ double sg[16], sv[16];
  __shared__ int si[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double g = 0.0, v = DBL_MAX;
  int idx = -1;
  for (int k = t; k < nb; k += 1024) {
    g += (p1 ? p1[k] : 0.0) + (p2 ? p2[k] : 0.0);
    if (pmin[k] < v || (pmin[k] == v && pidx[k] >= 0 && (idx < 0 || pidx[k] < idx))) { v = pmin[k]; idx = pidx[k]; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    g += __shfl_xor(g, off, 64);
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov < v || (ov == v && oi >= 0 && (idx < 0 || oi < idx))) { v = ov; idx = oi; }
  }
  if (lane == 0) { sg[w] = g; sv[w] = v; si[w] = idx; }
  __syncthreads();
  if (t != 0) return;
  g = 0.0;
  v = DBL_MAX;
  idx = -1;
  for (int q = 0; q < 16; ++q) {
    g += sg[q];
    if (sv[q] < v || (sv[q] == v && si[q] >= 0 && (idx < 0 || si[q] < idx))) { v = sv[q]; idx = si[q]; }
  }
  if (gh) g += gh[0];
  if (idx < 0) return;
  const int i = idx / M.ny, j = idx - i * M.ny;
  const Dual8 dt = dt_cell_dual(L, st, K, M, i, j);
  const size_t N = (size_t)M.nx * M.ny;
  const double s = eps * g;
  for (int v8 = 0; v8 < 8; ++v8) lam_L[v8 * N + idx] += s * dt.d[v8];
}

// populateVariablesFromState (equationset.cpp:96-104; idealmhd.cpp:226-252) reversed
__global__ void __launch_bounds__(256) k_adj_populate(const double* rho_s, const double* temp_s, const double* lam,
                                                      MhdConst K, size_t N, double* g) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double rho = rho_s[c], T = temp_s[c];
  const double x = rho / K.m;
  const double n0 = smax(x, K.nmin);
  const double Tm = smax(T, K.tmin);
  const double te = n0 * 2.0 * K_B * Tm / K.gm1;
  const double dte = te < K.temin ? 0.0 : 1.0;
  const double rho1 = n0 * K.m;
  const double drho = (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
  const double l_rho = lam[c], l_te = lam[4 * N + c];
  g[0 * N + c] = l_rho * drho + l_te * dte * (2.0 * K_B * Tm / K.gm1) * (x < K.nmin ? 0.0 : 1.0 / K.m);
  g[1 * N + c] = l_te * dte * (n0 * 2.0 * K_B / K.gm1) * (T < K.tmin ? 0.0 : 1.0);
  g[2 * N + c] = lam[1 * N + c];
  g[3 * N + c] = lam[2 * N + c];
  g[4 * N + c] = lam[3 * N + c];
  g[5 * N + c] = lam[5 * N + c];
  g[6 * N + c] = lam[6 * N + c];
  g[7 * N + c] = lam[7 * N + c];
}

__global__ void __launch_bounds__(256) k_fill(double* p, double v, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) p[c] = v;
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

int adj_blocks(const Mesh& m) {
  const dim3 g = g2(m, dim3(64, 4));
  return (int)(g.x * g.y);
}

void launch_adj_stage_cells(MhdSet B, const double* kv, const double* step, double mult, MhdSet S, MhdStatic st,
                            MhdConst k, Mesh mesh, const double* lam_T, double* lam_B, int assign_B, double* lam_S,
                            int assign_S, double* cx, double* cy, double* partial, hipStream_t s) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_adj_stage_cells, g2(mesh, b), b, 0, s, B, kv, step, mult, S, st, k, mesh, lam_T, lam_B,
                     assign_B, lam_S, assign_S, cx, cy, partial);
}

void launch_adj_faces(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, const double* cx, const double* cy,
                      double* lam_S, hipStream_t s) {
  const unsigned cols = (unsigned)((mesh.ny + 63) / 64);
  hipLaunchKernelGGL(k_adj_faces_x, dim3(cols, (unsigned)((mesh.nx + FX_ROWS - 1) / FX_ROWS)), dim3(64), 0, s, S, st,
                     k, mesh, cx, lam_S);
  hipLaunchKernelGGL(k_adj_faces_y, dim3(cols, (unsigned)((mesh.nx + 3) / 4)), dim3(64, 4), 0, s, S, st, k, mesh, cy,
                     lam_S);
}

void launch_adj_dt(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double eps, const double* p1, const double* p2,
                   const double* gh, double* pmin, int* pidx, double* lam_L, hipStream_t s) {
  const dim3 b(64, 4), g = g2(mesh, b);
  hipLaunchKernelGGL(k_adj_dt_blocks, g, b, 0, s, L, st, k, mesh, pmin, pidx);
  hipLaunchKernelGGL(k_adj_dt_apply, dim3(1), dim3(1024), 0, s, L, st, k, mesh, eps, p1, p2, gh, (const double*)pmin,
                     (const int*)pidx, (int)(g.x * g.y), lam_L);
}

void launch_adj_populate(const double* rho_s, const double* temp_s, const double* lam, MhdConst k, size_t N,
                         double* g_state, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_populate, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, rho_s, temp_s, lam, k, N,
                     g_state);
}

void launch_fill(double* p, double v, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}


void launch_adj_update(MhdSet B, const double* kv, const double* step, double mult, MhdConst k, Mesh mesh,
                       const double* lam_T, double* lam_B, int assign_B, double* lamk, double* partial,
                       hipStream_t s) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_adj_update, g2(mesh, b), b, 0, s, B, kv, step, mult, k, mesh, lam_T, lam_B, assign_B, lamk, partial);
}

void launch_adj_rhs_cells(MhdSet S, MhdStatic st, MhdConst k, Mesh mesh, const double* ka, double sa,
                          const double* kb, double* lam_S, int assign_S, double* cx, double* cy, hipStream_t s) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_adj_rhs_cells, g2(mesh, b), b, 0, s, S, st, k, mesh, ka, sa, kb, lam_S, assign_S, cx, cy);
}

}  // namespace spruce
