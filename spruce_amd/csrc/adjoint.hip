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

// One thread per face along AX; the face at cell (i, j) lies between (i, j) and its lower neighbour.
template <int AX>
__device__ __forceinline__ void adj_faces_axis(const MhdSet& S, const MhdStatic& st, const MhdConst& K, const Mesh& M,
                                               const double* coef, double* lam_S, int i, int j) {
  const Axis& ax = AX == 0 ? M.x : M.y;
  const int f = AX == 0 ? i : j;
  if (!ax.per && (f < ax.lo || f > ax.hi + 1)) return;      // faces of interior cells only
  const int fl = wrapi(f - 1, ax.n, ax.per);
  const size_t N = (size_t)M.nx * M.ny;
  const size_t cu = (size_t)i * M.ny + j;                     // cell above the face
  const size_t cl = AX == 0 ? (size_t)fl * M.ny + j : (size_t)i * M.ny + fl;
  double fa[ADJ_NCOEF];
  bool any = false;
  for (int q = 0; q < ADJ_NCOEF; ++q) {
    fa[q] = coef[q * N + cl] - coef[q * N + cu];
    any = any || fa[q] != 0.0;
  }
  if (!any) return;
  adj_face<AX>(S, st, K, M, i, j, fa, [&](size_t cell, int v, double val) { atomicAdd(&lam_S[v * N + cell], val); });
}

// x-faces for blockIdx.z == 0, y-faces for blockIdx.z == 1
__global__ void __launch_bounds__(256) k_adj_faces(MhdSet S, MhdStatic st, MhdConst K, Mesh M, const double* cx,
                                                   const double* cy, double* lam_S) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  if (blockIdx.z == 0) adj_faces_axis<0>(S, st, K, M, cx, lam_S, i, j);
  else adj_faces_axis<1>(S, st, K, M, cy, lam_S, i, j);
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

__global__ void __launch_bounds__(64) k_adj_dt_apply(MhdSet L, MhdStatic st, MhdConst K, Mesh M, double eps,
                                                     const double* p1, const double* p2, const double* gh,
                                                     const double* pmin, const int* pidx, int nb, double* lam_L) {
  // the adjoint of the step size: the sum of the stage kernels' block partials, plus the modules'
  double g = 0.0;
  for (int k = threadIdx.x; k < nb; k += 64) g += (p1 ? p1[k] : 0.0) + (p2 ? p2[k] : 0.0);
  for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off, 64);
  if (gh) g += gh[0];
  // (min, lowest index) over the block results: one wavefront, then a shuffle reduction; the rule
  // is order-independent, so the cell is the same however the blocks are visited
  double v = DBL_MAX;
  int idx = -1;
  for (int k = threadIdx.x; k < nb; k += 64)
    if (pmin[k] < v || (pmin[k] == v && pidx[k] >= 0 && (idx < 0 || pidx[k] < idx))) { v = pmin[k]; idx = pidx[k]; }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov < v || (ov == v && oi >= 0 && (idx < 0 || oi < idx))) { v = ov; idx = oi; }
  }
  if (threadIdx.x != 0 || idx < 0) return;
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
  const dim3 b(64, 4), g = g2(mesh, b);
  hipLaunchKernelGGL(k_adj_faces, dim3(g.x, g.y, 2), b, 0, s, S, st, k, mesh, cx, cy, lam_S);
}

void launch_adj_dt(MhdSet L, MhdStatic st, MhdConst k, Mesh mesh, double eps, const double* p1, const double* p2,
                   const double* gh, double* pmin, int* pidx, double* lam_L, hipStream_t s) {
  const dim3 b(64, 4), g = g2(mesh, b);
  hipLaunchKernelGGL(k_adj_dt_blocks, g, b, 0, s, L, st, k, mesh, pmin, pidx);
  hipLaunchKernelGGL(k_adj_dt_apply, dim3(1), dim3(64), 0, s, L, st, k, mesh, eps, p1, p2, gh, (const double*)pmin,
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
