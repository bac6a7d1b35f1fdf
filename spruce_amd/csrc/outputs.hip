// Module file output and multispecies accumulation (PlasmaDomain::storeGrids, fileio.cpp:163-194):
// elementwise kernels whose operation order follows each module's whole-grid expression.
#include "arith.hpp"
#include "outputs.hpp"

namespace spruce {
namespace {

__device__ __forceinline__ bool inside(const Mesh& M, int i, int j) {
  return i >= M.x.lo && i <= M.x.hi && j >= M.y.lo && j <= M.y.hi;
}

__global__ void __launch_bounds__(256) k_out_rate(const double* now, const double* before, double dt, size_t n,
                                                  double* out) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = (now[c] - before[c]) / dt;
}

__global__ void __launch_bounds__(256) k_ms_change(double* acc, const double* now, const double* before, double f,
                                                   int scaled, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double d = now[c] - before[c];
  acc[c] = acc[c] + (scaled ? f * d : d);
}

__global__ void __launch_bounds__(256) k_ms_source(double* acc, const double* src, double dt, const double* dt_dev,
                                                   double f, int mode, Mesh M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= M.nx || j >= M.ny) return;
  const size_t c = (size_t)i * M.ny + j;
  const double mask = inside(M, i, j) ? 1.0 : 0.0;
  const double h = dt_dev ? *dt_dev : dt;
  const double fm = f * mask;
  if (mode == MS_SRC_TEMPLATE) acc[c] = acc[c] + fm * (h * src[c]);
  else if (mode == MS_SRC_RATE) acc[c] = acc[c] + (fm * src[c]) * h;
  else acc[c] = acc[c] - (fm * src[c]) * h;
}

inline dim3 g2(const Mesh& m, dim3 b) { return dim3((m.ny + b.x - 1) / b.x, (m.nx + b.y - 1) / b.y); }

}  // namespace

void launch_out_rate(const double* now, const double* before, double dt, size_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_out_rate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, now, before, dt, n, out);
}

void launch_ms_change(double* acc, const double* now, const double* before, double f, bool scaled, size_t n,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_ms_change, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, acc, now, before, f,
                     scaled ? 1 : 0, n);
}

void launch_ms_source(double* acc, const double* src, double dt, const double* dt_dev, double f, int mode, Mesh mesh,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_ms_source, g2(mesh, dim3(64, 4)), dim3(64, 4), 0, s, acc, src, dt, dt_dev, f, mode, mesh);
}

}  // namespace spruce
