// Host check of the hand-written reverse of the conduction energy derivative
// (tc_core_reverse in spruce_amd/csrc/tc_cell.hpp) against the same function in forward mode
// (tc_core<Dn<15>>) and against central differences of tc_core<double>, at every interior cell of
// a small non-uniform mesh, periodic and with ghost rows.  Compiled by
// tests/test_adjoint_algebra.py with hipcc; runs on the host only.  Prints the worst relative
// differences: "dual <x>" and "fd <x>".
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../spruce_amd/csrc/tc_cell.hpp"

using namespace spruce;

struct AxisHost {
  std::vector<double> d, rd, h, hsq, rhsq, fden, rfden, gup, gdn;
  void build(const std::vector<double>& dc, bool per) {
    const int n = (int)dc.size();
    d = dc; rd.resize(n); h.resize(n); hsq.resize(n); rhsq.resize(n);
    fden.assign(n, 0.0); rfden.assign(n, 0.0); gup.assign(n, 0.0); gdn.assign(n, 0.0);
    for (int i = 0; i < n; ++i) { rd[i] = 1.0 / d[i]; h[i] = 0.5 * d[i]; hsq[i] = h[i] * h[i]; rhsq[i] = 1.0 / hsq[i]; }
    auto w = [&](int k) { return per ? ((k % n) + n) % n : k; };
    for (int c = 0; c < n; ++c) {
      const int cm1 = w(c - 1);
      if (cm1 >= 0 && cm1 < n) { fden[c] = h[c] + h[cm1]; rfden[c] = 1.0 / fden[c]; }
    }
  }
  Axis view(int n, int lo, int hi, int per) const {
    return Axis{d.data(), rd.data(), h.data(), hsq.data(), rhsq.data(), fden.data(), rfden.data(), gup.data(),
                gdn.data(), n, lo, hi, per};
  }
};

int main() {
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  double worst_dual = 0.0, worst_fd = 0.0;
  for (int per = 0; per < 2; ++per) {
    const int nx = 9, ny = 8;
    std::vector<double> dx(nx), dy(ny);
    for (int i = 0; i < nx; ++i) dx[i] = 5e6 * (1.0 + 0.2 * std::sin(0.9 * i));
    for (int j = 0; j < ny; ++j) dy[j] = 4e6 * std::pow(1.1, j);
    AxisHost ax, ay;
    ax.build(dx, per);
    ay.build(dy, per);
    const int xl = per ? 0 : 2, xu = per ? nx - 1 : nx - 3, yl = per ? 0 : 2, yu = per ? ny - 1 : ny - 3;
    Mesh M{ax.view(nx, xl, xu, per), ay.view(ny, yl, yu, per), nx, ny};
    for (int i = xl; i <= xu; ++i)
      for (int j = yl; j <= yu; ++j) {
        double T[3][3], fz[6];
        for (auto& r : T) for (double& v : r) v = 1e6 * (1.0 + 0.3 * U(rng));
        const double bx = U(rng), by = U(rng);
        const double nb = std::sqrt(bx * bx + by * by + 0.1);
        fz[0] = bx / nb; fz[1] = by / nb;
        for (int k = 2; k < 6; ++k) fz[k] = 1e-8 * U(rng);
        const double k0 = 1e-6, ld = U(rng);
        double gT[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gfz[6] = {0, 0, 0, 0, 0, 0};
        tcc::tc_core_reverse(M, i, j, T, fz[0], fz[1], fz[2], fz[3], fz[4], fz[5], k0, ld, gT, gfz);
        double g[15];
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) g[3 * a + b] = gT[a][b];
        for (int k = 0; k < 6; ++k) g[9 + k] = gfz[k];
        // forward mode
        Dn<15> xs[15];
        double x[15];
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) x[3 * a + b] = T[a][b];
        for (int k = 0; k < 6; ++k) x[9 + k] = fz[k];
        for (int k = 0; k < 15; ++k) xs[k] = dn_var<15>(x[k], k);
        const Dn<15> Td[3][3] = {{xs[0], xs[1], xs[2]}, {xs[3], xs[4], xs[5]}, {xs[6], xs[7], xs[8]}};
        const Dn<15> r = tcc::tc_core(M, i, j, Td, xs[9], xs[10], xs[11], xs[12], xs[13], xs[14], k0);
        double scale = 0.0;
        for (int k = 0; k < 15; ++k) scale = std::fmax(scale, std::fabs(ld * r.d[k] * x[k]));
        for (int k = 0; k < 15; ++k) {
          const double e = std::fabs(g[k] - ld * r.d[k]) * std::fabs(x[k]) / scale;
          worst_dual = std::fmax(worst_dual, e);
          // central difference in x[k]
          const double h = 1e-6 * std::fabs(x[k]);
          double xp[15], xm[15];
          for (int q = 0; q < 15; ++q) { xp[q] = x[q]; xm[q] = x[q]; }
          xp[k] += h; xm[k] -= h;
          auto F = [&](const double* v) {
            const double Tv[3][3] = {{v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]}};
            return tcc::tc_core<double>(M, i, j, Tv, v[9], v[10], v[11], v[12], v[13], v[14], k0);
          };
          const double fd = ld * (F(xp) - F(xm)) / (xp[k] - xm[k]);
          worst_fd = std::fmax(worst_fd, std::fabs(g[k] - fd) * std::fabs(x[k]) / scale);
        }
      }
  }
  std::printf("dual %.3e\nfd %.3e\n", worst_dual, worst_fd);
  return (worst_dual < 1e-12 && worst_fd < 1e-6) ? 0 : 1;
}
