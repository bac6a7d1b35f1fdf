// Host check of the hand-written adjoint algebra (spruce_amd/csrc/adj_cell.hpp) against central
// finite differences of the forward per-cell functions it reverses (mhd_cell.hpp), on a small
// periodic, non-uniform mesh.  Compiled by tests/test_adjoint_algebra.py with hipcc; runs on the
// host only (no GPU calls).  Prints one line per check: name, relative error.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../spruce_amd/csrc/adj_cell.hpp"

using namespace spruce;

struct AxisHost {
  std::vector<double> d, rd, h, hsq, rhsq, fden, rfden, gup, gdn;
  void build(const std::vector<double>& dc, bool per) {
    const int n = (int)dc.size();
    d = dc; rd.resize(n); h.resize(n); hsq.resize(n); rhsq.resize(n);
    fden.assign(n, 0.0); rfden.assign(n, 0.0); gup.assign(n, 0.0); gdn.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
      rd[i] = 1.0 / d[i]; h[i] = 0.5 * d[i]; hsq[i] = h[i] * h[i]; rhsq[i] = 1.0 / hsq[i];
    }
    auto w = [&](int k) { return per ? ((k % n) + n) % n : k; };
    for (int c = 0; c < n; ++c) {
      const int cm1 = w(c - 1), cm2 = w(c - 2), cp1 = w(c + 1);
      if (cm1 >= 0 && cm1 < n) { fden[c] = h[c] + h[cm1]; rfden[c] = 1.0 / fden[c]; }
      if (cm2 >= 0 && cm2 < n && cm1 >= 0) gup[c] = h[cm2] + 2.0 * h[cm1];
      if (cp1 >= 0 && cp1 < n) gdn[c] = h[cp1] + 2.0 * h[c];
    }
  }
  Axis view(int n) const {
    return Axis{d.data(), rd.data(), h.data(), hsq.data(), rhsq.data(), fden.data(), rfden.data(), gup.data(),
                gdn.data(), n, 0, n - 1, 1};
  }
};

struct State {
  std::vector<double> a[10];  // rho, n, mx, my, mz, te, bx, by, bz, rho_pre
  MhdSet view() {
    return MhdSet{a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(), a[5].data(), a[6].data(),
                  a[7].data(), a[8].data(), a[9].data()};
  }
};
// evolved index -> State slot
static const int SLOT[8] = {0, 2, 3, 4, 5, 6, 7, 8};

int main(int argc, char** argv) {
  // argv[1]: field strength factor (1: sound speed dominates; 100: Alfven speed dominates)
  const double bf = argc > 1 ? std::atof(argv[1]) : 1.0;
  const int nx = 12, ny = 10;
  const size_t N = (size_t)nx * ny;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> dx(nx), dy(ny);
  for (int i = 0; i < nx; ++i) dx[i] = 1e7 * (1.0 + 0.1 * std::sin(0.7 * i));
  for (int j = 0; j < ny; ++j) dy[j] = 8e6 * (1.0 + 0.15 * std::cos(0.9 * j));
  AxisHost ax, ay;
  ax.build(dx, true);
  ay.build(dy, true);
  Mesh M{ax.view(nx), ay.view(ny), nx, ny};
  MhdConst K{};
  K.m = 1.6726e-24; K.rm = 1.0 / K.m; K.gamma = 5.0 / 3.0; K.gm1 = K.gamma - 1.0; K.rgm1 = 1.0 / K.gm1;
  K.nmin = 1.0; K.temin = 1e-12; K.tmin = 1.0; K.eps = 0.2; K.r4pi = 1.0 / (4.0 * PI); K.r6 = 1.0 / 6.0; K.r3 = 1.0 / 3.0;
  std::vector<double> bex(N), bey(N), bez(N), gx(N, 0.0), gy(N, -2.74e4);
  State S;
  for (auto& v : S.a) v.assign(N, 0.0);
  for (size_t c = 0; c < N; ++c) {
    const double rho = 1e-15 * (1.0 + 0.3 * U(rng));
    S.a[0][c] = rho;
    S.a[1][c] = rho / K.m;
    S.a[2][c] = rho * 1e7 * U(rng);
    S.a[3][c] = rho * 1e7 * U(rng);
    S.a[4][c] = rho * 1e7 * U(rng);
    S.a[5][c] = 2.0 * (rho / K.m) * K_B * 1e6 * (1.0 + 0.3 * U(rng)) / K.gm1;
    S.a[6][c] = bf * U(rng);
    S.a[7][c] = bf * U(rng);
    S.a[8][c] = bf * U(rng);
    bex[c] = bf * (2.0 + 0.5 * U(rng)); bey[c] = bf * (1.0 + 0.5 * U(rng)); bez[c] = bf * 0.3;
  }
  MhdStatic st{bex.data(), bey.data(), bez.data(), gx.data(), gy.data()};
  const double scale[8] = {1e-15, 1e-8, 1e-8, 1e-8, 0.04, bf, bf, bf};

  // J(S) = sum_c lam(c) . k(c)
  std::vector<double> lam(8 * N);
  for (auto& v : lam) v = U(rng);
  auto J = [&](State& X) {
    MhdSet v = X.view();
    double s = 0.0;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        double k[8];
        mhd_rhs_cell(v, st, K, M, i, j, k);
        for (int q = 0; q < 8; ++q) s += lam[q * N + (size_t)i * ny + j] * k[q];
      }
    return s;
  };
  // adjoint: A = dJ/dS
  std::vector<double> A(8 * N, 0.0), cx(ADJ_NCOEF * N), cy(ADJ_NCOEF * N);
  {
    MhdSet v = S.view();
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const size_t c = (size_t)i * ny + j;
        double l[8], own[8], fx[ADJ_NCOEF], fy[ADJ_NCOEF];
        for (int q = 0; q < 8; ++q) l[q] = lam[q * N + c];
        adj_cell_coeffs(v, st, K, M, i, j, l, own, fx, fy);
        for (int q = 0; q < ADJ_NCOEF; ++q) { cx[q * N + c] = fx[q]; cy[q * N + c] = fy[q]; }
        for (int q = 0; q < 8; ++q) A[q * N + c] += own[q];
      }
    auto add = [&](size_t cell, int var, double val) { A[var * N + cell] += val; };
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const size_t cu = (size_t)i * ny + j;
        double fa[ADJ_NCOEF];
        const size_t clx = (size_t)((i + nx - 1) % nx) * ny + j;
        for (int q = 0; q < ADJ_NCOEF; ++q) fa[q] = cx[q * N + clx] - cx[q * N + cu];
        adj_face<0>(v, st, K, M, i, j, fa, add);
        const size_t cly = (size_t)i * ny + (j + ny - 1) % ny;
        for (int q = 0; q < ADJ_NCOEF; ++q) fa[q] = cy[q * N + cly] - cy[q * N + cu];
        adj_face<1>(v, st, K, M, i, j, fa, add);
      }
  }
  double worst = 0.0;
  for (int trial = 0; trial < 6; ++trial) {
    std::vector<double> dir(8 * N);
    for (int q = 0; q < 8; ++q)
      for (size_t c = 0; c < N; ++c) dir[q * N + c] = (trial < 5 ? U(rng) : (q == trial % 8 ? U(rng) : 0.0)) * scale[q];
    double ad = 0.0;
    for (size_t k = 0; k < 8 * N; ++k) ad += A[k] * dir[k];
    const double eps = 1e-6;
    State P = S, Q = S;
    for (int q = 0; q < 8; ++q)
      for (size_t c = 0; c < N; ++c) {
        P.a[SLOT[q]][c] += eps * dir[q * N + c];
        Q.a[SLOT[q]][c] -= eps * dir[q * N + c];
      }
    const double fd = (J(P) - J(Q)) / (2.0 * eps);
    const double rel = std::fabs(ad - fd) / (std::fabs(ad) + std::fabs(fd) + 1e-300);
    std::printf("rhs_adjoint_dir%d %.3e (adj %.10e fd %.10e)\n", trial, rel, ad, fd);
    worst = std::fmax(worst, rel);
  }
  // per-variable check: all weight on one evolved variable at a time
  for (int q = 0; q < 8; ++q) {
    std::vector<double> dir(8 * N, 0.0);
    for (size_t c = 0; c < N; ++c) dir[q * N + c] = U(rng) * scale[q];
    double ad = 0.0;
    for (size_t k = 0; k < 8 * N; ++k) ad += A[k] * dir[k];
    // a single variable can carry a small share of J: take the better of two step sizes, so the
    // cancellation error of the difference does not dominate
    double rel = 1.0;
    for (double eps : {1e-6, 1e-5}) {
      State P = S, Q = S;
      for (size_t c = 0; c < N; ++c) { P.a[SLOT[q]][c] += eps * dir[q * N + c]; Q.a[SLOT[q]][c] -= eps * dir[q * N + c]; }
      const double fd = (J(P) - J(Q)) / (2.0 * eps);
      rel = std::fmin(rel, std::fabs(ad - fd) / (std::fabs(ad) + std::fabs(fd) + 1e-300));
    }
    std::printf("rhs_adjoint_var%d %.3e\n", q, rel);
    worst = std::fmax(worst, rel);
  }
  // step-size gradient of one cell
  {
    MhdSet v = S.view();
    const int i = 5, j = 4;
    const size_t c = (size_t)i * ny + j;
    const Dual8 dt = dt_cell_dual(v, st, K, M, i, j);
    double num = 0.0, den = 0.0;
    for (int q = 0; q < 8; ++q) {
      const double e = 1e-6 * scale[q];
      State P = S, Q = S;
      P.a[SLOT[q]][c] += e;
      Q.a[SLOT[q]][c] -= e;
      const double fd = (dt_cell_dual(P.view(), st, K, M, i, j).v - dt_cell_dual(Q.view(), st, K, M, i, j).v) / (2 * e);
      // compared in the variables' own scales: the field derivatives vanish when the sound speed
      // exceeds the Alfven speed (the fast speed then equals the sound speed)
      num = std::fmax(num, std::fabs(dt.d[q] - fd) * scale[q]);
      den = std::fmax(den, std::fabs(fd) * scale[q]);
    }
    const double wd = num / den;
    std::printf("dt_gradient %.3e\n", wd);
    worst = std::fmax(worst, wd);
  }
  // ---- full RK2 steps: the step-level reverse the GPU driver performs (minimums/derived maps,
  // step-size dependence through the argmin cell), against finite differences of J = w . rho_N
  {
    auto rhs_all = [&](State& X, std::vector<double>& k) {
      MhdSet v = X.view();
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
          double kk[8];
          mhd_rhs_cell(v, st, K, M, i, j, kk);
          for (int q = 0; q < 8; ++q) k[q * N + (size_t)i * ny + j] = kk[q];
        }
    };
    auto phi = [&](const State& B, const std::vector<double>& k, double h, State& T) {
      for (size_t c = 0; c < N; ++c) {
        const double rho_new = B.a[0][c] + k[c] * h;
        const double rho1 = smax(rho_new / K.m, K.nmin) * K.m;
        const double n = smax(rho1 / K.m, K.nmin);
        T.a[0][c] = n * K.m;
        T.a[1][c] = n;
        for (int q = 1; q < 8; ++q) T.a[SLOT[q]][c] = B.a[SLOT[q]][c] + k[q * N + c] * h;
        T.a[5][c] = smax(smax(T.a[5][c], K.temin), K.temin);
      }
    };
    auto hstep = [&](State& X, int* arg) {
      MhdSet v = X.view();
      double m = DBL_MAX;
      int idx = -1;
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
          const double dt = dt_cell_dual(v, st, K, M, i, j).v;
          if (dt < m) { m = dt; idx = i * ny + j; }
        }
      if (arg) *arg = idx;
      return K.eps * m;
    };
    const int nsteps = 3;
    std::vector<double> w(N);
    for (auto& x : w) x = U(rng);
    auto forward = [&](State X, std::vector<State>* Ls, std::vector<double>* hs) {
      std::vector<double> k1(8 * N), k2(8 * N);
      State Wst = X;
      for (int n = 0; n < nsteps; ++n) {
        const double h = hstep(X, nullptr);
        if (Ls) Ls->push_back(X);
        if (hs) hs->push_back(h);
        rhs_all(X, k1);
        phi(X, k1, 0.5 * h, Wst);
        rhs_all(Wst, k2);
        State Y = X;
        phi(X, k2, h, Y);
        X = Y;
      }
      double j = 0.0;
      for (size_t c = 0; c < N; ++c) j += w[c] * X.a[0][c];
      return j;
    };
    std::vector<State> Ls;
    std::vector<double> hs;
    forward(S, &Ls, &hs);
    // reverse sweep
    std::vector<double> lam(8 * N, 0.0), lamN(8 * N), lamW(8 * N), lamk(8 * N), k1(8 * N), k2(8 * N);
    for (size_t c = 0; c < N; ++c) lam[c] = w[c];
    auto rhs_adj = [&](State& X, const std::vector<double>& lk, std::vector<double>& out) {
      MhdSet v = X.view();
      std::vector<double> gx_(ADJ_NCOEF * N), gy_(ADJ_NCOEF * N);
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
          const size_t c = (size_t)i * ny + j;
          double l[8], own[8], fx[ADJ_NCOEF], fy[ADJ_NCOEF];
          for (int q = 0; q < 8; ++q) l[q] = lk[q * N + c];
          adj_cell_coeffs(v, st, K, M, i, j, l, own, fx, fy);
          for (int q = 0; q < ADJ_NCOEF; ++q) { gx_[q * N + c] = fx[q]; gy_[q * N + c] = fy[q]; }
          for (int q = 0; q < 8; ++q) out[q * N + c] += own[q];
        }
      auto addf = [&](size_t cell, int var, double val) { out[var * N + cell] += val; };
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
          const size_t cu = (size_t)i * ny + j;
          double fa[ADJ_NCOEF];
          const size_t clx = (size_t)((i + nx - 1) % nx) * ny + j;
          for (int q = 0; q < ADJ_NCOEF; ++q) fa[q] = gx_[q * N + clx] - gx_[q * N + cu];
          adj_face<0>(v, st, K, M, i, j, fa, addf);
          const size_t cly = (size_t)i * ny + (j + ny - 1) % ny;
          for (int q = 0; q < ADJ_NCOEF; ++q) fa[q] = gy_[q * N + cly] - gy_[q * N + cu];
          adj_face<1>(v, st, K, M, i, j, fa, addf);
        }
    };
    auto phi_adj = [&](const State& B, const std::vector<double>& k, double h, double mult,
                       const std::vector<double>& lT, std::vector<double>& lB, std::vector<double>& lk) {
      double dot = 0.0;
      for (size_t c = 0; c < N; ++c) {
        const double hh = mult * h;
        const double x = (B.a[0][c] + k[c] * hh) / K.m;
        const double rho1 = smax(x, K.nmin) * K.m;
        const double drho = (x < K.nmin || rho1 / K.m < K.nmin) ? 0.0 : 1.0;
        const double dte = (B.a[5][c] + k[4 * N + c] * hh) < K.temin ? 0.0 : 1.0;
        for (int q = 0; q < 8; ++q) {
          const double mu = lT[q * N + c] * (q == 0 ? drho : (q == 4 ? dte : 1.0));
          lB[q * N + c] += mu;
          lk[q * N + c] = hh * mu;
          dot += mult * mu * k[q * N + c];
        }
      }
      return dot;
    };
    for (int n = nsteps - 1; n >= 0; --n) {
      State& Ln = Ls[n];
      const double h = hs[n];
      std::fill(lamN.begin(), lamN.end(), 0.0);
      rhs_all(Ln, k1);
      State Wst = Ln;
      phi(Ln, k1, 0.5 * h, Wst);
      rhs_all(Wst, k2);
      double gh = phi_adj(Ln, k2, h, 1.0, lam, lamN, lamk);
      std::fill(lamW.begin(), lamW.end(), 0.0);
      rhs_adj(Wst, lamk, lamW);
      gh += phi_adj(Ln, k1, h, 0.5, lamW, lamN, lamk);
      rhs_adj(Ln, lamk, lamN);
      int arg = -1;
      hstep(Ln, &arg);
      const Dual8 dt = dt_cell_dual(Ln.view(), st, K, M, arg / ny, arg % ny);
      for (int q = 0; q < 8; ++q) lamN[q * N + arg] += K.eps * gh * dt.d[q];
      lam.swap(lamN);
    }
    double wstep = 0.0;
    for (int trial = 0; trial < 4; ++trial) {
      std::vector<double> dir(8 * N);
      for (int q = 0; q < 8; ++q)
        for (size_t c = 0; c < N; ++c) dir[q * N + c] = U(rng) * scale[q];
      double ad = 0.0;
      for (size_t k = 0; k < 8 * N; ++k) ad += lam[k] * dir[k];
      double rel = 1.0;
      for (double eps : {1e-6, 1e-5}) {
        State P = S, Q = S;
        for (int q = 0; q < 8; ++q)
          for (size_t c = 0; c < N; ++c) { P.a[SLOT[q]][c] += eps * dir[q * N + c]; Q.a[SLOT[q]][c] -= eps * dir[q * N + c]; }
        const double fd = (forward(P, nullptr, nullptr) - forward(Q, nullptr, nullptr)) / (2 * eps);
        rel = std::fmin(rel, std::fabs(ad - fd) / (std::fabs(ad) + std::fabs(fd) + 1e-300));
      }
      std::printf("rk2_steps_dir%d %.3e\n", trial, rel);
      wstep = std::fmax(wstep, rel);
    }
    worst = std::fmax(worst, wstep);
  }
  std::printf("worst %.3e\n", worst);
  return worst < 1e-6 ? 0 : 1;
}
