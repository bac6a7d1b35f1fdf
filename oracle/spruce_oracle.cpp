// CPU FP64 restatement of the SPRUCE time step -- TEST INFRASTRUCTURE ONLY (the checker of the
// GPU path; see spruce_oracle.h).  Parity pins: tests/golden/*.npz, produced by the compiled
// reference, must be reproduced bit-for-bit (tests/test_oracle_golden.py).  ParametricHeating
// (spruce_oracle_modules.inc) has no reference counterpart: its restatement is its specification,
// so it has no reference pin.
//
// Every whole-array operation below is one separately rounded elementwise pass, evaluated left to
// right in the reference's expression order (Grid operators, source/mhd/grid.cpp:528-582); no
// fused multiply-add is used (the reference Makefile builds baseline x86-64, Makefile:34), and
// min/max follow std::min/std::max, i.e. NaN survives only as the first operand
// (grid.cpp:110-113, 161-163).
#include "spruce_oracle.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

static thread_local std::string g_err;

constexpr double K_B = 1.3807e-16;        // source/constants.hpp:8
constexpr double M_ELECTRON = 9.1094e-28; // constants.hpp:9
constexpr double KAPPA_0 = 1.0e-6;        // constants.hpp:14
constexpr double PI = 3.14159265358979323846;  // constants.hpp:15
constexpr int N_GHOST = 2;                // constants.hpp:4

struct fail : std::runtime_error { using std::runtime_error::runtime_error; };
static void require(bool c, const std::string& msg) { if (!c) throw fail(msg); }

// ---------------------------------------------------------------------------------------------
// Array type: row-major (x outer, y contiguous) like Grid (grid.cpp:516-526).
// ---------------------------------------------------------------------------------------------
struct A {
  int r = 1, c = 1;
  std::vector<double> v = std::vector<double>(1, 0.0);
  A() = default;
  A(int r_, int c_, double x = 0.0) : r(r_), c(c_), v((size_t)r_ * c_, x) {}
  double& operator()(int i, int j) { return v[(size_t)i * c + j]; }
  double operator()(int i, int j) const { return v[(size_t)i * c + j]; }
  size_t size() const { return v.size(); }
};

template <class F> static A map1(const A& a, F f) {
  A o(a.r, a.c);
  const size_t n = a.size();
#pragma omp parallel for if (n > 65536)
  for (size_t k = 0; k < n; ++k) o.v[k] = f(a.v[k]);
  return o;
}
template <class F> static A map2(const A& a, const A& b, F f) {
  require(a.r == b.r && a.c == b.c, "grid shape mismatch");
  A o(a.r, a.c);
  const size_t n = a.size();
#pragma omp parallel for if (n > 65536)
  for (size_t k = 0; k < n; ++k) o.v[k] = f(a.v[k], b.v[k]);
  return o;
}
static A operator+(const A& a, const A& b) { return map2(a, b, [](double x, double y) { return x + y; }); }
static A operator-(const A& a, const A& b) { return map2(a, b, [](double x, double y) { return x - y; }); }
static A operator*(const A& a, const A& b) { return map2(a, b, [](double x, double y) { return x * y; }); }
static A operator/(const A& a, const A& b) { return map2(a, b, [](double x, double y) { return x / y; }); }
static A operator+(const A& a, double s) { return map1(a, [s](double x) { return x + s; }); }
static A operator+(double s, const A& a) { return a + s; }
static A operator-(const A& a, double s) { return map1(a, [s](double x) { return x - s; }); }
static A operator-(double s, const A& a) { return map1(a, [s](double x) { return s - x; }); }
static A operator*(const A& a, double s) { return map1(a, [s](double x) { return x * s; }); }
static A operator*(double s, const A& a) { return a * s; }
static A operator/(const A& a, double s) { return map1(a, [s](double x) { return x / s; }); }
static A operator/(double s, const A& a) { return map1(a, [s](double x) { return s / x; }); }
static A operator-(const A& a) { return -1.0 * a; }
static A& operator+=(A& a, const A& b) { a = a + b; return a; }
static A& operator*=(A& a, const A& b) { a = a * b; return a; }
static A vmax(const A& a, double s) { return map1(a, [s](double x) { return (x < s) ? s : x; }); }
static A vmin(const A& a, double s) { return map1(a, [s](double x) { return (s < x) ? s : x; }); }
static A vmax(const A& a, const A& b) { return map2(a, b, [](double x, double y) { return (x < y) ? y : x; }); }
static A vmin(const A& a, const A& b) { return map2(a, b, [](double x, double y) { return (y < x) ? y : x; }); }
static A vsqrt(const A& a) { return map1(a, [](double x) { return std::sqrt(x); }); }
static A vabs(const A& a) { return map1(a, [](double x) { return std::fabs(x); }); }
static A vpow(const A& a, double p) { return map1(a, [p](double x) { return std::pow(x, p); }); }
static A sq(const A& a) { return a * a; }

// Range minimum with std::min semantics; the reference asserts the result is below DBL_MAX
// (grid.cpp:71-82).
static double range_min(const A& a, int il, int jl, int iu, int ju) {
  double m = DBL_MAX;
  for (int i = il; i <= iu; ++i)
    for (int j = jl; j <= ju; ++j) { double x = a(i, j); m = (x < m) ? x : m; }
  require(m < DBL_MAX, "min reduction found no finite value");
  return m;
}
static double whole_min(const A& a) { return range_min(a, 0, 0, a.r - 1, a.c - 1); }
static double whole_max(const A& a) {  // grid.cpp:115-123
  double m = -DBL_MAX;
  for (double x : a.v) m = (m < x) ? x : m;
  require(m > -DBL_MAX, "max reduction found no finite value");
  return m;
}

enum BC { Periodic = 0, Open = 1, Fixed = 2, Reflect = 3, OpenMoC = 4, OpenUCNP = 5 };
enum Integ { Euler = 0, RK2 = 1, RK4 = 2 };
enum PDG { d_x = 0, d_y, pos_x, pos_y, be_x, be_y, be_z };
static const char* PD_NAMES[7] = {"d_x", "d_y", "pos_x", "pos_y", "be_x", "be_y", "be_z"};

struct Domain;
using Grids = std::vector<A>;

// ---------------------------------------------------------------------------------------------
// EquationSet and Module interfaces (equationset.hpp:18-171; module.hpp:15-58)
// ---------------------------------------------------------------------------------------------
struct EqSet {
  Domain& pd;
  std::vector<std::string> names;
  Grids g;
  explicit EqSet(Domain& d, std::vector<std::string> n) : pd(d), names(std::move(n)), g(names.size()) {}
  virtual ~EqSet() = default;
  int idx(const std::string& n) const {
    for (size_t i = 0; i < names.size(); ++i) if (names[i] == n) return (int)i;
    throw fail("variable <" + n + "> not recognized");
  }
  bool has(const std::string& n) const { for (auto& s : names) if (s == n) return true; return false; }
  virtual std::vector<int> state_vars() const = 0;
  virtual std::vector<int> evolved() const = 0;
  virtual std::vector<int> densities() const = 0;
  virtual std::vector<int> thermal_energies() const = 0;
  virtual std::vector<std::vector<int>> momenta() const = 0;
  virtual std::vector<int> fields() const = 0;
  virtual int num_species() const { return (int)densities().size(); }
  virtual void parse(const std::vector<std::string>& k, const std::vector<std::string>& v) = 0;
  virtual Grids rhs(const Grids& grids) const = 0;
  virtual void evolved_from_state(Grids& grids) const = 0;
  virtual void derived(Grids& grids) const = 0;
  virtual void dt(Grids& grids) const = 0;
  virtual void minimums(Grids& grids) const = 0;
  virtual void setup_derived() {}

  Grids time_derivs(const Grids& grids) const;        // equationset.cpp:204-210
  void propagate(Grids& grids) const;                 // equationset.cpp:212-220
  void apply(Grids& grids, const Grids& k, double step) const;  // equationset.cpp:222-230
  void populate(Grids& grids) const;                  // equationset.cpp:96-104
  int evolved_index(const std::string& n) const {
    auto ev = evolved();
    for (size_t i = 0; i < ev.size(); ++i) if (names[ev[i]] == n) return (int)i;
    throw fail("<" + n + "> is not an evolved variable");
  }
};

struct Module {
  Domain& pd;
  explicit Module(Domain& d) : pd(d) {}
  virtual ~Module() = default;
  virtual void parse(const std::vector<std::string>& k, const std::vector<std::string>& v) = 0;
  virtual void setup() {}
  virtual void pre(double) {}
  virtual void iterate(double) {}
  virtual void post(double) {}
  virtual void rhs_add(const Grids&, Grids&) {}
  virtual void pre_recompute(Grids&) const {}
  virtual int subcycles() const { return -1; }
};

// ---------------------------------------------------------------------------------------------
// Domain: bounds, ghost fill, differential operators, time step
// ---------------------------------------------------------------------------------------------
struct Domain {
  int nx = 0, ny = 0;
  int bc[4] = {0, 0, 0, 0};
  int integ = Euler;
  double eps = 0, nmin = 0, tmin = 0, temin = 0, obs = 0, obdb = 1, m_ion = 0, gamma = 0;
  int xl = 0, xu = 0, yl = 0, yu = 0, xl_dt = 0, xu_dt = 0, yl_dt = 0, yu_dt = 0;
  A mask;
  A pd[7];
  bool pd_set[7] = {false, false, false, false, false, false, false};
  std::unique_ptr<EqSet> eqs;
  std::vector<std::unique_ptr<Module>> mods;
  double time = 0.0;
  int iter = 0;

  bool xper() const { return bc[0] == Periodic && bc[1] == Periodic; }
  bool yper() const { return bc[2] == Periodic && bc[3] == Periodic; }

  // plasmadomain.cpp:138-161
  void iteration_bounds() {
    require(nx > 2 * N_GHOST && ny > 2 * N_GHOST, "grid too small for ghost zones");
    xl = bc[0] == Periodic ? 0 : N_GHOST;
    xu = bc[1] == Periodic ? nx - 1 : nx - N_GHOST - 1;
    yl = bc[2] == Periodic ? 0 : N_GHOST;
    yu = bc[3] == Periodic ? ny - 1 : ny - N_GHOST - 1;
    require((bc[0] == Periodic) == (bc[1] == Periodic) && (bc[2] == Periodic) == (bc[3] == Periodic),
            "an axis must be periodic on both sides or on neither");
    mask = A(nx, ny, 0.0);
    for (int i = xl; i <= xu; ++i) for (int j = yl; j <= yu; ++j) mask(i, j) = 1.0;
    xl_dt = xl; xu_dt = xu; yl_dt = yl; yu_dt = yu;
    // open_moc sides evolve their ghost cells and take part in the step size (plasmadomain.cpp:155-159)
    if (bc[0] == OpenMoC) xl_dt -= N_GHOST;
    if (bc[1] == OpenMoC) xu_dt += N_GHOST;
    if (bc[2] == OpenMoC) yl_dt -= N_GHOST;
    if (bc[3] == OpenMoC) yu_dt += N_GHOST;
  }
  bool any_moc() const { return bc[0] == OpenMoC || bc[1] == OpenMoC || bc[2] == OpenMoC || bc[3] == OpenMoC; }

  // derivs.cpp:477-487: linear interpolation to the face between two cells
  double interp(const A& q, int i1, int j1, int i2, int j2) const {
    double a = q(i1, j1), b = q(i2, j2), da, db;
    if (i1 == i2) { da = 0.5 * pd[d_y](i1, j1); db = 0.5 * pd[d_y](i2, j2); }
    else { da = 0.5 * pd[d_x](i1, j1); db = 0.5 * pd[d_x](i2, j2); }
    return (a * db + b * da) / (db + da);
  }
  // derivs.cpp:490-499: linear extrapolation through two cells to the next face
  double extrap(const A& q, int i1, int j1, int i2, int j2) const {
    double a = q(i1, j1), b = q(i2, j2), da, db;
    if (i1 == i2) { da = 0.5 * pd[d_y](i1, j1); db = 0.5 * pd[d_y](i2, j2); }
    else { da = 0.5 * pd[d_x](i1, j1); db = 0.5 * pd[d_x](i2, j2); }
    return a + (b - a) * (da + 2.0 * db) / (da + db);
  }

  // Barton upwind face values, derivs.cpp:10-73.  Face (i,j) lies between (i-1,j) and (i,j)
  // for axis 0, between (i,j-1) and (i,j) for axis 1.
  A upwind(const A& q, const A& vel, int axis, int il, int jl, int iu, int ju) const {
    A s(nx + 1 - axis, ny + axis, 0.0);
    const bool per = axis == 0 ? xper() : yper();
    const int n = axis == 0 ? nx : ny;
#pragma omp parallel for collapse(2) if (nx * ny > 65536)
    for (int i = il; i <= iu + 1 - axis; ++i) {
      for (int j = jl; j <= ju + axis; ++j) {
        int c = axis == 0 ? i : j;  // index of the cell above the face along the axis
        int m2 = c - 2, m1 = c - 1, p1 = c + 1;
        if (per) {
          if (c == n) continue;
          m2 = (m2 + n) % n; m1 = (m1 + n) % n; p1 = (p1 + n) % n;
        }
        auto at = [&](int k, int& ii, int& jj) { if (axis == 0) { ii = k; jj = j; } else { ii = i; jj = k; } };
        int i0, j0, i1, j1, i2, j2, i3, j3;
        at(m2, i0, j0); at(m1, i1, j1); at(c, i2, j2); at(p1, i3, j3);
        double d2 = interp(q, i1, j1, i2, j2);
        double vf = interp(vel, i1, j1, i2, j2);
        double out;
        if (vf > 0.0) {
          double d3 = q(i1, j1), d1 = extrap(q, i0, j0, i1, j1);
          if (q(i2, j2) <= q(i1, j1)) out = std::min(d3, std::max(d1, d2));
          else out = std::max(d3, std::min(d1, d2));
        } else if (vf < 0.0) {
          double d3 = q(i2, j2), d1 = extrap(q, i3, j3, i2, j2);
          if (q(i2, j2) <= q(i1, j1)) out = std::max(d3, std::min(d1, d2));
          else out = std::min(d3, std::max(d1, d2));
        } else {
          out = d2;
        }
        s(i2, j2) = out;
      }
    }
    return s;
  }

  // derivs.cpp:122-162 (multiply_vel = true)
  A tderiv(const A& q, const A& vel, int axis) const {
    A s = upwind(q, vel, axis, xl, yl, xu, yu);
    A out(nx, ny, 0.0);
    const bool per = axis == 0 ? xper() : yper();
    const int n = axis == 0 ? nx : ny;
    // denom = d_x*(1-axis) + d_y*axis, built as whole-grid operations (derivs.cpp:128)
    A denom = pd[d_x] * (double)(1 - axis) + pd[d_y] * (double)axis;
#pragma omp parallel for collapse(2) if (nx * ny > 65536)
    for (int i = xl; i <= xu; ++i) {
      for (int j = yl; j <= yu; ++j) {
        int c = axis == 0 ? i : j, m = c - 1, p = c + 1;
        if (per) { m = (m + n) % n; p = (p + n) % n; }
        int im = axis == 0 ? m : i, jm = axis == 0 ? j : m, ip = axis == 0 ? p : i, jp = axis == 0 ? j : p;
        out(i, j) = (s(ip, jp) * interp(vel, i, j, ip, jp) - s(i, j) * interp(vel, im, jm, i, j)) / denom(i, j);
      }
    }
    return out;
  }
  A tdiv(const A& q, const A& vx, const A& vy) const { return tderiv(q, vx, 0) + tderiv(q, vy, 1); }  // derivs.cpp:216-220

  // derivs.cpp:223-264: central derivative through face interpolants
  A deriv(const A& q, int axis) const {
    A out(nx, ny, 0.0);
    const bool per = axis == 0 ? xper() : yper();
    const int n = axis == 0 ? nx : ny;
#pragma omp parallel for collapse(2) if (nx * ny > 65536)
    for (int i = xl; i <= xu; ++i) {
      for (int j = yl; j <= yu; ++j) {
        int c = axis == 0 ? i : j, m = c - 1, p = c + 1;
        if (per) { m = (m + n) % n; p = (p + n) % n; }
        int im = axis == 0 ? m : i, jm = axis == 0 ? j : m, ip = axis == 0 ? p : i, jp = axis == 0 ? j : p;
        double den = axis == 0 ? pd[d_x](i, j) : pd[d_y](i, j);
        out(i, j) = (interp(q, i, j, ip, jp) - interp(q, im, jm, i, j)) / den;
      }
    }
    return out;
  }
  // derivs.cpp:417-455
  A second(const A& q, int axis) const {
    A out(nx, ny, 0.0);
    A denom = sq(0.5 * pd[d_x] * (double)(1 - axis) + 0.5 * pd[d_y] * (double)axis);
    const bool per = axis == 0 ? xper() : yper();
    const int n = axis == 0 ? nx : ny;
#pragma omp parallel for collapse(2) if (nx * ny > 65536)
    for (int i = xl; i <= xu; ++i) {
      for (int j = yl; j <= yu; ++j) {
        int c = axis == 0 ? i : j, m = c - 1, p = c + 1;
        if (per) { m = (m + n) % n; p = (p + n) % n; }
        int im = axis == 0 ? m : i, jm = axis == 0 ? j : m, ip = axis == 0 ? p : i, jp = axis == 0 ? j : p;
        out(i, j) = (interp(q, i, j, ip, jp) - 2.0 * q(i, j) + interp(q, im, jm, i, j)) / denom(i, j);
      }
    }
    return out;
  }
  A laplacian(const A& q) const { A rx = second(q, 0); A ry = second(q, 1); return rx + ry; }   // derivs.cpp:458-462
  std::vector<A> curlZ(const A& z) const { A rx = deriv(z, 1); A ry = -deriv(z, 0); return {rx, ry}; }  // 465-469
  A curl2D(const A& x, const A& y) const { return deriv(y, 0) - deriv(x, 1); }  // 472-474
  A div2D(const A& x, const A& y) const { return deriv(x, 0) + deriv(y, 1); }   // 407-414

  // ----- the same operators over explicit ranges, and the one-sided operators of the method of
  // characteristics (derivs.cpp:77-117, 167-214, 223-264, 317-404, 417-455) -----
  struct R4 { int il, jl, iu, ju; };
  // neighbour c+off along the axis, wrapped only on a periodic axis
  int nb(int c, int off, int axis) const {
    const bool per = axis == 0 ? xper() : yper();
    const int n = axis == 0 ? nx : ny;
    return per ? (c + off + n) % n : c + off;
  }
  A deriv_r(const A& q, int axis, R4 r) const {  // derivative1D
    A out(nx, ny, 0.0);
    for (int i = r.il; i <= r.iu; ++i)
      for (int j = r.jl; j <= r.ju; ++j) {
        const int c = axis == 0 ? i : j, m = nb(c, -1, axis), p = nb(c, 1, axis);
        const int im = axis == 0 ? m : i, jm = axis == 0 ? j : m, ip = axis == 0 ? p : i, jp = axis == 0 ? j : p;
        const double den = axis == 0 ? pd[d_x](i, j) : pd[d_y](i, j);
        out(i, j) = (interp(q, i, j, ip, jp) - interp(q, im, jm, i, j)) / den;
      }
    return out;
  }
  A second_r(const A& q, int axis, R4 r) const {  // secondDerivative1D
    A out(nx, ny, 0.0);
    A denom = sq(0.5 * pd[d_x] * (double)(1 - axis) + 0.5 * pd[d_y] * (double)axis);
    for (int i = r.il; i <= r.iu; ++i)
      for (int j = r.jl; j <= r.ju; ++j) {
        const int c = axis == 0 ? i : j, m = nb(c, -1, axis), p = nb(c, 1, axis);
        const int im = axis == 0 ? m : i, jm = axis == 0 ? j : m, ip = axis == 0 ? p : i, jp = axis == 0 ? j : p;
        out(i, j) = (interp(q, i, j, ip, jp) - 2.0 * q(i, j) + interp(q, im, jm, i, j)) / denom(i, j);
      }
    return out;
  }
  A tderiv_r(const A& q, const A& vel, int axis, R4 r) const {  // transportDerivative1D
    A sf = upwind(q, vel, axis, r.il, r.jl, r.iu, r.ju);
    A out(nx, ny, 0.0);
    A denom = pd[d_x] * (double)(1 - axis) + pd[d_y] * (double)axis;
    for (int i = r.il; i <= r.iu; ++i)
      for (int j = r.jl; j <= r.ju; ++j) {
        const int c = axis == 0 ? i : j, m = nb(c, -1, axis), p = nb(c, 1, axis);
        const int im = axis == 0 ? m : i, jm = axis == 0 ? j : m, ip = axis == 0 ? p : i, jp = axis == 0 ? j : p;
        out(i, j) = (sf(ip, jp) * interp(vel, i, j, ip, jp) - sf(i, j) * interp(vel, im, jm, i, j)) / denom(i, j);
      }
    return out;
  }
  // upwindSurfaceDirected: Barton face values with the upwind side fixed
  A upwind_dir(const A& q, bool vpos, int axis, R4 r) const {
    A sf(nx + 1 - axis, ny + axis, 0.0);
    const int sg = vpos ? 1 : -1;
    for (int i = r.il; i <= r.iu; ++i)
      for (int j = r.jl; j <= r.ju; ++j) {
        int i0 = i, i1 = i, j0 = j, j1 = j;
        if (axis == 0) { i0 = i - 2 * sg; i1 = i - sg; } else { j0 = j - 2 * sg; j1 = j - sg; }
        const double d2 = interp(q, i1, j1, i, j), d3 = q(i1, j1), d1 = extrap(q, i0, j0, i1, j1);
        const int is = i + ((axis == 0 && !vpos) ? 1 : 0), js = j + ((axis == 1 && !vpos) ? 1 : 0);
        if (q(i, j) <= q(i1, j1)) sf(is, js) = std::min(d3, std::max(d1, d2));
        else sf(is, js) = std::max(d3, std::min(d1, d2));
      }
    return sf;
  }
  // characteristicBartonDerivative1D
  A char_barton(const A& q, bool pf, int axis, R4 r) const {
    A sf = upwind_dir(q, pf, axis, r);
    if (axis == 0) {
      for (int j = r.jl; j <= r.ju; ++j) {
        if (pf) sf(sf.r - 1, j) = q(q.r - 1, j) + (q(q.r - 1, j) - sf(sf.r - 2, j));
        else sf(0, j) = q(0, j) - (sf(1, j) - q(0, j));
      }
    } else {
      for (int i = r.il; i <= r.iu; ++i) {
        if (pf) sf(i, sf.c - 1) = q(i, q.c - 1) + (q(i, q.c - 1) - sf(i, sf.c - 2));
        else sf(i, 0) = q(i, 0) - (sf(i, 1) - q(i, 0));
      }
    }
    A denom = pd[d_x] * (double)(1 - axis) + pd[d_y] * (double)axis;
    A out(nx, ny, 0.0);
    for (int i = r.il; i <= r.iu; ++i)
      for (int j = r.jl; j <= r.ju; ++j) {
        const int i2 = axis == 0 ? i + 1 : i, j2 = axis == 0 ? j : j + 1;
        out(i, j) = (sf(i2, j2) - sf(i, j)) / denom(i, j);
      }
    return out;
  }
  // derivative1DBackward (first order, one-sided)
  A back1(const A& q, bool pf, int axis, R4 r) const {
    A out(nx, ny, 0.0);
    for (int i = r.il; i <= r.iu; ++i)
      for (int j = r.jl; j <= r.ju; ++j) {
        const int c = axis == 0 ? i : j, o = nb(c, pf ? -1 : 1, axis);
        const int i0 = axis == 0 ? o : i, j0 = axis == 0 ? j : o;
        const A& dd = axis == 0 ? pd[d_x] : pd[d_y];
        const double den = 0.5 * (dd(i, j) + dd(i0, j0));
        out(i, j) = (pf ? 1.0 : -1.0) * (q(i, j) - q(i0, j0)) / den;
      }
    return out;
  }
  // secondDerivative1DBackward
  A second_back(const A& q, bool pf, int axis, R4 r) const {
    A out(nx, ny, 0.0);
    for (int i = r.il; i <= r.iu; ++i)
      for (int j = r.jl; j <= r.ju; ++j) {
        const int c = axis == 0 ? i : j;
        const int c1 = nb(c, pf ? -1 : 1, axis), c0 = nb(c1, pf ? -1 : 1, axis);
        const int i1 = axis == 0 ? c1 : i, j1 = axis == 0 ? j : c1, i0 = axis == 0 ? c0 : i, j0 = axis == 0 ? j : c0;
        const A& dd = axis == 0 ? pd[d_x] : pd[d_y];
        const double den01 = 0.5 * (dd(i1, j1) + dd(i0, j0));
        const double den12 = 0.5 * (dd(i1, j1) + dd(i, j));
        const double t12 = q(std::max(i1, i), std::max(j1, j)) - q(std::min(i1, i), std::min(j1, j));
        const double t01 = q(std::max(i1, i0), std::max(j1, j0)) - q(std::min(i1, i0), std::min(j1, j0));
        out(i, j) = (pf ? 1.0 : -1.0) * (t12 / den12 - t01 / den01) / den12;
      }
    return out;
  }

  // ----- ghost-cell fill, evolution.cpp:126-334 -----
  void open_fill(Grids& grids, int i1, int i2, int i3, int j1, int j2, int j3) const {  // 158-224
    (void)grids;
    Grids& L = eqs->g;  // writes the live grids (evolution.cpp:169-223)
    const bool xb = (j1 == j2);
    const A& dd = xb ? pd[d_x] : pd[d_y];
    double delta_last = dd(i3, j3);
    double dist23 = 0.5 * (dd(i2, j2) + dd(i3, j3));
    double dist12 = 0.5 * (dd(i1, j1) + dd(i2, j2));
    double s2 = std::pow(obdb, dist23 / delta_last);
    double s1 = std::pow(obdb, dist12 / delta_last);
    for (int v : eqs->densities()) { L[v](i1, j1) = s1 * L[v](i3, j3); L[v](i2, j2) = s2 * L[v](i3, j3); }
    for (int v : eqs->thermal_energies()) { L[v](i1, j1) = s1 * L[v](i3, j3); L[v](i2, j2) = s2 * L[v](i3, j3); }
    double cs = 0.0;
    const int ns = eqs->num_species();
    for (int s = 0; s < ns; ++s) {
      double p3 = L[eqs->thermal_energies()[s]](i3, j3) * (gamma - 1.0);
      double r3 = L[eqs->densities()[s]](i3, j3);
      double c = std::sqrt(gamma * p3 / r3);
      if (c > cs) cs = c;
    }
    for (int s = 0; s < ns; ++s) {
      A& mx = L[eqs->momenta()[s][0]];
      A& my = L[eqs->momenta()[s][1]];
      A& rho = L[eqs->densities()[s]];
      double vx = mx(i3, j3) / rho(i3, j3);
      double vy = my(i3, j3) / rho(i3, j3);
      double boost = obs * cs;
      if (i2 > i1 || j2 > j1) boost *= -1.0;
      double dist = dist23;
      if (xb) {
        double bv = (i1 > i2) ? std::max(0.0, vx + boost) : std::min(0.0, vx + boost);
        double gv = (dist * bv - 0.5 * dd(i2, j2) * vx) / (0.5 * dd(i3, j3));
        mx(i1, j1) = rho(i1, j1) * gv; mx(i2, j2) = rho(i2, j2) * gv;
        my(i1, j1) = rho(i1, j1) * vy; my(i2, j2) = rho(i2, j2) * vy;
      } else {
        double bv = (j1 > j2) ? std::max(0.0, vy + boost) : std::min(0.0, vy + boost);
        double gv = (dist * bv - 0.5 * dd(i2, j2) * vy) / (0.5 * dd(i3, j3));
        mx(i1, j1) = rho(i1, j1) * vx; mx(i2, j2) = rho(i2, j2) * vx;
        my(i1, j1) = rho(i1, j1) * gv; my(i2, j2) = rho(i2, j2) * gv;
      }
    }
  }
  void reflect_fill(int i1, int i2, int i3, int j1, int j2, int j3) const {  // 231-266
    Grids& L = eqs->g;
    for (int v : eqs->thermal_energies()) { L[v](i1, j1) = L[v](i3, j3); L[v](i2, j2) = L[v](i3, j3); }
    for (int v : eqs->densities()) { L[v](i1, j1) = L[v](i3, j3); L[v](i2, j2) = L[v](i3, j3); }
    for (auto& mom : eqs->momenta())
      for (int v : mom) { L[v](i1, j1) = 0.0; L[v](i2, j2) = 0.0; L[v](i3, j3) = 0.0; }
  }
  void fixed_fill(int i1, int i2, int i3, int j1, int j2, int j3) const {  // 272-282
    Grids& L = eqs->g;
    for (auto& mom : eqs->momenta())
      for (int v : mom) { L[v](i1, j1) = 0.0; L[v](i2, j2) = 0.0; L[v](i3, j3) = 0.0; }
  }
  void ucnp_fill(Grids& grids, int side, int index) const {  // 290-334 (writes `grids`)
    for (int k = 0; k < N_GHOST; ++k) {
      int ix, iy, bx, by;
      switch (side) {
        case 0: ix = xl; iy = index; bx = k; by = index; break;
        case 1: ix = xu; iy = index; bx = xu + 1 + k; by = index; break;
        case 2: ix = index; iy = yl; bx = index; by = k; break;
        default: ix = index; iy = yu; bx = index; by = yu + 1 + k; break;
      }
      for (int v : eqs->densities()) grids[v](bx, by) = grids[v](ix, iy);
      for (int v : eqs->thermal_energies()) grids[v](bx, by) = grids[v](ix, iy);
      for (int v : eqs->fields()) grids[v](bx, by) = grids[v](ix, iy);
      const auto moms = eqs->momenta();
      for (int s = 0; s < eqs->num_species(); ++s)
        for (int v : moms[s]) grids[v](bx, by) = grids[v](ix, iy);
    }
  }
  void ghost(Grids& grids) const {  // updateGhostZones, evolution.cpp:126-152
    auto side = [&](int b, bool xaxis, bool upper) {
      int kind = bc[b];
      if (kind == Periodic || kind == OpenMoC) return;
      const int n = xaxis ? nx : ny;
      const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
      int lo = xaxis ? yl : xl, hi = xaxis ? yu : xu;
      if (kind == Fixed) { lo = 0; hi = (xaxis ? ny : nx) - 1; }
      for (int t = lo; t <= hi; ++t) {
        int i1, i2, i3, j1, j2, j3;
        if (xaxis) { i1 = e1; i2 = e2; i3 = e3; j1 = j2 = j3 = t; }
        else { j1 = e1; j2 = e2; j3 = e3; i1 = i2 = i3 = t; }
        if (kind == Open) open_fill(grids, i1, i2, i3, j1, j2, j3);
        else if (kind == Reflect) reflect_fill(i1, i2, i3, j1, j2, j3);
        else if (kind == Fixed) fixed_fill(i1, i2, i3, j1, j2, j3);
        else if (kind == OpenUCNP) ucnp_fill(grids, b, t);
      }
    };
    side(0, true, false);
    side(1, true, true);
    side(2, false, false);
    side(3, false, true);
  }

  // ----- time step, evolution.cpp:59-124 -----
  void advance() {
    int dti = eqs->idx("dt");
    double step = eps * range_min(eqs->g[dti], xl_dt, yl_dt, xu_dt, yu_dt);
    for (auto& m : mods) m->pre(step);
    for (auto& m : mods) m->iterate(step);
    if (integ == RK2) rk2(step);
    else if (integ == RK4) rk4(step);
    else euler(step);
    for (auto& m : mods) m->post(step);
    last_dt = step;
    time += step;
    iter++;
  }
  double last_dt = 0.0;
  void euler(double h) { Grids k = eqs->time_derivs(eqs->g); eqs->apply(eqs->g, k, h); }
  void rk2(double h) {
    Grids k = eqs->time_derivs(eqs->g);
    Grids mid = eqs->g;
    eqs->apply(mid, k, 0.5 * h);
    k = eqs->time_derivs(mid);
    eqs->apply(eqs->g, k, h);
  }
  void rk4(double h) {
    Grids k1 = eqs->time_derivs(eqs->g);
    Grids r = eqs->g;
    eqs->apply(r, k1, 0.5 * h);
    Grids k2 = eqs->time_derivs(r);
    r = eqs->g;
    eqs->apply(r, k2, 0.5 * h);
    Grids k3 = eqs->time_derivs(r);
    r = eqs->g;
    eqs->apply(r, k3, h);
    Grids k4 = eqs->time_derivs(r);
    Grids kf = k1;
    for (size_t i = 0; i < eqs->evolved().size(); ++i) kf[i] = (k1[i] + k4[i]) / 6.0 + (k2[i] + k3[i]) / 3.0;
    eqs->apply(eqs->g, kf, h);
  }
};

Grids EqSet::time_derivs(const Grids& grids) const {
  Grids k = rhs(grids);
  for (auto& m : pd.mods) m->rhs_add(grids, k);
  return k;
}
void EqSet::propagate(Grids& grids) const {
  minimums(grids);
  pd.ghost(grids);
  for (auto& m : pd.mods) m->pre_recompute(grids);
  derived(grids);
  dt(grids);
}
void EqSet::apply(Grids& grids, const Grids& k, double step) const {
  auto ev = evolved();
  for (size_t i = 0; i < ev.size(); ++i) grids[ev[i]] += step * k[i];
  propagate(grids);
}
void EqSet::populate(Grids& grids) const {
  evolved_from_state(grids);
  minimums(grids);
  pd.ghost(grids);
  derived(grids);
  dt(grids);
}

static double to_d(const std::string& s) { return std::stod(s); }

// ---------------------------------------------------------------------------------------------
// IdealMHD (source/equationsets/idealmhd.{hpp,cpp})
// ---------------------------------------------------------------------------------------------
struct IdealMHD : EqSet {
  enum V { rho, temp, mom_x, mom_y, mom_z, bi_x, bi_y, bi_z, grav_x, grav_y, n, press, thermal_energy,
           v_x, v_y, v_z, kinetic_energy, b_x, b_y, b_z, b_mag, b_hat_x, b_hat_y, b_hat_z, dt_ };
  double global_viscosity = 0.0;
  bool moc_b = false, moc_mom = false;
  double moc_b_lo = 0.1, moc_b_hi = 10.0, moc_mom_lo = 0.1, moc_mom_hi = 10.0;  // idealmhd.hpp:59-64
  explicit IdealMHD(Domain& d)
      : EqSet(d, {"rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z", "grav_x", "grav_y",
                  "n", "press", "thermal_energy", "v_x", "v_y", "v_z", "kinetic_energy",
                  "b_x", "b_y", "b_z", "b_mag", "b_hat_x", "b_hat_y", "b_hat_z", "dt"}) {}
  std::vector<int> state_vars() const override { return {rho, temp, mom_x, mom_y, mom_z, bi_x, bi_y, bi_z, grav_x, grav_y}; }
  std::vector<int> evolved() const override { return {rho, mom_x, mom_y, mom_z, thermal_energy, bi_x, bi_y, bi_z}; }
  std::vector<int> densities() const override { return {rho}; }
  std::vector<int> thermal_energies() const override { return {thermal_energy}; }
  std::vector<std::vector<int>> momenta() const override { return {{mom_x, mom_y, mom_z}}; }
  std::vector<int> fields() const override { return {bi_x, bi_y, bi_z}; }
  void parse(const std::vector<std::string>& k, const std::vector<std::string>& v) override {  // idealmhd.cpp:12-40
    for (size_t i = 0; i < k.size(); ++i) {
      if (k[i] == "global_viscosity") global_viscosity = to_d(v[i]);
      else if (k[i] == "viscosity_opt") {}
      else if (k[i] == "moc_b_limiting") moc_b = v[i] == "true";
      else if (k[i] == "moc_mom_limiting") moc_mom = v[i] == "true";
      else if (k[i] == "moc_b_lower_lim") { moc_b_lo = to_d(v[i]); require(moc_b_lo <= 1.0, "MoC B field lower threshold should be <=1.0"); }
      else if (k[i] == "moc_b_upper_lim") { moc_b_hi = to_d(v[i]); require(moc_b_hi >= 1.0, "MoC B field upper threshold should be >=1.0"); }
      else if (k[i] == "moc_mom_lower_lim") { moc_mom_lo = to_d(v[i]); require(moc_mom_lo <= 1.0, "MoC momentum threshold should be <=1.0"); }
      else if (k[i] == "moc_mom_upper_lim") { moc_mom_hi = to_d(v[i]); require(moc_mom_hi >= 1.0, "MoC momentum threshold should be >=1.0"); }
      else throw fail(k[i] + " is not recognized for this equation set.");
    }
  }
  // idealmhd.cpp:42-105
  Grids rhs(const Grids& G) const override {
    const Domain& D = pd;
    const A &bex = D.pd[be_x], &bey = D.pd[be_y], &bez = D.pd[be_z];
    const A &vx = G[v_x], &vy = G[v_y];
    A d_rho = -D.tdiv(G[rho], vx, vy);
    A curl_db = D.curl2D(G[bi_x], G[bi_y]) / (4.0 * PI);
    std::vector<A> ext = {-curl_db * bey, curl_db * bex};
    std::vector<A> in = {-curl_db * G[bi_y], curl_db * G[bi_x]};
    std::vector<A> cz = D.curlZ(G[bi_z]);
    // CrossProduct2DZ(a, b_z) = {b_z * a_y, -b_z * a_x} (grid.cpp:463-468)
    A bzi = G[bi_z] / (4.0 * PI), bze = bez / (4.0 * PI);
    std::vector<A> in_z = {-(-1.0 * bzi) * cz[1], (-1.0 * bzi) * cz[0]};
    std::vector<A> ext_z = {-(-1.0 * bze) * cz[1], (-1.0 * bze) * cz[0]};
    A fz_ext = (cz[0] * bey - cz[1] * bex) / (4.0 * PI);
    A fz_in = (cz[0] * G[bi_y] - cz[1] * G[bi_x]) / (4.0 * PI);
    A d_mx = -D.tdiv(G[mom_x], vx, vy) - D.deriv(G[press], 0) + G[rho] * G[grav_x] + ext[0] + in[0] + in_z[0] + ext_z[0];
    A d_my = -D.tdiv(G[mom_y], vx, vy) - D.deriv(G[press], 1) + G[rho] * G[grav_y] + ext[1] + in[1] + in_z[1] + ext_z[1];
    A d_mz = -D.tdiv(G[mom_z], vx, vy) + fz_ext + fz_in;
    A d_te = -D.tdiv(G[thermal_energy], vx, vy) - G[press] * D.div2D(vx, vy);
    A d_bx = -D.tdiv(G[bi_x], vx, vy) - D.tdiv(bex, vx, vy) + (G[bi_x] + bex) * D.deriv(G[v_x], 0) + (G[bi_y] + bey) * D.deriv(G[v_x], 1);
    A d_by = -D.tdiv(G[bi_y], vx, vy) - D.tdiv(bey, vx, vy) + (G[bi_x] + bex) * D.deriv(G[v_y], 0) + (G[bi_y] + bey) * D.deriv(G[v_y], 1);
    A d_bz = -D.tdiv(G[bi_z], vx, vy) - D.tdiv(bez, vx, vy) + (G[bi_x] + bex) * D.deriv(G[v_z], 0) + (G[bi_y] + bey) * D.deriv(G[v_z], 1);
    // global_visc_coeff (idealmhd.cpp:90) only feeds the characteristic-boundary terms, which are
    // identically zero without open_moc boundaries (idealmhd.cpp:308-332): rho*0 + v*0 for momenta.
    Grids ch;
    if (D.any_moc()) {
      A one(D.nx, D.ny, 1.0);
      const double gvc = global_viscosity * 0.5 *
          range_min(one / (one / sq(D.pd[d_x]) + one / sq(D.pd[d_y])) / G[dt_], D.xl_dt, D.yl_dt, D.xu_dt, D.yu_dt);
      ch = moc_boundary(G, gvc);
    } else {
      A zero(D.nx, D.ny, 0.0);
      ch.assign(8, zero);
      ch[1] = G[rho] * zero + G[v_x] * zero;
      ch[2] = G[rho] * zero + G[v_y] * zero;
      ch[3] = G[rho] * zero + G[v_z] * zero;
    }
    Grids k = {d_rho, d_mx, d_my, d_mz, d_te, d_bx, d_by, d_bz};
    for (int i = 0; i < 8; ++i) { k[i] *= D.mask; k[i] += ch[i]; }
    return k;
  }
  // computeTimeDerivativesCharacteristicBoundary (idealmhd.cpp:308-332): each open_moc side's
  // normal + parallel terms summed into zero grids in side order, then the momentum form
  Grids moc_boundary(const Grids& G, double vc) const {
    A zero(pd.nx, pd.ny, 0.0);
    Grids res(8, zero);
    const int sides[4][2] = {{0, 1}, {0, 0}, {1, 1}, {1, 0}};  // (axis, lower) of x_bound_1 .. y_bound_2
    for (int b = 0; b < 4; ++b) {
      if (pd.bc[b] != OpenMoC) continue;
      auto t = moc_side(G, sides[b][0], sides[b][1] != 0, vc);
      for (int i = 0; i < 8; ++i) res[i] += t[i][0] + t[i][1];
    }
    A mx = G[rho] * res[1] + G[v_x] * res[0];
    A my = G[rho] * res[2] + G[v_y] * res[0];
    A mz = G[rho] * res[3] + G[v_z] * res[0];
    res[1] = mx;
    res[2] = my;
    res[3] = mz;
    return res;
  }
  // singleBoundaryTermsMOC (idealmhd.cpp:336-615): {normal, parallel} terms of one open_moc side for
  // {rho, v_x, v_y, v_z, thermal_energy, b_x, b_y, b_z}
  std::vector<std::array<A, 2>> moc_side(const Grids& G, int bi, bool lower, double vc) const {
    const Domain& D = pd;
    using R4 = Domain::R4;
    const int NX = D.nx, NY = D.ny, NG = N_GHOST;
    int xi[2], yi[2];
    if (bi == 0) {
      xi[0] = lower ? 0 : NX - NG; xi[1] = lower ? NG - 1 : NX - 1;
      yi[0] = 0; yi[1] = NY - 1;
    } else {
      yi[0] = lower ? 0 : NY - NG; yi[1] = lower ? NG - 1 : NY - 1;
      xi[0] = 0; xi[1] = NX - 1;
    }
    int xin[2] = {xi[0], xi[1]}, yin[2] = {yi[0], yi[1]}, xpd[2] = {xi[0], xi[1]}, ypd[2] = {yi[0], yi[1]};
    int xlc[2] = {xi[0], xi[1]}, xuc[2] = {xi[0], xi[1]}, ylc[2] = {yi[0], yi[1]}, yuc[2] = {yi[0], yi[1]};
    if (bi == 0 && D.bc[2] != Periodic) {
      yin[0] = yi[0] + NG; yin[1] = yi[1] - NG;
      ypd[0] = yi[0] + 1; ypd[1] = yi[1] - 1;
      if (D.bc[2] != OpenMoC) { yi[0] = yin[0]; ypd[0] = yin[0]; }
      if (D.bc[3] != OpenMoC) { yi[1] = yin[1]; ypd[1] = yin[1]; }
      ylc[0] = yi[0]; ylc[1] = ypd[0] - 1;
      yuc[0] = ypd[1] + 1; yuc[1] = yi[1];
    } else if (bi == 1 && D.bc[0] != Periodic) {
      xin[0] = xi[0] + NG; xin[1] = xi[1] - NG;
      xpd[0] = xi[0] + 1; xpd[1] = xi[1] - 1;
      if (D.bc[0] != OpenMoC) { xi[0] = xin[0]; xpd[0] = xin[0]; }
      if (D.bc[1] != OpenMoC) { xi[1] = xin[1]; xpd[1] = xin[1]; }
      xlc[0] = xi[0]; xlc[1] = xpd[0] - 1;
      xuc[0] = xpd[1] + 1; xuc[1] = xi[1];
    }
    const R4 RB{xi[0], yi[0], xi[1], yi[1]}, RI{xin[0], yin[0], xin[1], yin[1]}, RP{xpd[0], ypd[0], xpd[1], ypd[1]};
    const R4 RU{xuc[0], yuc[0], xuc[1], yuc[1]}, RL{xlc[0], ylc[0], xlc[1], ylc[1]};
    A imask(NX, NY, 0.0);
    for (int i = RI.il; i <= RI.iu; ++i) for (int j = RI.jl; j <= RI.ju; ++j) imask(i, j) = 1.0;

    const int vperp = bi == 0 ? v_x : v_y, vpara = bi == 0 ? v_y : v_x, vgd = v_z;
    const A& be_perp = D.pd[bi == 0 ? be_x : be_y];
    const A& be_para = D.pd[bi == 0 ? be_y : be_x];
    const A& be_gd = D.pd[be_z];
    const A& bi_perp = G[bi == 0 ? bi_x : bi_y];
    const A& bi_para = G[bi == 0 ? bi_y : bi_x];
    const A& bi_gd = G[bi_z];
    const A& g_perp = G[bi == 0 ? grav_x : grav_y];
    const A& g_para = G[bi == 1 ? grav_x : grav_y];
    const bool pf = !lower;
    const A b_perp = be_perp + bi_perp, b_para = be_para + bi_para, b_gd = be_gd + bi_gd;
    const A bh = vsqrt(sq(b_para) + sq(b_gd));

    A s_perp(NX, NY, 0.0), R_para = s_perp, R_gd = s_perp, c_perp = s_perp, c_plus = s_perp, c_minus = s_perp;
    A al_p = s_perp, al_m = s_perp;
    A c_s(NX, NY, 1.0), c_s_sq(NX, NY, 1.0);
    for (int i = RB.il; i <= RB.iu; ++i)
      for (int j = RB.jl; j <= RB.ju; ++j) {
        s_perp(i, j) = b_perp(i, j) > 0.0 ? 1.0 : -1.0;
        R_para(i, j) = b_para(i, j) / bh(i, j);
        R_gd(i, j) = b_gd(i, j) / bh(i, j);
        const double r = G[rho](i, j);
        const double css = pd.gamma * G[press](i, j) / r;
        const double cas = G[b_mag](i, j) * G[b_mag](i, j) / (4.0 * PI * r);
        const double cps = b_perp(i, j) * b_perp(i, j) / (4.0 * PI * r);
        const double sum = cas + css;
        const double root = std::sqrt(std::fabs(sum * sum / 4.0 - cps * css));
        const double cpl = sum / 2.0 + root, cmi = sum / 2.0 - root;
        const double aps = std::max((css - cmi) / (cpl - cmi), 0.0), ams = std::max((cpl - css) / (cpl - cmi), 0.0);
        c_s_sq(i, j) = css;
        c_s(i, j) = std::sqrt(css);
        c_perp(i, j) = std::sqrt(cps);
        c_plus(i, j) = std::sqrt(cpl);
        c_minus(i, j) = std::sqrt(cmi);
        al_p(i, j) = std::sqrt(aps);
        al_m(i, j) = std::sqrt(ams);
      }
    for (size_t c = 0; c < s_perp.size(); ++c) if (b_perp.v[c] == 0.0) s_perp.v[c] = 0.0;

    auto cb = [&](const A& q) { return D.char_barton(q, pf, bi, RB); };
    const A b_para_grad = cb(be_para) + cb(bi_para);
    const A b_gd_grad = cb(be_gd) + cb(bi_gd);
    const A v_para_grad = cb(G[vpara]), v_perp_grad = cb(G[vperp]), v_gd_grad = cb(G[vgd]), p_grad = cb(G[press]);
    const A& vp = G[vperp];
    const A& RH = G[rho];
    const double sg = bi == 0 ? 1.0 : -1.0;
    const A sq4 = vsqrt(4.0 * PI * RH);
    A d1 = vp * cb(b_perp);
    A d2 = vp * ((G[thermal_energy] + G[press]) * cb(RH) - RH * cb(G[thermal_energy]));
    A d3 = s_perp * sg * (vp + c_perp) *
           (-s_perp * R_gd * v_para_grad + s_perp * R_para * v_gd_grad + R_gd / sq4 * b_para_grad - R_para / sq4 * b_gd_grad);
    A d4 = s_perp * sg * (vp - c_perp) *
           (s_perp * R_gd * v_para_grad - s_perp * R_para * v_gd_grad + R_gd / sq4 * b_para_grad - R_para / sq4 * b_gd_grad);
    A d5 = (vp + c_plus) * (al_p / RH * p_grad - s_perp * R_para * c_minus * al_m * v_para_grad
                            - s_perp * R_gd * c_minus * al_m * v_gd_grad + c_plus * al_p * v_perp_grad
                            + R_para * c_s * al_m / sq4 * b_para_grad + R_gd * c_s * al_m / sq4 * b_gd_grad);
    A d6 = (vp - c_plus) * (al_p / RH * p_grad + s_perp * R_para * c_minus * al_m * v_para_grad
                            + s_perp * R_gd * c_minus * al_m * v_gd_grad - c_plus * al_p * v_perp_grad
                            + R_para * c_s * al_m / sq4 * b_para_grad + R_gd * c_s * al_m / sq4 * b_gd_grad);
    A d7 = (vp + c_minus) * (al_m / RH * p_grad + s_perp * R_para * c_plus * al_p * v_para_grad
                             + s_perp * R_gd * c_plus * al_p * v_gd_grad + c_minus * al_m * v_perp_grad
                             - R_para * c_s * al_p / sq4 * b_para_grad - R_gd * c_s * al_p / sq4 * b_gd_grad);
    A d8 = (vp - c_minus) * (al_m / RH * p_grad - s_perp * R_para * c_plus * al_p * v_para_grad
                             - s_perp * R_gd * c_plus * al_p * v_gd_grad - c_minus * al_m * v_perp_grad
                             - R_para * c_s * al_p / sq4 * b_para_grad - R_gd * c_s * al_p / sq4 * b_gd_grad);
    const int pa = bi == 0 ? 1 : 0;
    // central derivative on the padded range plus the one-sided ones on the corner ranges
    auto alt = [&](const A& q) { return D.deriv_r(q, pa, RP) + D.back1(q, true, pa, RU) + D.back1(q, false, pa, RL); };
    auto alt2 = [&](const A& e, const A& in) {
      return D.deriv_r(e, pa, RP) + D.back1(e, true, pa, RU) + D.back1(e, false, pa, RL) + D.deriv_r(in, pa, RP) +
             D.back1(in, true, pa, RU) + D.back1(in, false, pa, RL);
    };
    const A bps_alt = 2.0 * b_perp * alt2(be_perp, bi_perp);
    const A p_alt = alt(G[press]);
    const A bp_alt = alt2(be_perp, bi_perp);
    const A bg_alt = alt2(be_gd, bi_gd);
    const A d3a = s_perp * s_perp * sg *
                  ((-1.0 * imask * R_gd * g_para + (R_gd * (p_alt + bps_alt / (8.0 * PI)) + bh * (bg_alt) / (4.0 * PI)) / RH));
    const A d5a = c_plus * al_p * (imask * g_perp + b_para * bp_alt / (4.0 * PI * RH)) +
                  c_minus * al_m * s_perp * (-imask * R_para * g_para + R_para * (p_alt + bps_alt / (8.0 * PI)) / RH);
    const A d7a = c_minus * al_m * (imask * g_perp + b_para * bp_alt / (4.0 * PI * RH)) -
                  c_plus * al_p * s_perp * (-imask * R_para * g_para + R_para * (p_alt + bps_alt / (8.0 * PI)) / RH);
    const int in_dir = lower ? 1 : -1;
    for (int i = RB.il; i <= RB.iu; ++i)
      for (int j = RB.jl; j <= RB.ju; ++j) {
        const double v = vp(i, j);
        if (in_dir * v > 0.0) { d1(i, j) = 0.0; d2(i, j) = 0.0; }
        if (in_dir * (v + c_perp(i, j)) > 0.0) d3(i, j) = d3a(i, j);
        if (in_dir * (v - c_perp(i, j)) > 0.0) d4(i, j) = -d3a(i, j);
        if (in_dir * (v + c_plus(i, j)) > 0.0) d5(i, j) = d5a(i, j);
        if (in_dir * (v - c_plus(i, j)) > 0.0) d6(i, j) = -d5a(i, j);
        if (in_dir * (v + c_minus(i, j)) > 0.0) d7(i, j) = d7a(i, j);
        if (in_dir * (v - c_minus(i, j)) > 0.0) d8(i, j) = -d7a(i, j);
      }
    const A& TE = G[thermal_energy];
    const A& P = G[press];
    const A sqp = vsqrt(PI * RH);
    std::array<A, 2> rho_t, vpara_t, vgd_t, vperp_t, te_t, bpara_t, bgd_t, bperp_t;
    rho_t[0] = -(pd.gamma / RH * d2 + 0.5 * RH * al_p * (d5 + d6) + 0.5 * RH * al_m * (d7 + d8)) / c_s_sq;
    vpara_t[0] = -0.5 * s_perp * (sg * R_gd * (-d3 + d4) + (c_minus * al_m / c_s_sq * R_para * (-d5 + d6) +
                                                           c_plus * al_p / c_s_sq * R_para * (d7 - d8))) +
                 vc * D.second_back(G[vpara], pf, bi, RB);
    vgd_t[0] = -0.5 * s_perp * (sg * R_para * (d3 - d4) + (c_minus * al_m / c_s_sq * R_gd * (-d5 + d6) +
                                                          c_plus * al_p / c_s_sq * R_gd * (d7 - d8))) +
               vc * D.second_back(G[vgd], pf, bi, RB);
    vperp_t[0] = -0.5 / c_s_sq * (c_plus * al_p * (d5 - d6) + c_minus * al_m * (d7 - d8)) +
                 vc * D.second_back(G[vperp], pf, bi, RB);
    te_t[0] = -(0.5 * (TE + P) * al_p * (d5 + d6) + 0.5 * (TE + P) * al_m * (d7 + d8)) / c_s_sq;
    bpara_t[0] = -sqp * (sg * R_gd * (d3 + d4) + al_m / c_s * R_para * (d5 + d6) - al_p / c_s * R_para * (d7 + d8));
    bgd_t[0] = -sqp * (sg * -1.0 * R_para * (d3 + d4) + al_m / c_s * R_gd * (d5 + d6) - al_p / c_s * R_gd * (d7 + d8));
    bperp_t[0] = -d1;

    const A& VA = G[vpara];
    auto dI = [&](const A& q) { return D.deriv_r(q, pa, RI); };
    auto tI = [&](const A& q) { return D.tderiv_r(q, VA, pa, RI); };
    rho_t[1] = -tI(RH);
    vpara_t[1] = -1.0 / RH * (dI(P) + 2.0 * b_para / (8.0 * PI) * (dI(be_para) + dI(bi_para)) +
                              2.0 * b_perp / (8.0 * PI) * (dI(be_perp) + dI(bi_perp)) +
                              2.0 * b_gd / (8.0 * PI) * (dI(be_gd) + dI(bi_gd))) -
                 (tI(RH * VA) - VA * tI(RH)) / RH + b_para / (4.0 * PI * RH) * (dI(be_para) + dI(bi_para)) +
                 imask * g_para + vc * D.second_r(VA, pa, RI);
    vgd_t[1] = -(tI(RH * G[vgd]) - G[vgd] * tI(RH)) / RH + b_para / (4.0 * PI * RH) * (dI(be_gd) + dI(bi_gd)) +
               vc * D.second_r(G[vgd], pa, RI);
    vperp_t[1] = -(tI(RH * vp) - vp * tI(RH)) / RH + b_para / (4.0 * PI * RH) * (dI(be_perp) + dI(bi_perp)) +
                 imask * g_perp + vc * D.second_r(vp, pa, RI);
    te_t[1] = -tI(TE) - P * dI(VA);
    bpara_t[1] = -(tI(be_para) + tI(bi_para)) + b_para * dI(VA);
    bgd_t[1] = -(tI(be_gd) + tI(bi_gd)) + b_para * dI(G[vgd]);
    bperp_t[1] = -(tI(be_perp) + tI(bi_perp)) + b_para * dI(vp);
    if (bi == 0) return {rho_t, vperp_t, vpara_t, vgd_t, te_t, bperp_t, bpara_t, bgd_t};
    return {rho_t, vpara_t, vperp_t, vgd_t, te_t, bpara_t, bperp_t, bgd_t};
  }
  // applyMomThresholdingMoC / applyBThresholdingMoC (idealmhd.cpp:107-224): clamp the open_moc ghost
  // cells and the first interior cell to [lo, hi] times the next interior cell, side by side
  static double clamp_ref(double x, double fl, double ce, double ref) {
    return ref >= 0.0 ? std::min(std::max(x, fl), ce) : std::max(std::min(x, fl), ce);
  }
  void moc_threshold(Grids& G) const {
    const int NX = pd.nx, NY = pd.ny, NG = N_GHOST;
    if (moc_mom) {
      for (int v : {mom_x, mom_y, mom_z}) {
        A& q = G[v];
        if (pd.bc[0] == OpenMoC)
          for (int j = 0; j < NY; ++j) for (int k = 0; k < NG + 1; ++k) {
            const double r = q(NG + 1, j);
            q(k, j) = clamp_ref(q(k, j), moc_mom_lo * r, moc_mom_hi * r, r);
          }
        if (pd.bc[1] == OpenMoC)
          for (int j = 0; j < NY; ++j) for (int k = 0; k < NG + 1; ++k) {
            const double r = q(NX - 2 - NG, j);
            q(NX - 1 - k, j) = clamp_ref(q(NX - 1 - k, j), moc_mom_lo * r, moc_mom_hi * r, r);
          }
        if (pd.bc[2] == OpenMoC)
          for (int i = 0; i < NX; ++i) for (int k = 0; k < NG + 1; ++k) {
            const double r = q(i, NG + 1);
            q(i, k) = clamp_ref(q(i, k), moc_mom_lo * r, moc_mom_hi * r, r);
          }
        if (pd.bc[3] == OpenMoC)
          for (int i = 0; i < NX; ++i) for (int k = 0; k < NG + 1; ++k) {
            const double r = q(i, NY - 2 - NG);
            q(i, NY - 1 - k) = clamp_ref(q(i, NY - 1 - k), moc_mom_lo * r, moc_mom_hi * r, r);
          }
      }
    }
    if (moc_b) {
      const int bes[3] = {be_x, be_y, be_z}, bis[3] = {bi_x, bi_y, bi_z};
      for (int c = 0; c < 3; ++c) {
        const A& e = pd.pd[bes[c]];
        A& q = G[bis[c]];
        if (pd.bc[0] == OpenMoC)
          for (int j = 0; j < NY; ++j) for (int k = 0; k < NG + 1; ++k) {
            const double r = e(NG + 1, j) + q(NG + 1, j);
            q(k, j) = clamp_ref(q(k, j), moc_b_lo * r - e(k, j), moc_b_hi * r - e(k, j), r);
          }
        if (pd.bc[1] == OpenMoC)
          for (int j = 0; j < NY; ++j) for (int k = 0; k < NG + 1; ++k) {
            const double r = e(NX - 2 - NG, j) + q(NX - 2 - NG, j);
            q(NX - 1 - k, j) = clamp_ref(q(NX - 1 - k, j), moc_b_lo * r - e(NX - 1 - k, j), moc_b_hi * r - e(NX - 1 - k, j), r);
          }
        if (pd.bc[2] == OpenMoC)
          for (int i = 0; i < NX; ++i) for (int k = 0; k < NG + 1; ++k) {
            const double r = e(i, NG + 1) + q(i, NG + 1);
            q(i, k) = clamp_ref(q(i, k), moc_b_lo * r - e(i, k), moc_b_hi * r - e(i, k), r);
          }
        if (pd.bc[3] == OpenMoC) {
          // the reference indexes the reference column with the x dimension (idealmhd.cpp:158)
          require(NX - 2 - NG < NY, "open_moc y_bound_2 with moc_b_limiting needs xdim - 4 < ydim (the reference reads outside the grid)");
          for (int i = 0; i < NX; ++i) for (int k = 0; k < NG + 1; ++k) {
            const double r = e(i, NX - 2 - NG) + q(i, NX - 2 - NG);
            q(i, NY - 1 - k) = clamp_ref(q(i, NY - 1 - k), moc_b_lo * r - e(i, NY - 1 - k), moc_b_hi * r - e(i, NY - 1 - k), r);
          }
        }
      }
    }
  }
  void evolved_from_state(Grids& G) const override {  // idealmhd.cpp:226-232
    G[n] = vmax(G[rho] / pd.m_ion, pd.nmin);
    G[press] = 2 * G[n] * K_B * vmax(G[temp], pd.tmin);
    G[thermal_energy] = G[press] / (pd.gamma - 1.0);
  }
  void minimums(Grids& G) const override {  // idealmhd.cpp:234-239
    G[rho] = pd.m_ion * vmax(G[rho] / pd.m_ion, pd.nmin);
    G[thermal_energy] = vmax(G[thermal_energy], pd.temin);
  }
  void derived(Grids& G) const override {  // idealmhd.cpp:241-277
    moc_threshold(G);
    G[n] = vmax(G[rho] / pd.m_ion, pd.nmin);
    G[rho] = G[n] * pd.m_ion;
    G[v_x] = G[mom_x] / G[rho];
    G[v_y] = G[mom_y] / G[rho];
    G[v_z] = G[mom_z] / G[rho];
    G[kinetic_energy] = 0.5 * G[rho] * (sq(G[v_x]) + sq(G[v_y]));
    G[thermal_energy] = vmax(G[thermal_energy], pd.temin);
    G[press] = (pd.gamma - 1.0) * G[thermal_energy];
    G[temp] = vmax(G[press] / (2 * K_B * G[n]), pd.tmin);
    G[b_x] = pd.pd[be_x] + G[bi_x];
    G[b_y] = pd.pd[be_y] + G[bi_y];
    G[b_z] = pd.pd[be_z] + G[bi_z];
    G[b_mag] = vsqrt(sq(G[b_x]) + sq(G[b_y]) + sq(G[b_z]));
    G[b_hat_x] = G[b_x] / G[b_mag];
    G[b_hat_y] = G[b_y] / G[b_mag];
    G[b_hat_z] = G[b_z] / G[b_mag];
    for (size_t c = 0; c < G[b_mag].size(); ++c)
      if (G[b_mag].v[c] == 0.0) { G[b_hat_x].v[c] = 0.0; G[b_hat_y].v[c] = 0.0; G[b_hat_z].v[c] = 0.0; }
  }
  void dt(Grids& G) const override {  // idealmhd.cpp:279-304
    A c_s = vsqrt(pd.gamma * G[press] / G[rho]);
    A c_s_sq = sq(c_s);
    A v_a = G[b_mag] / vsqrt(4.0 * PI * G[rho]);
    A v_a_sq = sq(v_a);
    A one(pd.nx, pd.ny, 1.0);
    A delta = vsqrt(one - 4.0 * c_s_sq * v_a_sq / sq(c_s_sq + v_a_sq));
    A v_fast = vsqrt(0.5 * (c_s_sq + v_a_sq) * (one + delta));
    A v_slow = vsqrt(0.5 * (c_s_sq + v_a_sq) * (one - delta));
    A vmx = vsqrt(sq(G[v_x]));
    A vmy = vsqrt(sq(G[v_y]));
    A cmax = vmax(vmax(vmax(c_s, v_a), v_fast), v_slow);
    G[dt_] = one / ((vmx + cmax) / pd.pd[d_x] + (vmy + cmax) / pd.pd[d_y]);
  }
};

// ---------------------------------------------------------------------------------------------
// Ideal2F: two fluids with electric and magnetic field evolution (source/equationsets/ideal2F.{hpp,cpp})
// ---------------------------------------------------------------------------------------------
constexpr double E_CH = 4.80320425e-10;   // constants.hpp:19
constexpr double C_LIGHT = 29979245800.0;  // constants.hpp:18 (an integer literal there)

struct Ideal2F : EqSet {
  enum V { i_rho, e_rho, i_mom_x, i_mom_y, e_mom_x, e_mom_y, i_temp, e_temp, bi_x, bi_y, bi_z, E_x, E_y, E_z,
           grav_x, grav_y, i_n, e_n, i_v_x, i_v_y, e_v_x, e_v_y, j_x, j_y, i_press, e_press, press,
           i_thermal_energy, e_thermal_energy, rho, rho_c, n, dn, dt_, dt_i, b_x, b_y, b_z, b_mag, b_mag_xy,
           b_hat_x, b_hat_y, curlE_z, divE, divB, i_dPdx, e_dPdx };
  bool use_sub_cycling = true, remove_curl_terms = false;
  explicit Ideal2F(Domain& d)
      : EqSet(d, {"i_rho", "e_rho", "i_mom_x", "i_mom_y", "e_mom_x", "e_mom_y", "i_temp", "e_temp", "bi_x", "bi_y",
                  "bi_z", "E_x", "E_y", "E_z", "grav_x", "grav_y", "i_n", "e_n", "i_v_x", "i_v_y", "e_v_x", "e_v_y",
                  "j_x", "j_y", "i_press", "e_press", "press", "i_thermal_energy", "e_thermal_energy", "rho",
                  "rho_c", "n", "dn", "dt", "dt_i", "b_x", "b_y", "b_z", "b_mag", "b_mag_xy", "b_hat_x", "b_hat_y",
                  "curlE_z", "divE", "divB", "i_dPdx", "e_dPdx"}) {}
  std::vector<int> state_vars() const override {
    return {i_rho, e_rho, i_mom_x, i_mom_y, e_mom_x, e_mom_y, i_temp, e_temp, bi_x, bi_y, bi_z, E_x, E_y, E_z, grav_x, grav_y};
  }
  std::vector<int> evolved() const override {
    return {i_rho, e_rho, i_mom_x, i_mom_y, e_mom_x, e_mom_y, i_thermal_energy, e_thermal_energy, E_x, E_y, E_z,
            bi_x, bi_y, bi_z};
  }
  std::vector<int> densities() const override { return {i_rho, e_rho}; }
  std::vector<int> thermal_energies() const override { return {i_thermal_energy, e_thermal_energy}; }
  std::vector<std::vector<int>> momenta() const override { return {{i_mom_x, i_mom_y}, {e_mom_x, e_mom_y}}; }
  std::vector<int> fields() const override { return {E_x, E_y, E_z}; }
  void parse(const std::vector<std::string>& k, const std::vector<std::string>& v) override {  // ideal2F.cpp:10-22
    for (size_t i = 0; i < k.size(); ++i) {
      if (k[i] == "use_sub_cycling") use_sub_cycling = v[i] == "true";
      else if (k[i] == "epsilon_courant" || k[i] == "verbose_2F") {}
      else if (k[i] == "remove_curl_terms") remove_curl_terms = v[i] == "true";
      else throw fail(k[i] + " is not recognized for this equation set.");
    }
  }
  void setup_derived() override { if (remove_curl_terms) use_sub_cycling = false; }  // :24-28
  // ideal2F.cpp:30-96
  Grids rhs(const Grids& G) const override {
    const Domain& D = pd;
    require(!use_sub_cycling, "ideal_2F with use_sub_cycling = true leaves 1x1 field derivatives (ideal2F.cpp:73-95)");
    const A &ivx = G[i_v_x], &ivy = G[i_v_y], &evx = G[e_v_x], &evy = G[e_v_y];
    A d_irho = -D.tdiv(G[i_rho], ivx, ivy);
    A d_erho = -D.tdiv(G[e_rho], evx, evy);
    std::vector<A> ivb = {G[i_v_y] * G[b_z], -G[i_v_x] * G[b_z]};
    std::vector<A> evb = {G[e_v_y] * G[b_z], -G[e_v_x] * G[b_z]};
    A iFx = E_CH * G[i_n] * (G[E_x] + ivb[0] / C_LIGHT);
    A iFy = E_CH * G[i_n] * (G[E_y] + ivb[1] / C_LIGHT);
    A eFx = -E_CH * G[e_n] * (G[E_x] + evb[0] / C_LIGHT);
    A eFy = -E_CH * G[e_n] * (G[E_y] + evb[1] / C_LIGHT);
    A d_imx = -D.tdiv(G[i_mom_x], ivx, ivy) - D.deriv(G[i_press], 0) + G[i_rho] * G[grav_x] + iFx;
    A d_imy = -D.tdiv(G[i_mom_y], ivx, ivy) - D.deriv(G[i_press], 1) + G[i_rho] * G[grav_y] + iFy;
    A d_emx = -D.tdiv(G[e_mom_x], evx, evy) - D.deriv(G[e_press], 0) + G[e_rho] * G[grav_x] + eFx;
    A d_emy = -D.tdiv(G[e_mom_y], evx, evy) - D.deriv(G[e_press], 1) + G[e_rho] * G[grav_y] + eFy;
    A d_ite = -D.tdiv(G[i_thermal_energy], ivx, ivy) - G[i_press] * D.div2D(ivx, ivy);
    A d_ete = -D.tdiv(G[e_thermal_energy], evx, evy) - G[e_press] * D.div2D(evx, evy);
    A dEx, dEy, dEz, dbx, dby, dbz;
    if (!remove_curl_terms) {
      dEx = (+C_LIGHT) * D.deriv(G[b_z], 1) - 4. * PI * G[j_x];
      dEy = (-C_LIGHT) * D.deriv(G[b_z], 0) - 4. * PI * G[j_y];
      dEz = (+C_LIGHT) * (D.deriv(G[b_y], 0) - D.deriv(G[b_x], 1));
      dbx = (-C_LIGHT) * D.deriv(G[E_z], 1);
      dby = (+C_LIGHT) * D.deriv(G[E_z], 0);
      dbz = (+C_LIGHT) * (D.deriv(G[E_x], 1) - D.deriv(G[E_y], 0));
    } else {
      dEx = (-4.) * PI * G[j_x];
      dEy = (-4.) * PI * G[j_y];
      A z(D.nx, D.ny, 0.0);
      dEz = z; dbx = z; dby = z; dbz = z;
    }
    Grids k = {d_irho, d_erho, d_imx, d_imy, d_emx, d_emy, d_ite, d_ete, dEx, dEy, dEz, dbx, dby, dbz};
    for (auto& g : k) g *= D.mask;
    return k;
  }
  void minimums(Grids& G) const override {  // :98-105
    G[i_rho] = pd.m_ion * vmax(G[i_rho] / pd.m_ion, pd.nmin);
    G[e_rho] = M_ELECTRON * vmax(G[e_rho] / M_ELECTRON, pd.nmin);
    G[i_thermal_energy] = vmax(G[i_thermal_energy], pd.temin);
    G[e_thermal_energy] = vmax(G[e_thermal_energy], pd.temin);
  }
  void evolved_from_state(Grids& G) const override {  // :107-117
    G[i_n] = G[i_rho] / pd.m_ion;
    G[e_n] = G[e_rho] / M_ELECTRON;
    G[i_press] = G[i_n] * K_B * G[i_temp];
    G[e_press] = G[e_n] * K_B * G[e_temp];
    G[press] = G[i_press] + G[e_press];
    G[i_thermal_energy] = G[i_press] / (pd.gamma - 1.0);
    G[e_thermal_energy] = G[e_press] / (pd.gamma - 1.0);
  }
  void derived(Grids& G) const override {  // :119-167
    const Domain& D = pd;
    G[i_n] = G[i_rho] / pd.m_ion;
    G[e_n] = G[e_rho] / M_ELECTRON;
    G[i_v_x] = G[i_mom_x] / G[i_rho];
    G[i_v_y] = G[i_mom_y] / G[i_rho];
    G[e_v_x] = G[e_mom_x] / G[e_rho];
    G[e_v_y] = G[e_mom_y] / G[e_rho];
    G[j_x] = G[i_n] * E_CH * G[i_v_x] - G[e_n] * E_CH * G[e_v_x];
    G[j_y] = G[i_n] * E_CH * G[i_v_y] - G[e_n] * E_CH * G[e_v_y];
    G[i_press] = (pd.gamma - 1.0) * G[i_thermal_energy];
    G[e_press] = (pd.gamma - 1.0) * G[e_thermal_energy];
    G[press] = G[i_press] + G[e_press];
    G[i_temp] = G[i_press] / (K_B * G[i_n]);
    G[e_temp] = G[e_press] / (K_B * G[e_n]);
    G[b_x] = D.pd[be_x] + G[bi_x];
    G[b_y] = D.pd[be_y] + G[bi_y];
    G[b_z] = G[bi_z];
    G[b_mag] = vsqrt(sq(G[b_x]) + sq(G[b_y]) + sq(G[b_z]));
    G[b_mag_xy] = vsqrt(sq(G[b_x]) + sq(G[b_y]));
    G[b_hat_x] = G[b_x] / G[b_mag_xy];
    G[b_hat_y] = G[b_y] / G[b_mag_xy];
    for (size_t c = 0; c < G[b_mag_xy].size(); ++c)
      if (G[b_mag_xy].v[c] == 0.0) { G[b_hat_x].v[c] = 0.0; G[b_hat_y].v[c] = 0.0; }
    G[rho] = G[i_rho] + G[e_rho];
    G[rho_c] = E_CH * (G[i_n] - G[e_n]);
    G[n] = G[i_n];
    G[dn] = G[i_n] - G[e_n];
    G[divE] = D.div2D(G[E_x], G[E_y]);
    G[divB] = D.div2D(G[b_x], G[b_y]);
    G[curlE_z] = -(D.deriv(G[E_x], 1) - D.deriv(G[E_y], 0));
    G[i_dPdx] = D.deriv(G[i_press], 0);
    G[e_dPdx] = D.deriv(G[e_press], 0);
  }
  void dt(Grids& G) const override {  // :169-208
    const A& dx = pd.pd[d_x];
    const A& dy = pd.pd[d_y];
    A lx = dx * 2;
    A kx = 2. * PI / lx;
    A ly = dy * 2;
    A ky = 2. * PI / ly;
    A vth = vsqrt(K_B * G[e_temp] / M_ELECTRON);
    A wpe = vsqrt(4 * PI * G[e_n] * E_CH * E_CH / M_ELECTRON);
    A wthx = vsqrt(3 * sq(kx) * sq(vth));
    A wthy = vsqrt(3 * sq(ky) * sq(vth));
    A wLx = vsqrt(sq(wpe) + sq(wthx));
    A wLy = vsqrt(sq(wpe) + sq(wthy));
    A vLx = wLx / kx;
    A vLy = wLy / ky;
    A vmx = vabs(G[e_v_x]) + vLx;
    A vmy = vabs(G[e_v_y]) + vLy;
    A dtv = 1. / (vmx / dx + vmy / dy);
    A dtem = dx * dy / (dx + dy) / C_LIGHT;
    G[dt_] = dtv;
    if (!use_sub_cycling && !remove_curl_terms) G[dt_] = vmin(G[dt_], dtem);
    A cs = vsqrt(pd.gamma * G[i_press] / G[i_rho]);
    A ivm = vabs(G[i_v_x]) + cs;
    A ivmy = vabs(G[i_v_y]) + cs;
    G[dt_i] = 1. / (ivm / dx + ivmy / dy);
  }
};

// Magnetoacoustic step-size grid shared by IdealMHD2E (idealmhd2E.cpp:115-137) and IdealMHDCons
// (idealmhdcons.cpp:54-73); unlike IdealMHD they take |v| with abs rather than sqrt(v^2).
static A mhd_dt_abs(const Domain& pd, const A& press, const A& rho, const A& bmag, const A& vx, const A& vy);

// ---------------------------------------------------------------------------------------------
// IdealMHD2E: one fluid, separate ion and electron energies (source/equationsets/idealmhd2E.{hpp,cpp})
// ---------------------------------------------------------------------------------------------
struct IdealMHD2E : EqSet {
  enum V { rho, i_temp, e_temp, mom_x, mom_y, bi_x, bi_y, grav_x, grav_y, n, i_press, e_press, press,
           i_thermal_energy, e_thermal_energy, v_x, v_y, kinetic_energy, b_x, b_y, b_mag, b_hat_x, b_hat_y, dt_ };
  explicit IdealMHD2E(Domain& d)
      : EqSet(d, {"rho", "i_temp", "e_temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x", "grav_y", "n", "i_press",
                  "e_press", "press", "i_thermal_energy", "e_thermal_energy", "v_x", "v_y", "kinetic_energy", "b_x",
                  "b_y", "b_mag", "b_hat_x", "b_hat_y", "dt"}) {}
  std::vector<int> state_vars() const override { return {rho, i_temp, e_temp, mom_x, mom_y, bi_x, bi_y, grav_x, grav_y}; }
  std::vector<int> evolved() const override { return {rho, mom_x, mom_y, i_thermal_energy, e_thermal_energy, bi_x, bi_y}; }
  // idealmhd2E.hpp:36-44: two species sharing one density and one momentum
  std::vector<int> densities() const override { return {rho, rho}; }
  std::vector<int> thermal_energies() const override { return {i_thermal_energy, e_thermal_energy}; }
  std::vector<std::vector<int>> momenta() const override { return {{mom_x, mom_y}, {mom_x, mom_y}}; }
  std::vector<int> fields() const override { return {bi_x, bi_y}; }
  void parse(const std::vector<std::string>& k, const std::vector<std::string>& v) override {  // idealmhd2E.cpp:10-20
    for (size_t i = 0; i < k.size(); ++i) {
      if (k[i] == "global_viscosity") (void)std::stod(v[i]);
      else if (k[i] == "viscosity_opt") {}
      else throw fail(k[i] + " is not recognized for this equation set.");
    }
  }
  Grids rhs(const Grids& G) const override {  // idealmhd2E.cpp:22-59
    const Domain& D = pd;
    const A &bex = D.pd[be_x], &bey = D.pd[be_y];
    const A &vx = G[v_x], &vy = G[v_y];
    A d_rho = -D.tdiv(G[rho], vx, vy);
    A curl_db = D.curl2D(G[bi_x], G[bi_y]) / (4.0 * PI);
    // CrossProductZ2D(a_z, {b_x, b_y}) = {-a_z * b_y, a_z * b_x} (grid.cpp:455-460)
    std::vector<A> ext = {-curl_db * bey, curl_db * bex};
    std::vector<A> in = {-curl_db * G[bi_y], curl_db * G[bi_x]};
    A d_mx = -D.tdiv(G[mom_x], vx, vy) - D.deriv(G[press], 0) + D.mask * (G[rho] * G[grav_x] + ext[0] + in[0]);
    A d_my = -D.tdiv(G[mom_y], vx, vy) - D.deriv(G[press], 1) + D.mask * (G[rho] * G[grav_y] + ext[1] + in[1]);
    A d_ite = -D.tdiv(G[i_thermal_energy], vx, vy) - G[i_press] * D.div2D(vx, vy);
    A d_ete = -D.tdiv(G[e_thermal_energy], vx, vy) - G[e_press] * D.div2D(vx, vy);
    // curlZ(CrossProduct2D(v, b)) with CrossProduct2D(a, b) = a_x b_y - a_y b_x (grid.cpp:448-452)
    std::vector<A> ind_ext = D.curlZ(vx * bey - vy * bex);
    std::vector<A> ind_in = D.curlZ(vx * G[bi_y] - vy * G[bi_x]);
    A d_bx = ind_ext[0] + ind_in[0];
    A d_by = ind_ext[1] + ind_in[1];
    Grids k = {d_rho, d_mx, d_my, d_ite, d_ete, d_bx, d_by};
    for (auto& g : k) g *= D.mask;
    return k;
  }
  void minimums(Grids& G) const override {  // idealmhd2E.cpp:61-67
    G[rho] = pd.m_ion * vmax(G[rho] / pd.m_ion, pd.nmin);
    G[i_thermal_energy] = vmax(G[i_thermal_energy], pd.temin);
    G[e_thermal_energy] = vmax(G[e_thermal_energy], pd.temin);
  }
  void evolved_from_state(Grids& G) const override {  // idealmhd2E.cpp:69-77
    G[n] = vmax(G[rho] / pd.m_ion, pd.nmin);
    G[i_press] = G[n] * K_B * vmax(G[i_temp], pd.tmin);
    G[e_press] = G[n] * K_B * vmax(G[e_temp], pd.tmin);
    G[i_thermal_energy] = vmax(G[i_press] / (pd.gamma - 1.0), pd.temin);
    G[e_thermal_energy] = vmax(G[e_press] / (pd.gamma - 1.0), pd.temin);
  }
  void derived(Grids& G) const override {  // idealmhd2E.cpp:79-113
    G[n] = vmax(G[rho] / pd.m_ion, pd.nmin);
    G[rho] = G[n] * pd.m_ion;
    G[v_x] = G[mom_x] / G[rho];
    G[v_y] = G[mom_y] / G[rho];
    G[kinetic_energy] = 0.5 * G[rho] * (sq(G[v_x]) + sq(G[v_y]));
    G[i_thermal_energy] = vmax(G[i_thermal_energy], pd.temin);
    G[e_thermal_energy] = vmax(G[e_thermal_energy], pd.temin);
    G[i_press] = (pd.gamma - 1.0) * G[i_thermal_energy];
    G[e_press] = (pd.gamma - 1.0) * G[e_thermal_energy];
    G[press] = G[i_press] + G[e_press];
    G[i_temp] = vmax(G[i_press] / (K_B * G[n]), pd.tmin);
    G[e_temp] = vmax(G[e_press] / (K_B * G[n]), pd.tmin);
    G[b_x] = pd.pd[be_x] + G[bi_x];
    G[b_y] = pd.pd[be_y] + G[bi_y];
    G[b_mag] = vsqrt(sq(G[b_x]) + sq(G[b_y]));
    G[b_hat_x] = G[b_x] / G[b_mag];
    G[b_hat_y] = G[b_y] / G[b_mag];
    for (size_t c = 0; c < G[b_mag].size(); ++c)
      if (G[b_mag].v[c] == 0.0) { G[b_hat_x].v[c] = 0.0; G[b_hat_y].v[c] = 0.0; }
  }
  void dt(Grids& G) const override { G[dt_] = mhd_dt_abs(pd, G[press], G[rho], G[b_mag], G[v_x], G[v_y]); }
};

// ---------------------------------------------------------------------------------------------
// IdealMHDCons: conservative total-energy form (source/equationsets/idealmhdcons.{hpp,cpp}).
// press_tot is read by the RHS but never computed (idealmhdcons.cpp:24-37 vs 96-128): it keeps
// the values of the .state file, and without them the reference asserts at setup
// (equationset.cpp:89-90).
// ---------------------------------------------------------------------------------------------
struct IdealMHDCons : EqSet {
  enum V { rho, temp, mom_x, mom_y, bi_x, bi_y, grav_x, grav_y, n, v_x, v_y, kinetic_energy, b_x, b_y, b_mag,
           mag_energy, b_hat_x, b_hat_y, thermal_energy, press, press_tot, energy, dt_ };
  explicit IdealMHDCons(Domain& d)
      : EqSet(d, {"rho", "temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x", "grav_y", "n", "v_x", "v_y",
                  "kinetic_energy", "b_x", "b_y", "b_mag", "mag_energy", "b_hat_x", "b_hat_y", "thermal_energy",
                  "press", "press_tot", "energy", "dt"}) {}
  std::vector<int> state_vars() const override { return {rho, temp, mom_x, mom_y, bi_x, bi_y, grav_x, grav_y}; }
  std::vector<int> evolved() const override { return {rho, mom_x, mom_y, energy, bi_x, bi_y}; }
  std::vector<int> densities() const override { return {rho}; }
  std::vector<int> thermal_energies() const override { return {thermal_energy}; }
  std::vector<std::vector<int>> momenta() const override { return {{mom_x, mom_y}}; }
  std::vector<int> fields() const override { return {bi_x, bi_y}; }
  void parse(const std::vector<std::string>&, const std::vector<std::string>&) override {}  // idealmhdcons.cpp:10-11
  // idealmhdcons.cpp:13-52 builds the derivative list {rho, mom_x, mom_y, bi_x, bi_y} (:48) for
  // six evolved variables and then indexes it up to 5 (:49): the energy derivative is never
  // returned and the reference aborts (or reads past the vector) on the first time derivative.
  // The restatement stops there too.
  Grids rhs(const Grids&) const override {
    throw fail("ideal_mhd_cons: the reference time derivative returns 5 grids for 6 evolved variables "
               "(idealmhdcons.cpp:48-49) and aborts; there is no step to restate");
  }
  void minimums(Grids& G) const override {  // idealmhdcons.cpp:89-94
    G[rho] = pd.m_ion * vmax(G[rho] / pd.m_ion, pd.nmin);
    G[thermal_energy] = vmax(G[thermal_energy], pd.temin);
  }
  void evolved_from_state(Grids& G) const override {  // idealmhdcons.cpp:96-109
    G[n] = vmax(G[rho] / pd.m_ion, pd.nmin);
    G[press] = 2 * G[n] * K_B * vmax(G[temp], pd.tmin);
    G[thermal_energy] = G[press] / (pd.gamma - 1.0);
    G[v_x] = G[mom_x] / G[rho];
    G[v_y] = G[mom_y] / G[rho];
    G[kinetic_energy] = 0.5 * G[rho] * (sq(G[v_x]) + sq(G[v_y]));
    G[b_x] = pd.pd[be_x] + G[bi_x];
    G[b_y] = pd.pd[be_y] + G[bi_y];
    G[b_mag] = vsqrt(sq(G[b_x]) + sq(G[b_y]));
    G[mag_energy] = sq(G[b_mag]) / (8. * PI);
    G[energy] = G[thermal_energy] + G[kinetic_energy] + G[mag_energy];
  }
  void derived(Grids& G) const override {  // idealmhdcons.cpp:111-128
    G[n] = vmax(G[rho] / pd.m_ion, pd.nmin);
    G[rho] = G[n] * pd.m_ion;
    G[v_x] = G[mom_x] / G[rho];
    G[v_y] = G[mom_y] / G[rho];
    G[kinetic_energy] = 0.5 * G[rho] * (sq(G[v_x]) + sq(G[v_y]));
    G[b_x] = pd.pd[be_x] + G[bi_x];
    G[b_y] = pd.pd[be_y] + G[bi_y];
    G[b_mag] = vsqrt(sq(G[b_x]) + sq(G[b_y]));
    G[mag_energy] = sq(G[b_mag]) / (8. * PI);
    G[b_hat_x] = G[b_x] / G[b_mag];
    G[b_hat_y] = G[b_y] / G[b_mag];
    // the floor applies to the kinetic energy term only (method call binds tighter than '-')
    G[thermal_energy] = G[energy] - G[mag_energy] - vmax(G[kinetic_energy], pd.temin);
    G[press] = (pd.gamma - 1.0) * G[thermal_energy];
    G[temp] = vmax(G[press] / (2 * K_B * G[n]), pd.tmin);
    for (size_t c = 0; c < G[b_mag].size(); ++c)
      if (G[b_mag].v[c] == 0.0) { G[b_hat_x].v[c] = 0.0; G[b_hat_y].v[c] = 0.0; }
  }
  void dt(Grids& G) const override { G[dt_] = mhd_dt_abs(pd, G[press], G[rho], G[b_mag], G[v_x], G[v_y]); }
};

static A mhd_dt_abs(const Domain& pd, const A& press, const A& rho, const A& bmag, const A& vx, const A& vy) {
  A c_s = vsqrt(pd.gamma * press / rho);
  A c_s_sq = sq(c_s);
  A v_a = bmag / vsqrt(4.0 * PI * rho);
  A v_a_sq = sq(v_a);
  A one(pd.nx, pd.ny, 1.0);
  A delta = vsqrt(one - 4.0 * c_s_sq * v_a_sq / sq(c_s_sq + v_a_sq));
  A v_fast = vsqrt(0.5 * (c_s_sq + v_a_sq) * (one + delta));
  A v_slow = vsqrt(0.5 * (c_s_sq + v_a_sq) * (one - delta));
  A vxm = vabs(vx) + vmax(vmax(vmax(c_s, v_a), v_fast), v_slow);
  A vym = vabs(vy) + vmax(vmax(vmax(c_s, v_a), v_fast), v_slow);
  return 1. / (vxm / pd.pd[d_x] + vym / pd.pd[d_y]);
}

static std::unique_ptr<EqSet> make_eqset(Domain& d, const std::string& name) {
  if (name == "ideal_mhd") return std::make_unique<IdealMHD>(d);
  if (name == "ideal_2F") return std::make_unique<Ideal2F>(d);
  if (name == "ideal_mhd_2E") return std::make_unique<IdealMHD2E>(d);
  if (name == "ideal_mhd_cons") return std::make_unique<IdealMHDCons>(d);
  throw fail("equation set <" + name + "> is not available in the restatement");
}

static std::unique_ptr<Module> make_module(Domain& d, const std::string& name);

}  // namespace orc

// =============================================================================================
// Module implementations live in spruce_oracle_modules.inc so the core above stays readable.
// =============================================================================================
#include "spruce_oracle_modules.inc"

// =============================================================================================
// C API
// =============================================================================================
using namespace orc;

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

#define GUARD_BEGIN try {
#define GUARD_END(ret) } catch (const std::exception& e) { g_err = e.what(); return ret; }

void* oracle_create(int xdim, int ydim, const int bc[4], int integrator, const double scal[8],
                    const char* eqset, int nopt, const char* const* keys, const char* const* vals) {
  GUARD_BEGIN
  auto* d = new Domain();
  d->nx = xdim; d->ny = ydim;
  for (int k = 0; k < 4; ++k) d->bc[k] = bc[k];
  d->integ = integrator;
  d->eps = scal[0]; d->nmin = scal[1]; d->tmin = scal[2]; d->temin = scal[3];
  d->obs = scal[4]; d->obdb = scal[5]; d->m_ion = scal[6]; d->gamma = scal[7];
  d->eqs = make_eqset(*d, eqset);
  std::vector<std::string> k, v;
  for (int i = 0; i < nopt; ++i) { k.emplace_back(keys[i]); v.emplace_back(vals[i]); }
  d->eqs->parse(k, v);
  return d;
  GUARD_END(nullptr)
}

int oracle_add_module(void* h, const char* name, int nopt, const char* const* keys, const char* const* vals) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  auto m = make_module(*d, name);
  std::vector<std::string> k, v;
  for (int i = 0; i < nopt; ++i) { k.emplace_back(keys[i]); v.emplace_back(vals[i]); }
  m->parse(k, v);
  d->mods.push_back(std::move(m));
  return 0;
  GUARD_END(-1)
}

int oracle_set_grid(void* h, const char* name, const double* data) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  A a(d->nx, d->ny);
  std::memcpy(a.v.data(), data, sizeof(double) * a.size());
  for (int k = 0; k < 7; ++k)
    if (std::strcmp(name, PD_NAMES[k]) == 0) { d->pd[k] = a; d->pd_set[k] = true; return 0; }
  d->eqs->g[d->eqs->idx(name)] = a;
  return 0;
  GUARD_END(-1)
}

int oracle_setup(void* h) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  for (int k = 0; k < 7; ++k) require(d->pd_set[k], std::string("grid ") + PD_NAMES[k] + " not set");
  for (int v : d->eqs->state_vars())
    require(d->eqs->g[v].size() > 1, "state variable " + d->eqs->names[v] + " not set");
  d->iteration_bounds();
  d->eqs->populate(d->eqs->g);
  for (size_t i = 0; i < d->eqs->g.size(); ++i)  // equationset.cpp:89-90
    require(d->eqs->g[i].size() == (size_t)d->nx * d->ny,
            "All variables must be initialized by <populateVariablesFromState> (" + d->eqs->names[i] + ")");
  d->eqs->setup_derived();
  for (auto& m : d->mods) m->setup();
  return 0;
  GUARD_END(-1)
}

int oracle_step(void* h, int nsteps, double* dt_out) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  for (int s = 0; s < nsteps; ++s) {
    d->advance();
    if (dt_out) dt_out[s] = d->last_dt;
  }
  return 0;
  GUARD_END(-1)
}

int oracle_get_grid(void* h, const char* name, double* out) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  const A* a = nullptr;
  for (int k = 0; k < 7; ++k) if (std::strcmp(name, PD_NAMES[k]) == 0) a = &d->pd[k];
  if (!a) a = &d->eqs->g[d->eqs->idx(name)];
  require((int)a->size() == d->nx * d->ny, std::string("grid ") + name + " is not initialised");
  std::memcpy(out, a->v.data(), sizeof(double) * a->size());
  return 0;
  GUARD_END(-1)
}

int oracle_rhs(void* h, double* out) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  Grids k = d->eqs->time_derivs(d->eqs->g);
  size_t n = (size_t)d->nx * d->ny;
  for (size_t i = 0; i < k.size(); ++i) std::memcpy(out + i * n, k[i].v.data(), sizeof(double) * n);
  return 0;
  GUARD_END(-1)
}

// EquationSet::propagateChanges on the live grids (equationset.cpp:212-220); *step_out = the next
// step's size, epsilon * min dt over the dt bounds (evolution.cpp:62).
int oracle_propagate(void* h, double* step_out) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  d->eqs->propagate(d->eqs->g);
  if (step_out) *step_out = d->eps * range_min(d->eqs->g[d->eqs->idx("dt")], d->xl_dt, d->yl_dt, d->xu_dt, d->yu_dt);
  return 0;
  GUARD_END(-1)
}

// PlasmaDomain::updateGhostZones on the live grids (evolution.cpp:126-152).
int oracle_ghost_fill(void* h) {
  GUARD_BEGIN
  auto* d = static_cast<Domain*>(h);
  d->ghost(d->eqs->g);
  return 0;
  GUARD_END(-1)
}

int oracle_num_evolved(void* h) { return (int)static_cast<Domain*>(h)->eqs->evolved().size(); }
double oracle_time(void* h) { return static_cast<Domain*>(h)->time; }

int oracle_subcycles(void* h, int* out, int nmax) {
  auto* d = static_cast<Domain*>(h);
  int n = 0;
  for (auto& m : d->mods) { if (n < nmax) out[n] = m->subcycles(); ++n; }
  return n;
}

void oracle_destroy(void* h) { delete static_cast<Domain*>(h); }

}  // extern "C"
