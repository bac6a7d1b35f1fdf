// Host driver of the time step and the C ABI declared in include/spruce_gpu.h.
//
// The Domain below plays PlasmaDomain's role for one GPU: it owns the device grids, keeps the
// reference's hook order (dt -> module pre -> module iterate -> integrator -> module post,
// evolution.cpp:59-82) and launches the kernels on one HIP stream.  The step size never leaves the
// device between steps: the dt reduction writes it into device memory and every kernel of the next
// step reads it from there, so a run of steps needs no host synchronisation.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/spruce_gpu.h"
#include "arith.hpp"
#include "mhd.hpp"
#include "host_common.hpp"
#include "solar.hpp"
#include "visc.hpp"
#include "pvisc.hpp"
#include "sources.hpp"
#include "adjoint.hpp"
#include "adj_modules.hpp"
#include "adj_pvisc.hpp"
#include "outputs.hpp"
#include "moc.hpp"
#include "crpow.hpp"
#include "tf_host.hpp"
#include "e2_host.hpp"

namespace spruce {

thread_local std::string g_err;

// Staging array of spruce_gpu_rhs_noalloc: the eight derivative grids of a grid of up to
// NOALLOC_CELLS cells (4 MiB).  Every GPU the library runs on holds its own copy until the process
// ends.  In host code the array's name is not a device address: the kernels get the address
// hipGetSymbolAddress returns for the device, looked up once per device with that device current.
constexpr size_t NOALLOC_CELLS = 65536;
__device__ __align__(256) double scratch[8 * NOALLOC_CELLS];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {};

static double* noalloc_scratch(int dev) {
  require(dev >= 0 && dev < 64, SPRUCE_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    hip_check(hipGetDevice(&cur), "hipGetDevice");
    if (cur != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    hip_check(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

static const char* PD_NAMES[7] = {"d_x", "d_y", "pos_x", "pos_y", "be_x", "be_y", "be_z"};

// IdealMHD variable names (idealmhd.hpp:19-26).
static const char* MHD_NAMES[25] = {"rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z",
                                    "grav_x", "grav_y", "n", "press", "thermal_energy", "v_x", "v_y",
                                    "v_z", "kinetic_energy", "b_x", "b_y", "b_z", "b_mag", "b_hat_x",
                                    "b_hat_y", "b_hat_z", "dt"};
static const char* MHD_STATE[10] = {"rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z", "grav_x", "grav_y"};
// Derived variables computed by k_mhd_derived, in its `which` order.
static const char* MHD_DERIVED[14] = {"temp", "press", "v_x", "v_y", "v_z", "kinetic_energy", "b_x", "b_y",
                                      "b_z", "b_mag", "b_hat_x", "b_hat_y", "b_hat_z", "dt"};

struct SetBufs {
  DBuf rho, n, mx, my, mz, te, bx, by, bz, rho_pre;
  void alloc(size_t N) { for (DBuf* b : all()) b->alloc(N); }
  std::vector<DBuf*> all() { return {&rho, &n, &mx, &my, &mz, &te, &bx, &by, &bz, &rho_pre}; }
  MhdSet view() { return MhdSet{rho.p, n.p, mx.p, my.p, mz.p, te.p, bx.p, by.p, bz.p, rho_pre.p}; }
  void swap(SetBufs& o) {
    auto a = all(), b = o.all();
    for (size_t k = 0; k < a.size(); ++k) a[k]->swap(*b[k]);
  }
};


// One Module of the reference plugin API (module.hpp:15-58) on the GPU.  pre_launch() enqueues the
// reductions its subcycle count needs; pre_finish() turns their results into the count on the host
// after the domain's one synchronisation per step; iterate() enqueues the subcycles and the live
// propagate that ends every reference iterateModule.
struct GpuModule {
  std::string name;
  int nsub = -1;
  virtual ~GpuModule() = default;
  virtual void setup(Domain& d) = 0;
  virtual void pre_launch(Domain& d, double* scal) = 0;
  virtual void pre_finish(double dt, const double* scal) = 0;
  virtual void iterate(Domain& d, double dt) = 0;
  virtual bool needs_bhat() const { return false; }
  virtual bool needs_dt_grid() const { return false; }
  // false when pre_launch/pre_finish do nothing (no subcycle-count reduction to read back)
  virtual bool needs_pre_scalars() const { return true; }
  // iterate() has work to do
  virtual bool iterates() const { return true; }
  // the step size (and so the simulation time) is needed on the host during the step
  virtual bool needs_host_dt() const { return iterates(); }
  // after every module's iterate, before the integrator (the live grids are final for this step)
  virtual void before_integrator(Domain&) {}
  // RHS terms of this module on grid set S (computeTimeDerivativesModule), appended to a stage
  virtual void rhs_terms(Domain&, const MhdSet&, StageArgs&) {}
  // the same terms added to a full RHS array (nevolved consecutive grids)
  virtual void add_rhs(Domain&, const MhdSet&, double*) {}
  // postIterateModule: after the integrator's final propagate, before the next step size
  virtual void post(Domain&) {}
  // the values commandLineMessage prints that live on the device (empty for most modules)
  virtual std::vector<double> report(Domain&) { return {}; }
  // Module::fileOutput (modulehandler.cpp:123-130): the grids this module appends to every stored
  // frame, name and device address, in the reference's order
  virtual void file_output(std::vector<std::pair<std::string, const double*>>&) {}
};

using OutList = std::vector<std::pair<std::string, const double*>>;
static void alloc_zero(DBuf& b, size_t n, hipStream_t s) {
  b.alloc(n);
  launch_fill(b.p, 0.0, n, s);
}

static double opt_d(const std::vector<std::pair<std::string, std::string>>& o, const std::string& k, bool& found) {
  for (auto& kv : o) if (kv.first == k) { found = true; return std::stod(kv.second); }
  found = false;
  return 0.0;
}
static std::string opt_s(const std::vector<std::pair<std::string, std::string>>& o, const std::string& k,
                         const std::string& def) {
  std::string v = def;
  for (auto& kv : o) if (kv.first == k) v = kv.second;
  return v;
}
// The largest double t with std::log10(t) <= c (the host C library's log10, which the reference
// calls): from 10^c, stepped down while log10 exceeds c, then up while the next double still
// satisfies it.  log10 is monotone there (checked over 2^16 doubles either side of each bound the
// radiative losses use), so log10(t) <= c exactly when t <= the result.
static double log10_upper(double c) {
  double t = std::pow(10.0, c);
  while (std::log10(t) > c) t = std::nextafter(t, 0.0);
  while (std::log10(std::nextafter(t, HUGE_VAL)) <= c) t = std::nextafter(t, HUGE_VAL);
  return t;
}
static int subcycle_count(double dt, double m) {
  const double q = dt / m;
  if (!(q < 1.0e8)) throw Error(SPRUCE_E_NUMERIC, "subcycle count out of range");
  return (int)q + 1;
}

// The state snapshots the adjoint keeps while it replays one step, for that step's reverse.
struct StepTape {
  std::vector<MhdSet> after_it;    // after each iterating module (its propagate included)
  MhdSet W{}, A3{};                // RK2 midpoint / RK4 first stage; after the integrator's final ghost fill
  MhdSet W2{}, W3{}, Lm{};         // RK4: second and third stages, and the live set they and the final
                                   // stage use as base (its ghost cells refilled from its interior)
  std::vector<MhdSet> after_post;  // after each post-iterate module
  double* k1 = nullptr;
  double* k2 = nullptr;
  double* k3 = nullptr;            // RK4
  double* kf = nullptr;            // RK4: (k1 + k4)/6 + (k2 + k3)/3
};

class Domain : public DomainBase {
 public:
  spruce_domain_desc desc{};
  std::string eqset;
  std::vector<std::pair<std::string, std::vector<std::pair<std::string, std::string>>>> modules;
  std::map<std::string, std::vector<double>> host_grids;
  bool ready = false;
  int nx = 0, ny = 0;
  size_t N = 0;
  int xl = 0, xu = 0, yl = 0, yu = 0;

  AxisBufs ax, ay;
  Mesh mesh{};
  MhdConst kc{};
  SetBufs L, W, W2;
  DBuf bex, bey, bez, gx, gy;
  DBuf kbuf[3];
  DBuf partial, step, history, scratch;
  int* err_flag = nullptr;
  double* hscal = nullptr;   // pinned host copy of the module scalars, step size and error flag
  int dt_blocks = 0;
  int kinds_all[4] = {0, 0, 0, 0}, kinds_ucnp[4] = {0, 0, 0, 0}, kinds_live[4] = {0, 0, 0, 0};
  double s1[4] = {1, 1, 1, 1}, s2[4] = {1, 1, 1, 1};
  bool any_open = false;
  int stage_kernel = -1;  // 0: wavefront strips (mhd2.hip), 1: one thread per cell, 2: tiles (mhd.hip), -1: by size
  // the final stage of a step also forms the next step size's per-block minima (tile form, no ghost
  // fill, limiter or module after the integrator)
  bool fuse_dt = false;
  int strip_rows = 0;     // rows per wavefront strip; 0 = chosen from the grid size
  int tc_kernel = -1;     // conduction update: 0 wavefront strips, 1 one thread per cell, -1 by size
  int tc_rows = 0;        // rows per conduction strip; 0 = chosen from the grid size
  // rows per strip of the adjoint's Euler conduction subcycle reverse; 0 = 8 where the grid allows
  // strips, -1 = the tiled one-thread-per-cell form
  int adj_tc_rows = 0;
  bool adj_tc_slots = true;   // keep the conduction records of whole segments where memory allows
  // multispecies mode (plasmadomain.hpp:133-137): cumulative electron, ion and Joule heating since
  // the last store, which the modules add to
  DBuf ms_grid[3];

  void set_option(const std::string& k, const std::string& v) override {
    if (k == "stage_kernel") {
      require(v == "strip" || v == "cell" || v == "tile" || v == "auto", SPRUCE_E_ARG,
              "stage_kernel must be strip, cell, tile or auto");
      stage_kernel = v == "cell" ? 1 : (v == "strip" ? 0 : (v == "tile" ? 2 : -1));
    } else if (k == "strip_rows") {
      const int r = std::stoi(v);
      require(r >= 4 && r <= 256 && r % 4 == 0, SPRUCE_E_ARG, "strip_rows must be a multiple of 4 in [4, 256]");
      strip_rows = r;
    } else if (k == "tc_kernel") {
      require(v == "strip" || v == "cell" || v == "auto", SPRUCE_E_ARG, "tc_kernel must be strip, cell or auto");
      tc_kernel = v == "cell" ? 1 : (v == "strip" ? 0 : -1);
    } else if (k == "multispecies_mode") {
      require(!ready, SPRUCE_E_STATE, "multispecies_mode must be set before setup");
      multispecies_on = v == "true";
    } else if (k == "tc_rows") {
      const int r = std::stoi(v);
      require(r >= 1 && r <= 256, SPRUCE_E_ARG, "tc_rows must be in [1, 256]");
      tc_rows = r;
    } else if (k == "adjoint_tc_records") {
      require(v == "auto" || v == "step", SPRUCE_E_ARG, "adjoint_tc_records must be auto or step");
      adj_tc_slots = v == "auto";
    } else if (k == "adjoint_tc_rows") {
      const int r = std::stoi(v);
      require(r >= -1 && r <= 256, SPRUCE_E_ARG, "adjoint_tc_rows must be -1, 0 or in [1, 256]");
      adj_tc_rows = r;
    } else {
      throw Error(SPRUCE_E_ARG, "unknown option " + k);
    }
  }

  explicit Domain(const spruce_domain_desc& d, const std::string& eq) : desc(d), eqset(eq) {
    require(d.xdim > 0 && d.ydim > 0, SPRUCE_E_ARG, "xdim and ydim must be positive");
    for (int k = 0; k < 4; ++k) require(d.bc[k] >= 0 && d.bc[k] <= 5, SPRUCE_E_ARG, "unknown boundary kind");
    require(d.time_integrator >= 0 && d.time_integrator <= 2, SPRUCE_E_ARG, "unknown time integrator");
    require(eq == "ideal_mhd" || eq == "ideal_mhd_cons" || eq == "ideal_mhd_2E" || eq == "ideal_2F", SPRUCE_E_ARG,
            "Equation set name <" + eq + "> not recognized.");
    require(eq == "ideal_mhd", SPRUCE_E_UNSUPPORTED, "equation set " + eq + " is not available in this build");
    nx = d.xdim;
    ny = d.ydim;
    N = (size_t)nx * ny;
    time = d.time;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  }
  ~Domain() {
    if (err_flag) (void)hipFree(err_flag);
    if (hscal) (void)hipHostFree(hscal);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void parse_eqset(const std::vector<std::string>& k, const std::vector<std::string>& v) {
    // IdealMHD::parseEquationSetConfigs (idealmhd.cpp:12-40): keys other than the MoC/viscosity
    // options are an error; global_viscosity only acts on open_moc boundaries.
    for (size_t i = 0; i < k.size(); ++i) {
      if (k[i] == "global_viscosity") gvisc = std::stod(v[i]);
      else if (k[i] == "viscosity_opt") {}
      else if (k[i] == "moc_b_limiting") lim.b = v[i] == "true" ? 1 : 0;
      else if (k[i] == "moc_mom_limiting") lim.mom = v[i] == "true" ? 1 : 0;
      else if (k[i] == "moc_b_lower_lim") {
        lim.b_lo = std::stod(v[i]);
        require(lim.b_lo <= 1.0, SPRUCE_E_ARG, "MoC B field lower threshold should be <=1.0");
      } else if (k[i] == "moc_b_upper_lim") {
        lim.b_hi = std::stod(v[i]);
        require(lim.b_hi >= 1.0, SPRUCE_E_ARG, "MoC B field upper threshold should be >=1.0");
      } else if (k[i] == "moc_mom_lower_lim") {
        lim.mom_lo = std::stod(v[i]);
        require(lim.mom_lo <= 1.0, SPRUCE_E_ARG, "MoC momentum threshold should be <=1.0");
      } else if (k[i] == "moc_mom_upper_lim") {
        lim.mom_hi = std::stod(v[i]);
        require(lim.mom_hi >= 1.0, SPRUCE_E_ARG, "MoC momentum threshold should be >=1.0");
      } else {
        throw Error(SPRUCE_E_ARG, k[i] + " is not recognized for this equation set.");
      }
    }
  }
  // open_moc boundaries (idealmhd.hpp:48, 59-64)
  double gvisc = 0.0;
  MocLimits lim{0, 0, 0.1, 10.0, 0.1, 10.0};
  bool any_moc = false;
  int moc_on[4] = {0, 0, 0, 0};
  MocSide moc_sides[4];
  DBuf mocbuf, moc_vc, moc_partial, moc_ops;
  int moc_base[5] = {0, 0, 0, 0, 0};
  // the characteristic terms of set S into mocbuf (and the global viscosity coefficient first)
  const double* moc_terms(const MhdSet& S) {
    if (!any_moc) return nullptr;
    MocArgs m{};
    m.S = S; m.st = stat(); m.K = kc; m.M = mesh;
    for (int b = 0; b < 4; ++b) m.side[b] = moc_sides[b];
    m.out = mocbuf.p;
    m.ops = moc_ops.p;
    for (int b = 0; b < 5; ++b) m.slot_base[b] = moc_base[b];
    m.vc = nullptr;
    if (gvisc != 0.0) {
      timed("moc_visc", [&] { launch_moc_visc(S, stat(), kc, mesh, gvisc, moc_partial.p, moc_vc.p, err_flag, stream); });
      m.vc = moc_vc.p;
    }
    timed("moc_rhs", [&] { launch_moc_rhs(m, stream); });
    return mocbuf.p;
  }
  // the ghost-cell limiters at the start of recomputeDerivedVarsFromEvolvedVars (idealmhd.cpp:244-245)
  void moc_limit(const MhdSet& G) {
    if (any_moc && (lim.mom || lim.b)) timed("moc_limit", [&] { launch_moc_threshold(G, stat(), mesh, moc_on, lim, stream); });
  }

  std::vector<std::unique_ptr<GpuModule>> mods;
  DBuf bhx, bhy, dtgrid, mod_partial, mod_scal;
  std::vector<int32_t> sub_hist;
  bool need_dt_grid = false;

  void add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>& opts) override;
  int num_evolved() const override { return 8; }

  bool is_pd(const std::string& n) const { for (auto* s : PD_NAMES) if (n == s) return true; return false; }
  bool is_state(const std::string& n) const { for (auto* s : MHD_STATE) if (n == s) return true; return false; }

  void set_grid(const std::string& name, const double* p, int64_t count) override {
    require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
    bool known = is_pd(name);
    for (auto* s : MHD_NAMES) known = known || name == s;
    require(known, SPRUCE_E_ARG, "Variable name <" + name + "> not recognized");
    host_grids[name].assign(p, p + N);
  }

  void upload(DBuf& b, const std::vector<double>& v) {
    b.alloc(N);
    hip_check(hipMemcpy(b.p, v.data(), N * sizeof(double), hipMemcpyHostToDevice), "upload grid");
  }

  void setup() override {
    require(!ready, SPRUCE_E_STATE, "setup called twice");
    for (auto* s : PD_NAMES) require(host_grids.count(s), SPRUCE_E_ARG, std::string("PlasmaDomain grid ") + s + " not set");
    for (auto* s : MHD_STATE) require(host_grids.count(s), SPRUCE_E_ARG, std::string("state variable ") + s + " not set");
    // computeIterationBounds (plasmadomain.cpp:138-161)
    require(nx > 2 * N_GHOST && ny > 2 * N_GHOST, SPRUCE_E_ARG, "Grid too small for ghost zones");
    const int* bc = desc.bc;
    for (int k = 0; k < 4; ++k) {
      moc_on[k] = bc[k] == SPRUCE_BC_OPEN_MOC ? 1 : 0;
      any_moc = any_moc || moc_on[k];
    }
    require(!any_moc || !(lim.mom || lim.b) || (nx >= 8 && ny >= 8), SPRUCE_E_UNSUPPORTED,
            "the open_moc limiters need xdim >= 8 and ydim >= 8 in this build");
    // the y_bound_2 field limiter reads column xdim-4 (idealmhd.cpp:158): outside the grid unless xdim-4 < ydim
    require(!(moc_on[3] && lim.b) || nx - 4 < ny, SPRUCE_E_ARG,
            "moc_b_limiting on y_bound_2 reads column xdim-4, outside the grid when xdim-4 >= ydim");
    require((bc[0] == 0) == (bc[1] == 0) && (bc[2] == 0) == (bc[3] == 0), SPRUCE_E_ARG,
            "an axis must be periodic on both sides or on neither");
    xl = bc[0] == 0 ? 0 : N_GHOST;
    xu = bc[1] == 0 ? nx - 1 : nx - N_GHOST - 1;
    yl = bc[2] == 0 ? 0 : N_GHOST;
    yu = bc[3] == 0 ? ny - 1 : ny - N_GHOST - 1;
    // The mesh must be separable: d_x constant along y, d_y constant along x (README.md:41).
    const auto& dxg = host_grids["d_x"];
    const auto& dyg = host_grids["d_y"];
    std::vector<double> dx(nx), dy(ny);
    for (int i = 0; i < nx; ++i) dx[i] = dxg[(size_t)i * ny];
    for (int j = 0; j < ny; ++j) dy[j] = dyg[j];
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        require(dxg[(size_t)i * ny + j] == dx[i], SPRUCE_E_UNSUPPORTED, "d_x must not vary along y");
        require(dyg[(size_t)i * ny + j] == dy[j], SPRUCE_E_UNSUPPORTED, "d_y must not vary along x");
      }
    for (double v : dx) require(std::isfinite(v) && v > 0.0, SPRUCE_E_ARG, "d_x must be positive and finite");
    for (double v : dy) require(std::isfinite(v) && v > 0.0, SPRUCE_E_ARG, "d_y must be positive and finite");
    // validateCellSizesAndPositions (plasmadomain.cpp:196-211)
    const auto& px = host_grids["pos_x"];
    const auto& py = host_grids["pos_y"];
    for (int i = 0; i + 1 < nx; ++i)
      for (int j = 0; j + 1 < ny; ++j) {
        double dpx = px[(size_t)(i + 1) * ny + j] - px[(size_t)i * ny + j];
        double ddx = 0.5 * dxg[(size_t)i * ny + j] + 0.5 * dxg[(size_t)(i + 1) * ny + j];
        double dpy = py[(size_t)i * ny + j + 1] - py[(size_t)i * ny + j];
        double ddy = 0.5 * dyg[(size_t)i * ny + j] + 0.5 * dyg[(size_t)i * ny + j + 1];
        require(std::abs(1.0 - dpx / ddx) < 1.0e-4 && std::abs(1.0 - dpy / ddy) < 1.0e-4, SPRUCE_E_ARG,
                "Cell sizes and positions must correspond");
      }
    const bool xper = bc[0] == 0, yper = bc[2] == 0;
    ax.build(dx, xper);
    ay.build(dy, yper);
    mesh.x = ax.view(nx, xl, xu, xper);
    mesh.y = ay.view(ny, yl, yu, yper);
    mesh.nx = nx;
    mesh.ny = ny;
    // open_moc sides take part in the step size (plasmadomain.cpp:155-159)
    mesh.x.dlo = xl - (moc_on[0] ? N_GHOST : 0);
    mesh.x.dhi = xu + (moc_on[1] ? N_GHOST : 0);
    mesh.y.dlo = yl - (moc_on[2] ? N_GHOST : 0);
    mesh.y.dhi = yu + (moc_on[3] ? N_GHOST : 0);
    for (int b = 0; b < 4; ++b) moc_sides[b] = moc_side(b, bc, nx, ny);

    kc.m = desc.ion_mass;
    kc.rm = 1.0 / desc.ion_mass;
    kc.gamma = desc.adiabatic_index;
    kc.gm1 = desc.adiabatic_index - 1.0;
    kc.rgm1 = 1.0 / kc.gm1;
    kc.nmin = desc.density_min;
    kc.temin = desc.thermal_energy_min;
    kc.tmin = desc.temp_min;
    kc.eps = desc.epsilon;
    kc.obs = desc.open_boundary_strength;
    kc.r4pi = 1.0 / (4.0 * PI);
    kc.r6 = 1.0 / 6.0;
    kc.r3 = 1.0 / 3.0;
    require(std::isfinite(kc.m) && kc.m > 0 && std::isnormal(kc.rm), SPRUCE_E_ARG, "ion_mass must be positive");
    require(std::isfinite(kc.gm1) && kc.gm1 > 0 && std::isnormal(kc.rgm1), SPRUCE_E_ARG, "adiabatic_index must exceed 1");

    // boundary kinds per ghost-fill call site and the open-boundary decay factors
    for (int k = 0; k < 4; ++k) {
      kinds_all[k] = bc[k];
      kinds_ucnp[k] = bc[k] == SPRUCE_BC_OPEN_UCNP ? bc[k] : 0;
      kinds_live[k] = (bc[k] == SPRUCE_BC_OPEN || bc[k] == SPRUCE_BC_FIXED || bc[k] == SPRUCE_BC_REFLECT) ? bc[k] : 0;
      any_open = any_open || bc[k] == SPRUCE_BC_OPEN;
      const std::vector<double>& d = k < 2 ? dx : dy;
      const int n = (int)d.size();
      const bool upper = k & 1;
      const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
      const double delta_last = d[e3];
      const double dist23 = 0.5 * (d[e2] + d[e3]);
      const double dist12 = 0.5 * (d[e1] + d[e2]);
      s2[k] = std::pow(desc.open_boundary_decay_base, dist23 / delta_last);   // evolution.cpp:166-167
      s1[k] = std::pow(desc.open_boundary_decay_base, dist12 / delta_last);
    }

    L.alloc(N);
    W.alloc(N);
    if (desc.time_integrator == SPRUCE_RK4) {
      W2.alloc(N);
      for (auto& b : kbuf) b.alloc(8 * N);
    }
    upload(bex, host_grids["be_x"]);
    upload(bey, host_grids["be_y"]);
    upload(bez, host_grids["be_z"]);
    upload(gx, host_grids["grav_x"]);
    upload(gy, host_grids["grav_y"]);
    hip_check(hipMemcpy(L.rho.p, host_grids["rho"].data(), N * 8, hipMemcpyHostToDevice), "upload rho");
    hip_check(hipMemcpy(L.mx.p, host_grids["mom_x"].data(), N * 8, hipMemcpyHostToDevice), "upload mom_x");
    hip_check(hipMemcpy(L.my.p, host_grids["mom_y"].data(), N * 8, hipMemcpyHostToDevice), "upload mom_y");
    hip_check(hipMemcpy(L.mz.p, host_grids["mom_z"].data(), N * 8, hipMemcpyHostToDevice), "upload mom_z");
    hip_check(hipMemcpy(L.bx.p, host_grids["bi_x"].data(), N * 8, hipMemcpyHostToDevice), "upload bi_x");
    hip_check(hipMemcpy(L.by.p, host_grids["bi_y"].data(), N * 8, hipMemcpyHostToDevice), "upload bi_y");
    hip_check(hipMemcpy(L.bz.p, host_grids["bi_z"].data(), N * 8, hipMemcpyHostToDevice), "upload bi_z");
    upload(scratch, host_grids["temp"]);

    if (strip_rows == 0) strip_rows = mhd_strip_rows_default(nx, ny);
    if (tc_rows == 0) tc_rows = tc_strip_rows_default(nx, ny);
    if (tc_kernel < 0) tc_kernel = tc_kernel_default(nx, ny);
    if (stage_kernel < 0) stage_kernel = mhd_stage_kernel_default(nx, ny);
    fuse_dt = stage_kernel == 2 && !(kinds_all[0] || kinds_all[1] || kinds_all[2] || kinds_all[3]) && !any_moc &&
              mods.empty();
    dt_blocks = mhd_dt_blocks(mesh);
    partial.alloc(std::max(dt_blocks, mhd_tile_blocks(mesh)));
    step.alloc(1);
    history.alloc(1);
    hip_check(hipMalloc(&err_flag, sizeof(int)), "hipMalloc");
    hip_check(hipMemset(err_flag, 0, sizeof(int)), "hipMemset");

    if (any_moc) {
      mocbuf.alloc(8 * N);   // zero: cells outside the open_moc regions keep a zero term
      moc_ops.alloc((size_t)moc_slot_bases(moc_sides, moc_base) * moc_ops_per_slot());
      moc_vc.alloc(1);
      moc_partial.alloc(dt_blocks);
    }
    // populateVariablesFromState (equationset.cpp:96-104)
    launch_mhd_populate(L.view(), scratch.p, kc, nx, ny, stream);
    launch_mhd_ghost(L.view(), kc, mesh, kinds_all, s1, s2, GHOST_NORMAL, stream);
    moc_limit(L.view());
    finish_step(-1);
    sync_and_check();
    if (!mods.empty()) {
      bhx.alloc(N);
      bhy.alloc(N);
      mod_partial.alloc(2 * (size_t)solar_blocks(mesh));
      mod_scal.alloc(2 * mods.size());
      hip_check(hipHostMalloc((void**)&hscal, (2 * mods.size() + 2) * sizeof(double)), "hipHostMalloc");
      for (auto& m : mods) {
        m->setup(*this);
        need_dt_grid = need_dt_grid || m->needs_dt_grid();
      }
      if (need_dt_grid) dtgrid.alloc(N);
    }
    if (multispecies_on)
      for (auto& g : ms_grid) alloc_zero(g, N, stream);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    ready = true;
  }

  size_t num_cells() const override { return N; }
  // Module::fileOutput of module k: the grids it writes to every stored frame
  int module_output(int k, int idx, std::string* name, double* host, int64_t count) override {
    require(ready, SPRUCE_E_STATE, "setup must precede module_output");
    require(k >= 0 && k < (int)mods.size(), SPRUCE_E_ARG, "module index out of range");
    OutList out;
    mods[k]->file_output(out);
    if (idx < 0) return (int)out.size();
    require(idx < (int)out.size(), SPRUCE_E_ARG, "module output index out of range");
    if (name) *name = out[idx].first;
    if (host) {
      require(count == (int64_t)N, SPRUCE_E_ARG, "count must equal xdim*ydim");
      hip_check(hipMemcpyAsync(host, out[idx].second, N * 8, hipMemcpyDeviceToHost, stream), "download output");
      hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    return (int)out.size();
  }
  // the multispecies accumulators (0 electron, 1 ion, 2 Joule heating); null unless multispecies_mode
  double* ms(int w) { return multispecies_on ? ms_grid[w].p : nullptr; }
  void multispecies(int which, double* host, int64_t count) override {
    require(ready, SPRUCE_E_STATE, "setup must precede multispecies");
    require(multispecies_on, SPRUCE_E_STATE, "multispecies_mode is off");
    require(which >= 0 && which < 3, SPRUCE_E_ARG, "multispecies grid index must be 0, 1 or 2");
    require(count == (int64_t)N, SPRUCE_E_ARG, "count must equal xdim*ydim");
    hip_check(hipMemcpyAsync(host, ms_grid[which].p, N * 8, hipMemcpyDeviceToHost, stream), "download multispecies");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }
  // cumulative heating restarts after every store (evolution.cpp:35-41)
  void multispecies_reset() override {
    require(ready, SPRUCE_E_STATE, "setup must precede multispecies_reset");
    require(multispecies_on, SPRUCE_E_STATE, "multispecies_mode is off");
    for (auto& g : ms_grid) launch_fill(g.p, 0.0, N, stream);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  // The live dt grid (recomputeDT after the latest propagate), for modules that read it.
  void refresh_dt_grid() {
    launch_mhd_derived(L.view(), stat(), kc, mesh, 13, dtgrid.p, stream);
  }

  // propagateChanges() on the live grids (equationset.cpp:212-220), as every reference
  // iterateModule ends (e.g. thermalconduction.cpp:154-155).
  void propagate_live() {
    timed("propagate_live", [&] { launch_propagate_live(L.view(), kc, nx, ny, any_open ? 1 : 0, stream); });
    ghost(L, kinds_all, GHOST_NORMAL);
    moc_limit(L.view());
  }

  // ModuleHandler pre/iterate loops (modulehandler.cpp:34-50; evolution.cpp:65-66).
  void run_modules(int s) {
    bool bh = false;
    for (auto& m : mods) bh = bh || m->needs_bhat();
    if (bh) timed("bhat", [&] { launch_bhat(L.view(), stat(), mesh, bhx.p, bhy.p, stream); });
    bool host = false;
    for (auto& m : mods) host = host || m->needs_pre_scalars() || m->needs_host_dt();
    for (size_t k = 0; k < mods.size(); ++k) mods[k]->pre_launch(*this, mod_scal.p + 2 * k);
    double dt = std::nan("");
    if (host) {
      const int ns = 2 * (int)mods.size();
      launch_gather_scalars(mod_scal.p, ns, step.p, err_flag, hscal, stream);
      hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
      if (hscal[ns + 1] != 0.0) {
        hip_check(hipMemset(err_flag, 0, sizeof(int)), "hipMemset");
        throw Error(SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (no finite dt)");
      }
      const std::vector<double> sc(hscal, hscal + ns + 1);
      dt = sc[ns];
      for (size_t k = 0; k < mods.size(); ++k) mods[k]->pre_finish(dt, sc.data() + 2 * k);
    }
    host_dt = dt;
    have_host_dt = host;
    size_t ni = 0;
    for (auto& m : mods)
      if (m->iterates()) {
        m->iterate(*this, dt);
        if (rec) copy_set(L.view(), rec->after_it[ni++], N, stream);
      }
    // counts are recorded after iterate(): viscosity modules fix theirs there (viscosity.cpp:125-180)
    for (size_t k = 0; k < mods.size(); ++k) sub_hist[(size_t)s * mods.size() + k] = mods[k]->nsub;
    for (auto& m : mods) m->before_integrator(*this);
  }
  // the step size of the current step and the simulation time at its start, on the host (valid
  // while have_host_dt: some module asked for them)
  double host_dt = 0.0, tstep = 0.0;
  bool have_host_dt = false;
  // set while the adjoint replays a step: where the step keeps its intermediate states, and which
  // modules act after the integrator
  StepTape* rec = nullptr;
  std::vector<bool> rec_post;

  MhdStatic stat() { return MhdStatic{bex.p, bey.p, bez.p, gx.p, gy.p}; }

  void stage(SetBufs& S, SetBufs& B, SetBufs& T, double mult, int mode, int kslot, const double* step_ptr = nullptr,
             double* dt_partial = nullptr) {
    StageArgs a = stage_args(S.view(), B.view(), T.view(), mult, mode, kslot, step_ptr, nullptr);
    a.dt_partial = dt_partial;
    launch_stage_mods(a);
  }
  // a stage with the modules' right-hand-side terms at its state a.S
  void launch_stage_mods(StageArgs a) {
    for (auto& m : mods) m->rhs_terms(*this, a.S, a);
    launch_stage(a);
  }
  // a stage on plain set views, for sets that live in a caller's buffer (no module terms)
  void stage_view(const MhdSet& S, const MhdSet& B, const MhdSet& T, double mult, int mode, const double* step_ptr,
                  double* kout) {
    require(mods.empty(), SPRUCE_E_UNSUPPORTED, "module terms need the domain's own sets");
    launch_stage(stage_args(S, B, T, mult, mode, 0, step_ptr, kout));
  }
  StageArgs stage_args(const MhdSet& S, const MhdSet& B, const MhdSet& T, double mult, int mode, int kslot,
                       const double* step_ptr, double* kout) {
    StageArgs a{};
    a.S = S; a.B = B; a.T = T;
    a.st = stat(); a.k = kc; a.mesh = mesh;
    a.step = step_ptr ? step_ptr : step.p; a.step_mult = mult; a.mode = mode;
    a.write_pre = any_open ? 1 : 0;
    a.write_n = 1;
    a.moc = moc_terms(S);
    if (mode == STAGE_STORE_K) a.kbuf = kout ? kout : kbuf[kslot].p;
    if (mode == STAGE_RK4_FINAL) { a.kbuf = kbuf[0].p; a.k2buf = kbuf[1].p; a.k3buf = kbuf[2].p; }
    return a;
  }
  void launch_stage(const StageArgs& a) {
    if (stage_kernel == 1) timed("mhd_stage", [&] { launch_mhd_stage(a, stream); });
    else if (stage_kernel == 2) timed("mhd_stage", [&] { launch_mhd_stage_tile(a, stream); });
    else timed("mhd_stage", [&] { launch_mhd_stage_strip(a, strip_rows, stream); });
  }

  void ghost(SetBufs& G, const int* kinds, int mode) { ghost_set(G.view(), kinds, mode); }
  void ghost_set(const MhdSet& G, const int* kinds, int mode) {
    if (!(kinds[0] || kinds[1] || kinds[2] || kinds[3])) return;
    timed("mhd_ghost", [&] { launch_mhd_ghost(G, kc, mesh, kinds, s1, s2, mode, stream); });
  }

  // End of a propagate on the live grids: per-cell dt, then the step size of the next step
  // (evolution.cpp:62) written on the device.
  void finish_step(int idx, bool fused = false) {
    if (!fused) timed("mhd_dt", [&] { launch_mhd_dt(L.view(), stat(), kc, mesh, nullptr, partial.p, stream); });
    const int nb = fused ? mhd_tile_blocks(mesh) : dt_blocks;
    timed("reduce_step", [&] {
      launch_reduce_step(partial.p, nb, kc.eps, step.p, idx >= 0 ? history.p : nullptr, idx, err_flag, stream);
    });
  }

  void sync_and_check() {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    int e = 0;
    hip_check(hipMemcpy(&e, err_flag, sizeof(int), hipMemcpyDeviceToHost), "read error flag");
    if (e) {
      hip_check(hipMemset(err_flag, 0, sizeof(int)), "hipMemset");
      throw Error(SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (no finite dt)");
    }
  }

  void advance(int s) {
    if (!mods.empty()) run_modules(s);
    const int integ = desc.time_integrator;
    const bool fused = fuse_dt && !rec;
    double* dtp = fused ? partial.p : nullptr;
    if (integ == SPRUCE_EULER) {                       // evolution.cpp:84-88
      if (rec) launch_stage_mods(stage_args(L.view(), L.view(), W.view(), 1.0, STAGE_STORE_K, 0, nullptr, rec->k1));
      else stage(L, L, W, 1.0, STAGE_PLAIN, 0, nullptr, dtp);
      L.swap(W);
    } else if (integ == SPRUCE_RK2) {                  // evolution.cpp:90-101
      if (rec) {                                       // the same stages, keeping k1, k2 and the midpoint
        launch_stage_mods(stage_args(L.view(), L.view(), rec->W, 0.5, STAGE_STORE_K, 0, nullptr, rec->k1));
        ghost_set(rec->W, kinds_ucnp, GHOST_NORMAL);
        launch_stage_mods(stage_args(rec->W, L.view(), L.view(), 1.0, STAGE_STORE_K, 0, nullptr, rec->k2));
      } else {
        stage(L, L, W, 0.5, STAGE_PLAIN, 0);
        ghost(W, kinds_ucnp, GHOST_NORMAL);
        // the midpoint's live-grid ghost writes are overwritten unread, except where a fixed side
        // zeroes momenta inside an open_moc side's ghost cells, which the final stage reads as its base
        if (any_moc) ghost(L, kinds_live, GHOST_LIVE_RAW);
        moc_limit(W.view());
        stage(W, L, L, 1.0, STAGE_PLAIN, 0, nullptr, dtp);
      }
    } else if (rec) {                                  // RK4, keeping every stage and derivative
      launch_stage_mods(stage_args(L.view(), L.view(), rec->W, 0.5, STAGE_STORE_K, 0, nullptr, rec->k1));
      ghost_set(rec->W, kinds_ucnp, GHOST_NORMAL);
      ghost(L, kinds_live, GHOST_LIVE_RAW);
      launch_stage_mods(stage_args(rec->W, L.view(), rec->W2, 0.5, STAGE_STORE_K, 0, nullptr, rec->k2));
      ghost_set(rec->W2, kinds_ucnp, GHOST_NORMAL);
      ghost(L, kinds_live, GHOST_LIVE_RAW);
      launch_stage_mods(stage_args(rec->W2, L.view(), rec->W3, 1.0, STAGE_STORE_K, 0, nullptr, rec->k3));
      ghost_set(rec->W3, kinds_ucnp, GHOST_NORMAL);
      ghost(L, kinds_live, GHOST_LIVE_RAW);
      copy_set(L.view(), rec->Lm, N, stream);
      StageArgs a = stage_args(rec->W3, L.view(), L.view(), 1.0, STAGE_RK4_FINAL, 0, nullptr, nullptr);
      a.kbuf = rec->k1; a.k2buf = rec->k2; a.k3buf = rec->k3; a.kf_out = rec->kf;
      launch_stage_mods(a);
    } else {                                           // evolution.cpp:103-124
      stage(L, L, W, 0.5, STAGE_STORE_K, 0);
      ghost(W, kinds_ucnp, GHOST_NORMAL);
      ghost(L, kinds_live, GHOST_LIVE_RAW);            // open/fixed/reflect write the live grids
      moc_limit(W.view());
      stage(W, L, W2, 0.5, STAGE_STORE_K, 1);
      ghost(W2, kinds_ucnp, GHOST_NORMAL);
      ghost(L, kinds_live, GHOST_LIVE_RAW);
      moc_limit(W2.view());
      stage(W2, L, W, 1.0, STAGE_STORE_K, 2);
      ghost(W, kinds_ucnp, GHOST_NORMAL);
      ghost(L, kinds_live, GHOST_LIVE_RAW);
      moc_limit(W.view());
      stage(W, L, L, 1.0, STAGE_RK4_FINAL, 0, nullptr, dtp);
    }
    ghost(L, kinds_all, GHOST_NORMAL);
    moc_limit(L.view());
    if (rec) copy_set(L.view(), rec->A3, N, stream);
    size_t np = 0;
    for (size_t k = 0; k < mods.size(); ++k) {
      mods[k]->post(*this);
      if (rec && k < rec_post.size() && rec_post[k]) copy_set(L.view(), rec->after_post[np++], N, stream);
    }
    if (have_host_dt) tstep += host_dt;               // evolution.cpp:80, after the post-iterate hooks
    finish_step(s + 1, fused);
  }

  void run(int nsteps, double* dt_out, int32_t* sub_out) override {
    require(ready, SPRUCE_E_STATE, "setup must precede step");
    require(nsteps >= 0, SPRUCE_E_ARG, "nsteps must be non-negative");
    if (nsteps == 0) return;
    if ((int)history.n < nsteps + 1) history.alloc(nsteps + 1);
    hip_check(hipMemcpyAsync(history.p, step.p, sizeof(double), hipMemcpyDeviceToDevice, stream), "copy step");
    sub_hist.assign((size_t)nsteps * std::max<size_t>(1, mods.size()), -1);
    tstep = time;
    for (int s = 0; s < nsteps; ++s) advance(s);
    sync_and_check();
    if (clock.enabled) clock.collect();
    std::vector<double> h(nsteps);
    hip_check(hipMemcpy(h.data(), history.p, nsteps * sizeof(double), hipMemcpyDeviceToHost), "read dt history");
    for (int s = 0; s < nsteps; ++s) {
      time += h[s];                                    // evolution.cpp:80
      iter++;
      if (dt_out) dt_out[s] = h[s];
    }
    if (sub_out) std::memcpy(sub_out, sub_hist.data(), sub_hist.size() * sizeof(int32_t));
  }

  void get_grid(const std::string& name, double* out, int64_t count) override {
    require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
    if (is_pd(name)) {
      auto it = host_grids.find(name);
      require(it != host_grids.end(), SPRUCE_E_ARG, "grid " + name + " not set");
      std::memcpy(out, it->second.data(), N * 8);
      return;
    }
    require(ready, SPRUCE_E_STATE, "setup must precede get_grid");
    const double* src = nullptr;
    if (name == "rho") src = L.rho.p;
    else if (name == "n") src = L.n.p;
    else if (name == "mom_x") src = L.mx.p;
    else if (name == "mom_y") src = L.my.p;
    else if (name == "mom_z") src = L.mz.p;
    else if (name == "thermal_energy") src = L.te.p;
    else if (name == "bi_x") src = L.bx.p;
    else if (name == "bi_y") src = L.by.p;
    else if (name == "bi_z") src = L.bz.p;
    else if (name == "grav_x") src = gx.p;
    else if (name == "grav_y") src = gy.p;
    else {
      int which = -1;
      for (int k = 0; k < 14; ++k) if (name == MHD_DERIVED[k]) which = k;
      require(which >= 0, SPRUCE_E_ARG, "Variable name <" + name + "> not recognized");
      launch_mhd_derived(L.view(), stat(), kc, mesh, which, scratch.p, stream);
      src = scratch.p;
    }
    hip_check(hipMemcpyAsync(out, src, N * 8, hipMemcpyDeviceToHost, stream), "download grid");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  double* device_grid(const std::string& name) override {
    require(ready, SPRUCE_E_STATE, "setup must precede device_grid");
    if (name == "rho") return L.rho.p;
    if (name == "n") return L.n.p;
    if (name == "mom_x") return L.mx.p;
    if (name == "mom_y") return L.my.p;
    if (name == "mom_z") return L.mz.p;
    if (name == "thermal_energy") return L.te.p;
    if (name == "bi_x") return L.bx.p;
    if (name == "bi_y") return L.by.p;
    if (name == "bi_z") return L.bz.p;
    throw Error(SPRUCE_E_ARG, "no device-resident grid named " + name);
  }

  void rhs(double* out, int64_t count) override {
    require(ready, SPRUCE_E_STATE, "setup must precede rhs");
    require(count == (int64_t)(8 * N), SPRUCE_E_ARG, "rhs output must hold nevolved*xdim*ydim doubles");
    DBuf tmp;
    tmp.alloc(8 * N);
    rhs_staged(tmp.p, out);
  }

  // The eight derivative grids of the live state, computed into dev (8*N doubles of device memory,
  // every element written) and copied to out.  The kernels and module hooks use buffers allocated
  // at set-up only.
  void rhs_staged(double* dev, double* out) {
    launch_mhd_rhs_full(L.view(), stat(), kc, mesh, dev, stream, moc_terms(L.view()));
    for (auto& m : mods) m->before_integrator(*this);
    for (auto& m : mods) m->add_rhs(*this, L.view(), dev);
    hip_check(hipMemcpyAsync(out, dev, 8 * N * 8, hipMemcpyDeviceToHost, stream), "download rhs");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  void rhs_noalloc(double* out, int64_t count) override {
    require(ready, SPRUCE_E_STATE, "setup must precede rhs");
    require(count == (int64_t)(8 * N), SPRUCE_E_ARG, "rhs output must hold nevolved*xdim*ydim doubles");
    require(N <= NOALLOC_CELLS, SPRUCE_E_ARG,
            "spruce_gpu_rhs_noalloc handles grids of at most " + std::to_string(NOALLOC_CELLS) + " cells (this one has " +
                std::to_string(N) + "); use spruce_gpu_rhs");
    rhs_staged(noalloc_scratch(device), out);
  }

  static void copy_set(SetBufs& from, SetBufs& to, size_t N, hipStream_t s) {
    auto a = from.all(), b = to.all();
    for (size_t k = 0; k < a.size(); ++k)
      hip_check(hipMemcpyAsync(b[k]->p, a[k]->p, N * 8, hipMemcpyDeviceToDevice, s), "copy set");
  }
  static void copy_set(const MhdSet& from, const MhdSet& to, size_t N, hipStream_t s) {
    const double* a[10] = {from.rho, from.n, from.mx, from.my, from.mz, from.te, from.bx, from.by, from.bz, from.rho_pre};
    double* b[10] = {to.rho, to.n, to.mx, to.my, to.mz, to.te, to.bx, to.by, to.bz, to.rho_pre};
    for (int k = 0; k < 10; ++k)
      hip_check(hipMemcpyAsync(b[k], a[k], N * 8, hipMemcpyDeviceToDevice, s), "copy set");
  }
  // a set of 10 grids laid out one after another from p
  static MhdSet set_at(double* p, size_t N) {
    return MhdSet{p, p + N, p + 2 * N, p + 3 * N, p + 4 * N, p + 5 * N, p + 6 * N, p + 7 * N, p + 8 * N, p + 9 * N};
  }
  // evolved grid e (rho, mom_x, mom_y, mom_z, thermal_energy, bi_x, bi_y, bi_z) of a set
  static double* evolved_of(const MhdSet& S, int e) {
    double* const g[8] = {S.rho, S.mx, S.my, S.mz, S.te, S.bx, S.by, S.bz};
    return g[e];
  }

  // Discrete adjoint of nsteps steps (reverse of advance(): modules, stages, ghost fills,
  // minimums/derived, step size), with the evolved grids checkpointed every `ckpt` steps and each
  // segment recomputed from its checkpoint during the backward sweep.  Defined after the modules.
  void adjoint(int nsteps, int ckpt, const double* seed, double* grad, double* dt_out) override;
  void replay_step(int n, const MhdSet& start, StepTape& tp, const std::vector<double>& times);
  void adj_ghost(const MhdSet& P, double* lam, const int* kinds = nullptr) {
    const int* k = kinds ? kinds : kinds_all;
    launch_adj_ghost_axis(P, kc, mesh, k, s1, s2, 1, lam, stream);   // y sides were filled last
    launch_adj_ghost_axis(P, kc, mesh, k, s1, s2, 0, lam, stream);
  }
  // dJ/dtheta of the parametrised heating modules after the last adjoint call
  std::vector<double> param_grad() override;
  std::vector<double> module_report(int k) override {
    require(k >= 0 && k < (int)mods.size(), SPRUCE_E_ARG, "module index out of range");
    return mods[k]->report(*this);
  }
  // device memory the last adjoint call allocated for itself (checkpoints, segment, replay
  // snapshots, adjoint vectors, conduction subcycle record)
  int64_t adj_bytes = 0;
  int64_t adjoint_bytes() const override { return adj_bytes; }

  // reverse of populateVariablesFromState for the .state grids kept from set-up
  void adjoint_state(const double* grad_evolved, double* grad_state) override {
    require(ready, SPRUCE_E_STATE, "setup must precede adjoint_state");
    require(grad_evolved && grad_state, SPRUCE_E_ARG, "null argument");
    DBuf rho_s, temp_s, lam, g;
    upload(rho_s, host_grids["rho"]);
    upload(temp_s, host_grids["temp"]);
    lam.alloc(8 * N);
    g.alloc(8 * N);
    hip_check(hipMemcpy(lam.p, grad_evolved, 8 * N * 8, hipMemcpyHostToDevice), "upload gradient");
    if (kinds_all[0] || kinds_all[1] || kinds_all[2] || kinds_all[3]) {
      // set-up ends with a ghost fill (equationset.cpp:96-104): rebuild the set-up state and reverse it
      SetBufs P;
      P.alloc(N);
      const char* src[7] = {"rho", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z"};
      DBuf* dst[7] = {&P.rho, &P.mx, &P.my, &P.mz, &P.bx, &P.by, &P.bz};
      for (int k = 0; k < 7; ++k)
        hip_check(hipMemcpy(dst[k]->p, host_grids[src[k]].data(), N * 8, hipMemcpyHostToDevice), "upload grid");
      launch_mhd_populate(P.view(), temp_s.p, kc, nx, ny, stream);
      launch_mhd_ghost(P.view(), kc, mesh, kinds_all, s1, s2, GHOST_NORMAL, stream);
      adj_ghost(P.view(), lam.p);
    }
    launch_adj_populate(rho_s.p, temp_s.p, lam.p, kc, N, g.p, stream);
    hip_check(hipMemcpyAsync(grad_state, g.p, 8 * N * 8, hipMemcpyDeviceToHost, stream), "download gradient");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  // Overwrites one evolved grid of the live state without propagating (a plugin writing
  // m_eqs->grid(v), e.g. thermalconduction.cpp:153); derived grids keep their values until the
  // next propagate.
  void put_grid(const std::string& name, const double* p, int64_t count) override {
    require(ready, SPRUCE_E_STATE, "setup must precede put_grid");
    require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
    DBuf* b = nullptr;
    if (name == "rho") b = &L.rho;
    else if (name == "mom_x") b = &L.mx;
    else if (name == "mom_y") b = &L.my;
    else if (name == "mom_z") b = &L.mz;
    else if (name == "thermal_energy") b = &L.te;
    else if (name == "bi_x") b = &L.bx;
    else if (name == "bi_y") b = &L.by;
    else if (name == "bi_z") b = &L.bz;
    require(b != nullptr, SPRUCE_E_ARG, "put_grid takes an evolved variable of ideal_mhd, not <" + name + ">");
    hip_check(hipMemcpyAsync(b->p, p, N * 8, hipMemcpyHostToDevice, stream), "upload grid");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  // PlasmaDomain::updateGhostZones on the live grids (evolution.cpp:126-152): the ghost cells'
  // evolved variables only (derived ones follow at the next propagate).
  void ghost_fill() override {
    require(ready, SPRUCE_E_STATE, "setup must precede ghost_fill");
    ghost(L, kinds_all, GHOST_LIVE_RAW);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  // EquationSet::propagateChanges on the live grids (equationset.cpp:212-220): minimums, ghost
  // fill, derived variables, dt; the next step's size epsilon * min dt (evolution.cpp:62) becomes
  // the domain's and is returned.
  double propagate() override {
    require(ready, SPRUCE_E_STATE, "setup must precede propagate");
    propagate_live();
    finish_step(-1);
    sync_and_check();
    return dt_min();
  }

  double dt_min() override {
    require(ready, SPRUCE_E_STATE, "setup must precede dt_min");
    double v = 0.0;
    hip_check(hipMemcpy(&v, step.p, sizeof(double), hipMemcpyDeviceToHost), "read step");
    return v;
  }
};


// ---------------------------------------------------------------------------------------------
// ThermalConduction (source/modules/solar/thermalconduction.cpp)
// ---------------------------------------------------------------------------------------------
struct TcModule : GpuModule {
  TcParams p{};
  double eps = 0.0;
  std::string integ = "euler";
  bool inactive = false;
  DBuf ta, tb, tc, k1, k2, k3, rx, ry, coef, dbxx, dbxy, dbyx, dbyy;
  bool to_file = false;           // output_to_file: the average heating rate and the saturation factor
  double msf = 1.0;               // ms_electron_heating_fraction (thermalconduction.hpp:34)
  DBuf avg, sat;
  // while recording (the adjoint's replay of a step), per subcycle: Euler: te after it; RK2: the
  // midpoint, te after it; RK4: i1, i2, i3, k1, k2, k3, te after it
  bool recording = false;
  std::vector<DBuf> tape;
  double h_last = 0.0;
  // The adjoint's segment recompute records every step of a segment into a slot of its own (slot >=
  // 0), and the replay of that step then takes its subcycle results from the slot (reuse) instead of
  // computing them again; slot_n = the subcycle count a slot holds
  // (slot_pool: one allocation; slot i holds slot_cap[i] records from record slot_off[i] on)
  DBuf slot_pool;
  std::vector<size_t> slot_off, slot_cap;
  std::vector<int> slot_n;
  int slot = -1;
  bool reuse = false;
  size_t cells = 0;
  int rec_per_sub() const { return integ == "euler" ? 1 : (integ == "rk2" ? 2 : 7); }
  void record(Domain& d, int s, int k, const double* src);
  double* rec_buf(Domain& d, int s, int k);   // the record's buffer (allocated on first use)
  const double* rec_at(int s, int k) const {
    const size_t idx = (size_t)s * rec_per_sub() + k;
    return slot >= 0 ? slot_pool.p + (slot_off[(size_t)slot] + idx) * cells : tape[idx].p;
  }
  const double* rec_out(int s) const { return rec_at(s, rec_per_sub() - 1); }
  explicit TcModule(const std::vector<std::pair<std::string, std::string>>& o) {
    name = "thermal_conduction";
    bool f1, f2, f3;
    const std::string sat = opt_s(o, "flux_saturation", "");
    eps = opt_d(o, "epsilon", f1);
    p.dsm = opt_d(o, "dt_subcycle_min", f2);
    require(!sat.empty() && f1 && f2, SPRUCE_E_ARG,
            "thermal_conduction needs flux_saturation, epsilon and dt_subcycle_min (uninitialised otherwise)");
    p.saturation = sat == "true";
    double wf = opt_d(o, "weakening_factor", f3);
    if (!f3) wf = 1.0;
    integ = opt_s(o, "time_integrator", "euler");
    if (integ.empty()) integ = "euler";
    inactive = opt_s(o, "inactive_mode", "false") == "true";
    to_file = opt_s(o, "output_to_file", "false") == "true";
    bool f4;
    const double m = opt_d(o, "ms_electron_heating_fraction", f4);
    if (f4) msf = m;
    require(msf >= 0.0 && msf <= 1.0, SPRUCE_E_ARG,
            "Thermal Conduction MS electron heating fraction must be between 0 and 1");
    require(integ == "euler" || integ == "rk2" || integ == "rk4", SPRUCE_E_ARG,
            "Invalid time integrator given for Thermal Conduction module");
    p.k0 = wf * KAPPA_0;
    p.c1 = K_B / (wf * KAPPA_0);
    p.c_sat = (1.0 / 6.0) * (3.0 / 2.0);
    p.sqrt_me = std::sqrt(M_ELECTRON);
  }
  bool needs_bhat() const override { return true; }
  void setup(Domain& d) override {
    cells = d.N;
    ta.alloc(d.N); tb.alloc(d.N); tc.alloc(d.N);
    dbxx.alloc(d.N); dbxy.alloc(d.N); dbyx.alloc(d.N); dbyy.alloc(d.N);
    if (integ == "rk4") { k1.alloc(d.N); k2.alloc(d.N); k3.alloc(d.N); }
    if (p.saturation) { rx.alloc(d.N); ry.alloc(d.N); coef.alloc(d.N); }
    if (to_file) { alloc_zero(avg, d.N, d.stream); alloc_zero(sat, d.N, d.stream); }
  }
  void file_output(OutList& o) override {  // thermalconduction.cpp:229-239
    if (!to_file) return;
    o.emplace_back("thermal_conduction", avg.p);
    o.emplace_back("flux_saturation", sat.p);
  }
  void pre_launch(Domain& d, double* scal) override {
    // the strip form of the update forms the b_hat derivatives itself; the grids serve the cell
    // form and the adjoint's replay
    if (d.tc_kernel == 1 || (recording && (reuse || slot < 0)))
      d.timed("bhat_derivs",
              [&] { launch_bhat_derivs(d.bhx.p, d.bhy.p, d.mesh, dbxx.p, dbxy.p, dbyx.p, dbyy.p, d.stream); });
    d.timed("tc_pre", [&] {
      launch_tc_pre(d.L.view(), d.bhx.p, d.bhy.p, p, d.kc, d.mesh, d.mod_partial.p, solar_blocks(d.mesh), scal,
                    d.stream);
    });
  }
  void pre_finish(double dt, const double* scal) override {  // thermalconduction.cpp:179-193
    if (p.saturation && scal[1] == 0.0) { nsub = 0; return; }
    require(scal[0] < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (conduction)");
    const double m = std::max(eps * scal[0], p.dsm);
    nsub = subcycle_count(dt, m);
  }
  void step(Domain& d, const double* src, const double* base, double* out, double mult, int mode, double* kout,
            double h) {
    if (p.saturation)
      d.timed("tc_sat", [&] {
        launch_tc_sat(src, d.L.n.p, d.L.rho.p, d.bhx.p, d.bhy.p, p, d.kc, d.mesh, rx.p, ry.p, coef.p, d.stream);
      });
    TcStepArgs a{};
    a.te_src = src; a.te_base = base; a.n = d.L.n.p; a.bhx = d.bhx.p; a.bhy = d.bhy.p;
    a.rx = rx.p; a.ry = ry.p; a.coef = coef.p; a.te_out = out; a.k_out = kout;
    a.k1 = k1.p; a.k2 = k2.p; a.k3 = k3.p; a.h = h; a.mult = mult; a.mode = mode; a.p = p; a.k = d.kc;
    a.dbxx = dbxx.p; a.dbxy = dbxy.p; a.dbyx = dbyx.p; a.dbyy = dbyy.p;
    a.mesh = d.mesh;
    a.rows = d.tc_kernel == 1 ? 0 : d.tc_rows;
    d.timed("tc_step", [&] { launch_tc_step(a, d.stream); });
  }
  void iterate(Domain& d, double dt) override {  // thermalconduction.cpp:91-156
    const double h = dt / (double)nsub;
    h_last = h;
    const double* te0 = d.L.te.p;
    if (to_file && p.saturation) {   // saturationTerms(temp) at the start of the step (:51-57)
      launch_tc_sat(te0, d.L.n.p, d.L.rho.p, d.bhx.p, d.bhy.p, p, d.kc, d.mesh, rx.p, ry.p, coef.p, d.stream);
      hip_check(hipMemcpyAsync(sat.p, coef.p, d.N * 8, hipMemcpyDeviceToDevice, d.stream), "copy saturation");
    }
    const double* cur = d.L.te.p;
    double* bufs[3] = {ta.p, tb.p, tc.p};
    auto other = [&](const double* a, const double* b) {
      for (double* q : bufs) if (q != a && q != b) return q;
      return bufs[0];
    };
    if (reuse) {   // the adjoint's replay of a step whose subcycles its slot holds
      require(slot >= 0 && slot_n[(size_t)slot] == nsub, SPRUCE_E_STATE,
              "conduction replay: the subcycle count differs from the recorded one");
      if (nsub) cur = rec_out(nsub - 1);
    }
    for (int s = 0; s < (reuse ? 0 : nsub); ++s) {
      if (integ == "euler") {
        // while recording, each subcycle writes its output into its record
        double* out = recording ? rec_buf(d, s, 0) : other(cur, nullptr);
        step(d, cur, cur, out, 1.0, TC_UPDATE, nullptr, h);
        cur = out;
      } else if (integ == "rk2") {
        double* half = other(cur, nullptr);
        step(d, cur, cur, half, 0.5, TC_UPDATE, nullptr, h);
        double* out = other(cur, half);
        step(d, half, cur, out, 1.0, TC_UPDATE, nullptr, h);
        if (recording) { record(d, s, 0, half); record(d, s, 1, out); }
        cur = out;
      } else {
        double* i1 = other(cur, nullptr);
        step(d, cur, cur, i1, 0.5, TC_STORE_K, k1.p, h);
        if (recording) { record(d, s, 0, i1); record(d, s, 3, k1.p); }
        double* i2 = other(cur, i1);
        step(d, i1, cur, i2, 0.5, TC_STORE_K, k2.p, h);
        if (recording) { record(d, s, 1, i2); record(d, s, 4, k2.p); }
        step(d, i2, cur, i1, 1.0, TC_STORE_K, k3.p, h);
        if (recording) { record(d, s, 2, i1); record(d, s, 5, k3.p); }
        step(d, i1, cur, i2, 1.0, TC_RK4_FINAL, nullptr, h);
        if (recording) record(d, s, 6, i2);
        cur = i2;
      }
    }
    if (recording && !reuse && slot >= 0) slot_n[(size_t)slot] = nsub;
    if (to_file) launch_out_rate(cur, te0, dt, d.N, avg.p, d.stream);   // (:101-103)
    if (d.multispecies_on) {                                            // (:105-108)
      if (msf < 1.0) launch_ms_change(d.ms(1), cur, te0, 1.0 - msf, true, d.N, d.stream);
      if (msf > 0.0) launch_ms_change(d.ms(0), cur, te0, msf, true, d.N, d.stream);
    }
    if (inactive) return;
    // the last subcycle's buffer becomes the live thermal energy (no copy unless it is a record)
    bool live = cur == d.L.te.p;
    for (DBuf* b : {&ta, &tb, &tc})
      if (!live && b->p == cur) { d.L.te.swap(*b); live = true; }
    if (!live) hip_check(hipMemcpyAsync(d.L.te.p, cur, d.N * 8, hipMemcpyDeviceToDevice, d.stream), "copy conduction");
    d.propagate_live();
  }
};

double* TcModule::rec_buf(Domain& d, int s, int k) {
  const size_t idx = (size_t)s * rec_per_sub() + k;
  if (slot >= 0) {
    require(idx < slot_cap[(size_t)slot], SPRUCE_E_STATE, "conduction record beyond its slot");
    return slot_pool.p + (slot_off[(size_t)slot] + idx) * d.N;
  }
  while (tape.size() <= idx) {
    tape.emplace_back();
    tape.back().alloc(d.N);
  }
  return tape[idx].p;
}
void TcModule::record(Domain& d, int s, int k, const double* src) {
  hip_check(hipMemcpyAsync(rec_buf(d, s, k), src, d.N * 8, hipMemcpyDeviceToDevice, d.stream), "record conduction");
}

// ---------------------------------------------------------------------------------------------
// RadiativeLosses (source/modules/solar/radiativelosses.cpp)
// ---------------------------------------------------------------------------------------------
struct RlModule : GpuModule {
  RlParams p{};
  std::string integ = "euler";
  bool inactive = false;
  bool to_file = false;           // output_to_file: the average loss rate ("rad")
  double msf = 1.0;               // ms_electron_heating_fraction (radiativelosses.hpp:31)
  DBuf tout, avg;
  explicit RlModule(const std::vector<std::pair<std::string, std::string>>& o) {
    name = "radiative_losses";
    bool f1, f2, f3;
    p.ramp = opt_d(o, "cutoff_ramp", f1);
    p.cut = opt_d(o, "cutoff_temp", f2);
    p.eps = opt_d(o, "epsilon", f3);
    require(f1 && f2 && f3, SPRUCE_E_ARG,
            "radiative_losses needs cutoff_ramp, cutoff_temp and epsilon (uninitialised otherwise)");
    p.cut_plus_ramp = p.cut + p.ramp;
    const double lim[6] = {4.97, 5.67, 6.18, 6.55, 6.90, 7.63};
    for (int k = 0; k < 6; ++k) p.tb[k] = log10_upper(lim[k]);
    p.prevent = opt_s(o, "prevent_subcycling", "false") == "true";
    integ = opt_s(o, "time_integrator", "euler");
    if (integ.empty()) integ = "euler";
    inactive = opt_s(o, "inactive_mode", "false") == "true";
    to_file = opt_s(o, "output_to_file", "false") == "true";
    bool f4;
    const double m = opt_d(o, "ms_electron_heating_fraction", f4);
    if (f4) msf = m;
    require(msf >= 0.0 && msf <= 1.0, SPRUCE_E_ARG,
            "Radiative Losses MS electron heating fraction must be between 0 and 1");
    require(integ == "euler" || integ == "rk2" || integ == "rk4", SPRUCE_E_ARG,
            "Invalid time integrator given for radiative losses module");
  }
  bool needs_dt_grid() const override { return p.prevent != 0; }
  void setup(Domain& d) override {
    tout.alloc(d.N);
    p.pd_eps = d.kc.eps;
    if (to_file) alloc_zero(avg, d.N, d.stream);
  }
  void file_output(OutList& o) override {  // radiativelosses.cpp:174-181
    if (to_file) o.emplace_back("rad", avg.p);
  }
  void pre_launch(Domain& d, double* scal) override {
    if (p.prevent) d.refresh_dt_grid();
    d.timed("rl_pre", [&] {
      launch_rl_pre(d.L.view(), d.dtgrid.p, p, d.kc, d.mesh, d.mod_partial.p, solar_blocks(d.mesh), scal, d.stream);
    });
  }
  void pre_finish(double dt, const double* scal) override {  // radiativelosses.cpp:200-205
    if (scal[1] == 0.0) { nsub = 0; return; }
    require(scal[0] < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (radiation)");
    nsub = subcycle_count(dt, p.eps * scal[0]);
  }
  double h_last = 0.0;
  void iterate(Domain& d, double dt) override {  // radiativelosses.cpp:84-140
    const double h = dt / (double)nsub;
    h_last = h;
    if (p.prevent) d.refresh_dt_grid();
    const int ig = integ == "euler" ? 0 : (integ == "rk2" ? 1 : 2);
    d.timed("rl_iterate", [&] { launch_rl_iterate(d.L.view(), d.dtgrid.p, p, d.kc, d.mesh, nsub, h, ig, tout.p, d.stream); });
    if (to_file) launch_out_rate(tout.p, d.L.te.p, dt, d.N, avg.p, d.stream);   // (:93)
    if (d.multispecies_on) {                                                   // (:94-97)
      if (msf < 1.0) launch_ms_change(d.ms(1), tout.p, d.L.te.p, 1.0 - msf, true, d.N, d.stream);
      if (msf > 0.0) launch_ms_change(d.ms(0), tout.p, d.L.te.p, msf, true, d.N, d.stream);
    }
    if (inactive) return;
    d.L.te.swap(tout);   // the result buffer becomes the live thermal energy (no copy)
    d.propagate_live();
  }
};

// ---------------------------------------------------------------------------------------------
// Artificial viscosity (source/modules/viscosity.cpp), module name "artificial_viscosity"
// ---------------------------------------------------------------------------------------------
static std::string no_space(std::string v) {  // clearWhitespace (utils.cpp)
  v.erase(std::remove_if(v.begin(), v.end(), [](unsigned char ch) { return std::isspace(ch); }), v.end());
  return v;
}
static std::vector<std::string> split_list(const std::string& v) {  // splitString(.., ',')
  std::vector<std::string> out;
  size_t a = 0;
  if (v.empty()) return out;
  while (true) {
    const size_t b = v.find(',', a);
    out.push_back(v.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  if (!out.empty() && out.back().empty()) out.pop_back();  // getline-based splitting drops a trailing empty field
  return out;
}

struct ViscModule : GpuModule {
  struct Term {
    std::string opt, diff, evol, species;
    double strength = 0.0, length = 0.0;
    int var = 0, evol_idx = 0, scale = 0;
    double scale_c = 1.0;
    std::vector<double> sgrid_host;
    DBuf sgrid, out;
    DBuf o_dq, o_lap, o_str, o_dt;   // file output of the last evaluation (hyperviscous dqdt: last subcycle)
    bool hyper() const { return strength > 1.0; }
  };
  bool out_visc = false, out_lap = false, out_str = false, out_dt = false;   // visc_output_*
  std::string in_opt, in_strength, in_diff, in_evol, in_length, in_species, shape, hv_integ;
  bool grad = false;
  double hv_eps = 1.0;  // viscosity.hpp:68
  std::vector<Term> terms;
  DBuf dtg, dtmin, part, init, d1, d2, d3, d4, dq;
  bool any_rhs = false, any_hyper = false;

  explicit ViscModule(const std::vector<std::pair<std::string, std::string>>& o) {  // viscosity.cpp:6-33
    name = "artificial_viscosity";
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "hv_time_integrator") hv_integ = v;
      else if (k == "gradient_correction") grad = v == "true";
      else if (k == "hv_epsilon") hv_eps = std::stod(v);
      else if (k == "visc_opt") in_opt = v;
      else if (k == "boundary_falloff_shape") shape = v;
      else if (k == "visc_strength") in_strength = v;
      else if (k == "visc_vars_to_diff") in_diff = v;
      else if (k == "visc_vars_to_evol") in_evol = v;
      else if (k == "visc_length") in_length = v;
      else if (k == "visc_species") in_species = v;
      else if (k == "visc_output_visc") out_visc = v == "true";
      else if (k == "visc_output_lap") out_lap = v == "true";
      else if (k == "visc_output_strength") out_str = v == "true";
      else if (k == "visc_output_timescale") out_dt = v == "true";
      // other keys only warn in the reference
    }
  }
  bool needs_pre_scalars() const override { return false; }
  bool iterates() const override { return any_hyper; }
  void pre_launch(Domain&, double*) override {}
  void pre_finish(double, const double*) override {}

  static int var_code(const std::string& n) {
    static const char* names[VV_COUNT] = {"rho", "n", "mom_x", "mom_y", "mom_z", "thermal_energy", "bi_x", "bi_y",
                                          "bi_z", "grav_x", "grav_y", "temp", "press", "v_x", "v_y", "v_z",
                                          "kinetic_energy", "b_x", "b_y", "b_z", "b_mag", "b_hat_x", "b_hat_y",
                                          "b_hat_z"};
    for (int k = 0; k < VV_COUNT; ++k) if (n == names[k]) return k;
    return -1;
  }
  static int evolved_code(const std::string& n) {
    static const char* ev[8] = {"rho", "mom_x", "mom_y", "mom_z", "thermal_energy", "bi_x", "bi_y", "bi_z"};
    for (int k = 0; k < 8; ++k) if (n == ev[k]) return k;
    return -1;
  }

  // getBoundaryViscosity (viscosity.cpp:288-338), evaluated once on the host with the C library's
  // exp: it depends on the cell positions only.
  std::vector<double> boundary(const Domain& d, double st, double len) const {
    const auto& x = d.host_grids.at("pos_x");
    const auto& y = d.host_grids.at("pos_y");
    const size_t N = d.N;
    auto gmax = [](const std::vector<double>& a) {
      double m = -DBL_MAX;
      for (double v : a) m = (m < v) ? v : m;
      require(m > -DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the max calculation");
      return m;
    };
    auto gmin = [](const std::vector<double>& a) {
      double m = DBL_MAX;
      for (double v : a) m = (v < m) ? v : m;
      require(m < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the min calculation");
      return m;
    };
    const double xmax = gmax(x), ymax = gmax(y), xmin = gmin(x), ymin = gmin(y);
    std::vector<double> r(N, 0.0);
    if (shape == "gaussian") {
      for (size_t c = 0; c < N; ++c) {
        const double a0 = (x[c] - xmin) / len, a1 = (x[c] - xmax) / len;
        const double a2 = (y[c] - ymax) / len, a3 = (y[c] - ymin) / len;
        r[c] = r[c] + std::exp(-2.3 * (a0 * a0)) * st;
        r[c] = r[c] + std::exp(-2.3 * (a1 * a1)) * st;
        r[c] = r[c] + std::exp(-2.3 * (a2 * a2)) * st;
        r[c] = r[c] + std::exp(-2.3 * (a3 * a3)) * st;
      }
    } else if (shape == "exp") {
      for (size_t c = 0; c < N; ++c) {
        r[c] = r[c] + std::exp(-2.3 * (x[c] - xmin) / len) * st;
        r[c] = r[c] + std::exp(2.3 * (x[c] - xmax) / len) * st;
        r[c] = r[c] + std::exp(2.3 * (y[c] - ymax) / len) * st;
        r[c] = r[c] + std::exp(-2.3 * (y[c] - ymin) / len) * st;
      }
    } else {
      const double xc = 0.5 * (xmin + xmax), yc = 0.5 * (ymin + ymax);
      const double ax = xmax - xc, ay = ymax - yc;
      const double sl = len / std::min(ax, ay);
      for (size_t c = 0; c < N; ++c) {
        const double dx = x[c] - xc, dy = y[c] - yc;
        const double rs = dx * dx / (ax * ax) + dy * dy / (ay * ay);
        const double m = ((rs - 1.0) < 0.0) ? (rs - 1.0) : 0.0;   // min(rs - 1, 0) with std::min semantics
        if (shape == "exp_elliptical") r[c] = std::exp(2.3 * m / sl) * st;
        else { const double t = m / sl; r[c] = std::exp(-2.3 * (t * t)) * st; }
      }
    }
    for (auto& v : r) v = (st < v) ? st : v;
    return r;
  }

  void setup(Domain& d) override {  // viscosity.cpp:35-108
    const auto opts = split_list(no_space(in_opt));
    const auto strs = split_list(no_space(in_strength));
    const auto diffs = split_list(no_space(in_diff));
    const auto evols = split_list(no_space(in_evol));
    const auto lens = split_list(no_space(in_length));
    const auto specs = split_list(in_species);
    const size_t n = opts.size();
    require(n > 0, SPRUCE_E_ARG, "artificial_viscosity: no terms given");
    require(strs.size() == n && diffs.size() == n && evols.size() == n && lens.size() == n && specs.size() == n,
            SPRUCE_E_ARG, "artificial_viscosity: every list must have one entry per term");
    if (hv_integ.empty()) hv_integ = "euler";
    require(hv_integ == "euler" || hv_integ == "rk2" || hv_integ == "rk4", SPRUCE_E_ARG,
            "Invalid hyperviscous time integrator given for Viscosity module");
    if (shape.empty()) shape = "gaussian";
    require(shape == "gaussian" || shape == "exp" || shape == "exp_elliptical" || shape == "gaussian_elliptical",
            SPRUCE_E_ARG, "Invalid boundary falloff shape given for Viscosity module");
    terms.resize(n);
    int n_rhs = 0;
    for (size_t i = 0; i < n; ++i) {
      Term& t = terms[i];
      t.opt = opts[i]; t.diff = diffs[i]; t.evol = evols[i]; t.species = specs[i];
      t.strength = std::stod(strs[i]);
      t.length = std::stod(lens[i]);
      require(t.opt == "local" || t.opt == "global" || t.opt == "boundary" || t.opt == "boundary_global",
              SPRUCE_E_ARG, "Viscosity option must be global, local, or boundary.");
      require(t.length >= 0, SPRUCE_E_ARG, "Length constants must greater than or equal to zero.");
      t.var = var_code(t.diff);
      require(t.var >= 0, SPRUCE_E_ARG,
              "Each variable to differentiate must be a valid variable within the chosen equation set.");
      t.evol_idx = evolved_code(t.evol);
      require(t.evol_idx >= 0, SPRUCE_E_ARG,
              "Each variable to evolve must be a valid evolved variable within the chosen equation set.");
      const bool mom = t.evol_idx >= 1 && t.evol_idx <= 3, vel = t.var >= VV_VX && t.var <= VV_VZ;
      if (mom && vel) {            // viscosity.cpp:231-239: ions scale by n*m_i; IdealMHD has no e_n
        require(t.species == "i", SPRUCE_E_UNSUPPORTED, "electron momentum viscosity needs e_n (not in ideal_mhd)");
        t.scale = 1;
        t.scale_c = d.kc.m;
      }
      if (t.evol_idx == 4 && t.var == VV_TEMP && (t.species == "i" || t.species == "e")) {  // :244-254
        t.scale = 2;
        t.scale_c = K_B / (d.desc.adiabatic_index - 1);
      }
      if (t.opt == "boundary" || t.opt == "boundary_global") {
        t.sgrid_host = boundary(d, t.strength, t.length);
        t.sgrid.alloc(d.N);
        hip_check(hipMemcpy(t.sgrid.p, t.sgrid_host.data(), d.N * 8, hipMemcpyHostToDevice), "upload strength");
      }
      if (t.hyper()) any_hyper = true;
      else { any_rhs = true; alloc_zero(t.out, d.N, d.stream); ++n_rhs; }
      // the output grids start at zero (:96-99); a term's RHS result is its masked dqdt (:117)
      if (out_visc && t.hyper()) alloc_zero(t.o_dq, d.N, d.stream);
      if (out_lap) alloc_zero(t.o_lap, d.N, d.stream);
      if (out_str) alloc_zero(t.o_str, d.N, d.stream);
      if (out_dt) alloc_zero(t.o_dt, d.N, d.stream);
    }
    require(n_rhs <= MAX_RHS_TERMS, SPRUCE_E_UNSUPPORTED, "at most 4 artificial_viscosity RHS terms");
    dtg.alloc(d.N);
    dtmin.alloc(1);
    part.alloc(d.dt_blocks);
    if (any_hyper) {
      init.alloc(d.N); d1.alloc(d.N);
      if (hv_integ == "rk4") { d2.alloc(d.N); d3.alloc(d.N); d4.alloc(d.N); dq.alloc(d.N); }
    }
  }

  // the live dt grid and its minimum over the dt bounds (viscosity.cpp:195-208)
  void refresh(Domain& d) {
    d.timed("visc_dt", [&] {
      launch_mhd_dt(d.L.view(), d.stat(), d.kc, d.mesh, dtg.p, part.p, d.stream);
      launch_reduce_step(part.p, d.dt_blocks, 1.0, dtmin.p, nullptr, 0, d.err_flag, d.stream);
    });
  }
  // the arguments of one term's evaluation on S (the adjoint reuses them)
  ViscTermArgs term_args(Domain& d, const Term& t, const MhdSet& S) {
    ViscTermArgs a{};
    a.S = S; a.st = d.stat(); a.k = d.kc; a.mesh = d.mesh;
    a.strength = t.sgrid.p;
    a.strength_c = t.strength;
    const bool local = t.opt == "local" || t.opt == "boundary";
    a.dtgrid = local ? dtg.p : nullptr;
    a.dtmin = dtmin.p;
    a.var = t.var; a.scale = t.scale; a.scale_c = t.scale_c; a.grad = grad ? 1 : 0;
    return a;
  }
  void term(Domain& d, const Term& t, const MhdSet& S, double* out, bool masked) {
    ViscTermArgs a = term_args(d, t, S);
    a.masked = masked ? 1 : 0;
    a.out = out;
    a.lap_out = t.o_lap.p; a.str_out = t.o_str.p; a.dt_out = t.o_dt.p;
    d.timed("visc_term", [&] { launch_visc_term(a, d.stream); });
  }
  void file_output(OutList& o) override {  // viscosity.cpp:351-376
    if (out_visc) for (auto& t : terms) o.emplace_back(t.evol + "_dqdt", t.hyper() ? t.o_dq.p : t.out.p);
    if (out_lap) for (auto& t : terms) o.emplace_back(t.evol + "_lap", t.o_lap.p);
    if (out_str) for (auto& t : terms) o.emplace_back(t.evol + "_str", t.o_str.p);
    if (out_dt) for (auto& t : terms) o.emplace_back(t.evol + "_dt", t.o_dt.p);
  }
  // The reference evaluates every term, hyperviscous ones included, in each right-hand side
  // (constructViscosityGrids, viscosity.cpp:112-121) and adds only the others: a hyperviscous
  // term's Laplacian, strength and timescale outputs come from that evaluation when the step's
  // integrator runs after the module's subcycles.
  bool side_out() const { return out_lap || out_str || out_dt; }
  void hyper_side(Domain& d, const MhdSet& S) {
    if (!side_out()) return;
    for (auto& t : terms)
      if (t.hyper()) term(d, t, S, d1.p, true);   // the dqdt itself is discarded
  }
  void before_integrator(Domain& d) override { if (any_rhs || (any_hyper && side_out())) refresh(d); }
  void rhs_terms(Domain& d, const MhdSet& S, StageArgs& a) override {
    hyper_side(d, S);
    for (auto& t : terms) {
      if (t.hyper()) continue;
      term(d, t, S, t.out.p, true);
      a.extra[a.n_extra] = t.out.p;
      a.extra_var[a.n_extra] = t.evol_idx;
      a.n_extra++;
    }
  }
  void add_rhs(Domain& d, const MhdSet& S, double* out) override {
    hyper_side(d, S);
    for (auto& t : terms) {
      if (t.hyper()) continue;
      term(d, t, S, t.out.p, true);
      launch_add(out + (size_t)t.evol_idx * d.N, t.out.p, d.N, d.stream);
    }
  }
  static double* evolved_ptr(SetBufs& L, int e) {
    DBuf* b[8] = {&L.rho, &L.mx, &L.my, &L.mz, &L.te, &L.bx, &L.by, &L.bz};
    return b[e]->p;
  }
  // while recording (the adjoint's replay of a step), per hyperviscous term and subcycle, for each
  // evaluation the live state it read (10 grids) and its dqdt, and with RK4 the combination after
  // the fourth; hv_rec: per hyperviscous term, where its entries start, its subcycle count and size
  bool recording = false;
  std::vector<DBuf> tape;
  size_t tape_n = 0;
  struct HvRec { size_t start; int n; double h; };
  std::vector<HvRec> hv_rec;
  void record(Domain& d, const double* src, size_t count) {
    if (tape_n >= tape.size()) tape.emplace_back();
    if (tape[tape_n].n != count) tape[tape_n].alloc(count);
    hip_check(hipMemcpyAsync(tape[tape_n].p, src, count * 8, hipMemcpyDeviceToDevice, d.stream), "record viscosity");
    ++tape_n;
  }
  void record_state(Domain& d) {
    if (tape_n >= tape.size()) tape.emplace_back();
    if (tape[tape_n].n != 10 * d.N) tape[tape_n].alloc(10 * d.N);
    Domain::copy_set(d.L.view(), Domain::set_at(tape[tape_n].p, d.N), d.N, d.stream);
    ++tape_n;
  }
  void iterate(Domain& d, double dt) override {  // viscosity.cpp:125-180
    nsub = -1;
    if (recording) { tape_n = 0; hv_rec.clear(); }
    for (auto& t : terms) {
      if (!t.hyper()) continue;
      const int n = (int)(std::ceil(t.strength / hv_eps) + 0.1);
      require(n > 0, SPRUCE_E_ARG, "hyperviscous subcycle count must be positive");
      if (nsub < 0) nsub = n;
      const double h = dt / (double)n;
      if (recording) hv_rec.push_back(HvRec{tape_n, n, h});
      double* g = evolved_ptr(d.L, t.evol_idx);
      auto eval = [&](double* out) {
        refresh(d);
        if (recording) record_state(d);
        term(d, t, d.L.view(), out, false);
        if (recording) record(d, out, d.N);
      };
      auto upd = [&](const double* base, const double* dd, double hh) {
        d.timed("visc_update", [&] { launch_visc_update(g, base, dd, hh, d.mesh, d.stream); });
        d.propagate_live();
      };
      for (int sc = 0; sc < n; ++sc) {
        if (hv_integ == "euler") {
          eval(d1.p);
          upd(g, d1.p, h);
        } else {
          hip_check(hipMemcpyAsync(init.p, g, d.N * 8, hipMemcpyDeviceToDevice, d.stream), "copy");
          if (hv_integ == "rk2") {
            eval(d1.p);
            upd(init.p, d1.p, 0.5 * h);
            eval(d1.p);
            upd(init.p, d1.p, h);
          } else {
            eval(d1.p);
            upd(init.p, d1.p, 0.5 * h);
            eval(d2.p);
            upd(init.p, d2.p, 0.5 * h);
            eval(d3.p);
            upd(init.p, d3.p, h);
            eval(d4.p);
            launch_visc_rk4(d1.p, d2.p, d3.p, d4.p, dq.p, d.N, d.stream);
            if (recording) record(d, dq.p, d.N);
            upd(init.p, dq.p, h);
          }
        }
      }
      // the stored dqdt: the last subcycle's final evaluation (rk2) or combination (rk4); Euler
      // subcycles leave it unchanged (:133-173)
      if (out_visc && hv_integ != "euler")
        hip_check(hipMemcpyAsync(t.o_dq.p, hv_integ == "rk2" ? d1.p : dq.p, d.N * 8, hipMemcpyDeviceToDevice, d.stream),
                  "copy dqdt");
      d.propagate_live();
    }
  }
};

// ---------------------------------------------------------------------------------------------
// AmbientHeating (source/modules/solar/ambientheating.cpp)
// ---------------------------------------------------------------------------------------------
struct HeatModule : GpuModule {
  double rate = 0.0, exp_base = 0.0, exp_h = 0.0, split_h = 0.0, split_start = 0.0;
  double msf = 0.5;   // ms_electron_heating_fraction (ambientheating.hpp:28)
  bool exp_mode = false, split_mode = false;
  DBuf heating;
  explicit HeatModule(const std::vector<std::pair<std::string, std::string>>& o) {  // :12-26
    name = "ambient_heating";
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "heating_rate") rate = std::stod(v);
      if (k == "exp_mode") exp_mode = v == "true";
      else if (k == "exp_base_heating_rate") exp_base = std::stod(v);
      else if (k == "exp_scale_height") exp_h = std::stod(v);
      else if (k == "split_exp_mode") split_mode = v == "true";
      else if (k == "split_exp_scale_height") split_h = std::stod(v);
      else if (k == "split_exp_start_height") split_start = std::stod(v);
      else if (k == "ms_electron_heating_fraction") msf = std::stod(v);
    }
  }
  bool needs_pre_scalars() const override { return false; }
  bool iterates() const override { return false; }
  void pre_launch(Domain&, double*) override {}
  void pre_finish(double, const double*) override {}
  void iterate(Domain&, double) override {}
  // the heating grid depends on positions only: built once on the host with the C library's exp
  // (setupModule, :28-40)
  void setup(Domain& d) override {
    const auto& y = d.host_grids.at("pos_y");
    std::vector<double> h(d.N);
    for (int i = 0; i < d.nx; ++i)
      for (int j = 0; j < d.ny; ++j) {
        const size_t c = (size_t)i * d.ny + j;
        const double mask = (i >= d.xl && i <= d.xu && j >= d.yl && j <= d.yu) ? 1.0 : 0.0;
        if (exp_mode) {
          double v = mask * exp_base * std::exp(-1.0 * y[c] / exp_h);
          if (split_mode) {
            const double sb = exp_base * std::exp((exp_h - split_h) * split_start / (exp_h * split_h));
            const double o = mask * sb * std::exp(-1.0 * y[c] / split_h);
            v = (v < o) ? o : v;
          }
          h[c] = v;
        } else {
          h[c] = mask * rate;
        }
      }
    require(msf >= 0.0 && msf <= 1.0, SPRUCE_E_ARG,
            "Ambient Heating MS electron heating fraction must be between 0 and 1");
    heating.alloc(d.N);
    hip_check(hipMemcpy(heating.p, h.data(), d.N * 8, hipMemcpyHostToDevice), "upload heating");
  }
  void post(Domain& d) override {  // postIterateModule (:42-49)
    d.timed("heating", [&] { launch_add_step_times(d.L.te.p, heating.p, d.step.p, d.N, d.stream); });
    d.propagate_live();
    if (d.multispecies_on) {
      if (msf < 1.0) launch_ms_source(d.ms(1), heating.p, 0.0, d.step.p, 1.0 - msf, MS_SRC_RATE, d.mesh, d.stream);
      if (msf > 0.0) launch_ms_source(d.ms(0), heating.p, 0.0, d.step.p, msf, MS_SRC_RATE, d.mesh, d.stream);
    }
  }
};

// ---------------------------------------------------------------------------------------------
// PhysicalViscosity (source/modules/solar/physicalviscosity.cpp)
// ---------------------------------------------------------------------------------------------
struct PvModule : GpuModule {
  double coeff = 0.0, ramp = 0.0, epsilon = 1.0;
  double msf = 0.0;   // ms_electron_heating_fraction (physicalviscosity.hpp:33)
  bool heating_on = true, force_on = true, inactive = false, grad_corr = false, to_file = false;
  std::string integ;
  DBuf coef, bhx, bhy, bhz, part, tsmin;
  DBuf avg_heat, avg_f[3];   // output_to_file: subcycle averages of the heating and the force
  DBuf sa[8], sb[8], sh[8], dv[7];
  // while recording (the adjoint's replay of a step): per subcycle the evaluation sets (8 grids
  // each): Euler: its start; RK2: its start and half step; and the subcycle size
  bool recording = false;
  std::vector<DBuf> tape;
  double h_last = 0.0;
  void record(Domain& d, size_t idx, const PvSet& src) {
    while (tape.size() <= idx) tape.emplace_back();
    if (tape[idx].n != 8 * d.N) tape[idx].alloc(8 * d.N);
    const double* g[8] = {src.te, src.T, src.mx, src.my, src.mz, src.vx, src.vy, src.vz};
    for (int k = 0; k < 8; ++k)
      hip_check(hipMemcpyAsync(tape[idx].p + (size_t)k * d.N, g[k], d.N * 8, hipMemcpyDeviceToDevice, d.stream),
                "record physical viscosity");
  }
  static PvSet set_of(double* p, size_t N) {
    return PvSet{p, p + N, p + 2 * N, p + 3 * N, p + 4 * N, p + 5 * N, p + 6 * N, p + 7 * N};
  }
  explicit PvModule(const std::vector<std::pair<std::string, std::string>>& o) {  // :17-34
    name = "physical_viscosity";
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "coeff") coeff = std::stod(v);
      else if (k == "ramp_length") ramp = std::stod(v);
      else if (k == "buffer_length") (void)std::stod(v);
      else if (k == "epsilon") epsilon = std::stod(v);
      else if (k == "heating_on") heating_on = v == "true";
      else if (k == "force_on") force_on = v == "true";
      else if (k == "inactive_mode") inactive = v == "true";
      else if (k == "gradient_correction") grad_corr = v == "true";
      else if (k == "time_integrator") integ = v;
      else if (k == "output_to_file") to_file = v == "true";
      else if (k == "ms_electron_heating_fraction") msf = std::stod(v);
    }
    if (integ.empty()) integ = "euler";
    require(integ == "euler" || integ == "rk2", SPRUCE_E_ARG, "Invalid time integrator given for Physical Viscosity module");
    require(msf >= 0.0 && msf <= 1.0, SPRUCE_E_ARG,
            "Physical Viscosity MS electron heating fraction must be between 0 and 1");
  }
  bool needs_pre_scalars() const override { return false; }
  void pre_launch(Domain&, double*) override {}
  void pre_finish(double, const double*) override {}
  static PvSet view(DBuf* b) { return PvSet{b[0].p, b[1].p, b[2].p, b[3].p, b[4].p, b[5].p, b[6].p, b[7].p}; }
  // constructCoefficientGrid (:252-270), once on the host with the C library's exp
  void setup(Domain& d) override {
    std::vector<double> g(d.N, coeff * 1.0);
    if (ramp != 0.0) {
      const auto& x = d.host_grids.at("pos_x");
      const auto& y = d.host_grids.at("pos_y");
      double xmax = -DBL_MAX, ymax = -DBL_MAX, xmin = DBL_MAX, ymin = DBL_MAX;
      for (double v : x) { xmax = (xmax < v) ? v : xmax; xmin = (v < xmin) ? v : xmin; }
      for (double v : y) { ymax = (ymax < v) ? v : ymax; ymin = (v < ymin) ? v : ymin; }
      const double xc = 0.5 * (xmin + xmax), yc = 0.5 * (ymin + ymax);
      const double ax = xmax - xc, ay = ymax - yc;
      const double sl = ramp / std::min(ax, ay);
      for (size_t c = 0; c < d.N; ++c) {
        const double dx = x[c] - xc, dy = y[c] - yc;
        const double rs = dx * dx / (ax * ax) + dy * dy / (ay * ay);
        double m = rs + 2.0 * sl - 1.0;
        m = (m < 0.0) ? 0.0 : m;
        const double t = m / sl;
        double e = std::exp(-2.3 * (t * t)) - 0.01;
        e = (e < 0.0) ? 0.0 : e;
        g[c] = coeff * (1.0 / 0.99 * e);
      }
    }
    coef.alloc(d.N);
    hip_check(hipMemcpy(coef.p, g.data(), d.N * 8, hipMemcpyHostToDevice), "upload coefficient");
    bhx.alloc(d.N); bhy.alloc(d.N); bhz.alloc(d.N);
    for (auto& b : sa) b.alloc(d.N);
    for (auto& b : sb) b.alloc(d.N);
    if (integ == "rk2") for (auto& b : sh) b.alloc(d.N);
    for (auto& b : dv) b.alloc(d.N);
    part.alloc(pv_blocks(d.mesh));
    tsmin.alloc(1);
    if (to_file) {
      alloc_zero(avg_heat, d.N, d.stream);
      for (auto& b : avg_f) alloc_zero(b, d.N, d.stream);
    }
  }
  void file_output(OutList& o) override {  // :292-310
    if (!to_file) return;
    if (heating_on) o.emplace_back("viscous_heating", avg_heat.p);
    if (force_on) {
      o.emplace_back("viscous_force_x", avg_f[0].p);
      o.emplace_back("viscous_force_y", avg_f[1].p);
      o.emplace_back("viscous_force_z", avg_f[2].p);
    }
  }
  void iterate(Domain& d, double dt) override {  // :148-250
    PvFrozen fz{bhx.p, bhy.p, bhz.p, coef.p, d.L.n.p, d.L.rho.p};
    d.timed("pv_init", [&] { launch_pv_init(d.L.view(), d.stat(), d.kc, d.mesh, bhx.p, bhy.p, bhz.p, view(sa), d.stream); });
    // computeViscousSubcycles (:67-80): epsilon * min(timescale) over the interior
    d.timed("pv_timescale", [&] {
      launch_pv_timescale(view(sa), fz, 0.000001 * coeff, d.mesh, part.p, d.stream);
      launch_reduce_step(part.p, pv_blocks(d.mesh), 1.0, tsmin.p, nullptr, 0, d.err_flag, d.stream);
    });
    double m = 0.0;
    hip_check(hipMemcpyAsync(&m, tsmin.p, 8, hipMemcpyDeviceToHost, d.stream), "read timescale");
    hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    require(m < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (physical viscosity)");
    nsub = subcycle_count(dt, epsilon * m);
    const double h = dt / nsub;
    h_last = h;
    PvDerivs dd{dv[0].p, dv[1].p, dv[2].p, dv[3].p, dv[4].p, dv[5].p, dv[6].p};
    PvArgs a{};
    a.d = dd; a.f = fz; a.k = d.kc; a.mesh = d.mesh; a.h = h;
    a.heating_on = heating_on ? 1 : 0; a.force_on = force_on ? 1 : 0; a.apply = inactive ? 0 : 1;
    a.coeff_zero = coeff == 0.0 ? 1 : 0; a.grad_corr = grad_corr ? 1 : 0;
    a.nsd = (double)nsub;
    if (to_file) {   // the averages restart every step (:151-152)
      launch_fill(avg_heat.p, 0.0, d.N, d.stream);
      for (int q = 0; q < 3; ++q) { launch_fill(avg_f[q].p, 0.0, d.N, d.stream); a.avg_f[q] = avg_f[q].p; }
      a.avg_heat = avg_heat.p;
    }
    if (d.multispecies_on) { a.ms_ion = d.ms(1); a.ms_el = d.ms(0); a.msf = msf; }
    DBuf* cur = sa;
    DBuf* nxt = sb;
    for (int s = 0; s < nsub; ++s) {
      if (recording) record(d, integ == "euler" ? (size_t)s : 2 * (size_t)s, view(cur));
      d.timed("pv_derivs", [&] { launch_pv_derivs(view(cur), d.mesh, dd, d.stream); });
      if (integ == "euler") {
        a.eval = view(cur); a.base = view(cur); a.out = view(nxt); a.mode = PV_FULL;
        d.timed("pv_apply", [&] { launch_pv_apply(a, d.stream); });
      } else {
        a.eval = view(cur); a.base = view(cur); a.out = view(sh); a.mode = PV_HALF;
        d.timed("pv_apply", [&] { launch_pv_apply(a, d.stream); });
        if (recording) record(d, 2 * (size_t)s + 1, view(sh));
        d.timed("pv_derivs", [&] { launch_pv_derivs(view(sh), d.mesh, dd, d.stream); });
        a.eval = view(sh); a.base = view(cur); a.out = view(nxt); a.mode = PV_FULL;
        d.timed("pv_apply", [&] { launch_pv_apply(a, d.stream); });
      }
      std::swap(cur, nxt);
    }
    const size_t bytes = d.N * 8;
    hip_check(hipMemcpyAsync(d.L.te.p, cur[0].p, bytes, hipMemcpyDeviceToDevice, d.stream), "copy te");
    hip_check(hipMemcpyAsync(d.L.mx.p, cur[2].p, bytes, hipMemcpyDeviceToDevice, d.stream), "copy mom_x");
    hip_check(hipMemcpyAsync(d.L.my.p, cur[3].p, bytes, hipMemcpyDeviceToDevice, d.stream), "copy mom_y");
    hip_check(hipMemcpyAsync(d.L.mz.p, cur[4].p, bytes, hipMemcpyDeviceToDevice, d.stream), "copy mom_z");
    d.propagate_live();
  }
};

// ---------------------------------------------------------------------------------------------
// Source-term modules: localized heating, mass and momentum injection, ambient heating sink,
// field heating, divergence cleaning (source/modules/solar/*.cpp)
// ---------------------------------------------------------------------------------------------
using Opts = std::vector<std::pair<std::string, std::string>>;

// SolarUtils::GaussianGrid (solarutils.cpp:65-82), on the host with the C library's exp and pow
static std::vector<double> gaussian_grid(int nx, int ny, double mn, double mx, double sx, double sy, double cx,
                                         double cy) {
  std::vector<double> gx(nx), gy(ny), r((size_t)nx * ny);
  for (int i = 0; i < nx; ++i) gx[i] = std::exp(-0.5 * std::pow(((double)i - cx) / sx, 2.0));
  for (int j = 0; j < ny; ++j) gy[j] = std::exp(-0.5 * std::pow(((double)j - cy) / sy, 2.0));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) r[(size_t)i * ny + j] = (mx - mn) * gx[i] * gy[j] + mn;
  return r;
}
// SolarUtils::GaussianGridRotated (solarutils.cpp:85-98)
static std::vector<double> gaussian_grid_rot(int nx, int ny, double mn, double mx, double sx, double sy, double cx,
                                             double cy, double ang) {
  require(ang > -90.0 && ang < 90.0, SPRUCE_E_ARG, "GaussianGridRotated requires angles between -90 deg and +90 deg");
  if (ang == 0.0) return gaussian_grid(nx, ny, mn, mx, sx, sy, cx, cy);
  std::vector<double> r((size_t)nx * ny);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const double xe = ((double)i - cx) * std::cos(ang * PI / 180.0) - ((double)j - cy) * std::sin(ang * PI / 180.0);
      const double ye = ((double)i - cx) * std::sin(ang * PI / 180.0) + ((double)j - cy) * std::cos(ang * PI / 180.0);
      r[(size_t)i * ny + j] = (mx - mn) * std::exp(-0.5 * std::pow(xe / sx, 2.0)) * std::exp(-0.5 * std::pow(ye / sy, 2.0)) + mn;
    }
  return r;
}
static void vmax_into(std::vector<double>& a, const std::vector<double>& b) {
  for (size_t k = 0; k < a.size(); ++k) a[k] = smax(a[k], b[k]);
}
static void vmin_into(std::vector<double>& a, const std::vector<double>& b) {
  for (size_t k = 0; k < a.size(); ++k) a[k] = smin(a[k], b[k]);
}
// The clipped Gaussian template of localized heating and mass injection with its periodic images
// (localizedheating.cpp:34-49, massinjection.cpp:32-47).
static std::vector<double> injection_template(const Domain& d, double rate, double sx, double sy, double cx, double cy) {
  auto g = [&](double x, double y) {
    std::vector<double> v = gaussian_grid(d.nx, d.ny, -1.0e-2 * rate, rate, sx, sy, x, y);
    for (double& e : v) e = smax(e, 0.0);
    return v;
  };
  std::vector<double> t = g(cx, cy);
  if (d.desc.bc[0] == SPRUCE_BC_PERIODIC) {
    vmax_into(t, g(cx + (double)d.nx, cy));
    vmax_into(t, g(cx - (double)d.nx, cy));
  }
  if (d.desc.bc[2] == SPRUCE_BC_PERIODIC) {
    vmax_into(t, g(cx, cy + (double)d.ny));
    vmax_into(t, g(cx, cy - (double)d.ny));
  }
  return t;
}
static void upload_vec(DBuf& b, const std::vector<double>& v) {
  b.alloc(v.size());
  hip_check(hipMemcpy(b.p, v.data(), v.size() * 8, hipMemcpyHostToDevice), "upload template");
}

// What the adjoint needs of one source module's post-integrator update at simulation time `now`:
// the launch_source mode and templates, and the factor g(now) of the step size in its amplitude
// with its time derivative gp.
struct SourceRev {
  int mode = SRC_ADD;
  const double *t = nullptr, *t2 = nullptr;
  double g = 1.0, gp = 0.0;
};

// Base of the modules that act only after the integrator, at a simulation time window.
struct PostSourceModule : GpuModule {
  bool needs_pre_scalars() const override { return false; }
  bool iterates() const override { return false; }
  void pre_launch(Domain&, double*) override {}
  void pre_finish(double, const double*) override {}
  void iterate(Domain&, double) override {}
  // false when post() at `now` leaves the state alone (outside the module's time window)
  virtual bool source_rev(double now, SourceRev& r) const { (void)now; (void)r; return false; }
};

// LocalizedHeating (localizedheating.cpp)
struct LocalHeatModule : PostSourceModule {
  double start = 0.0, ramp = 0.0, dur = 0.0, rate = 0.0, sx = 1.0, sy = 1.0, cx = 0.0, cy = 0.0, msf = 0.0;
  DBuf tmpl;
  explicit LocalHeatModule(const Opts& o) {  // :14-29
    name = "localized_heating";
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "start_time") start = std::stod(v);
      else if (k == "duration") dur = std::stod(v);
      else if (k == "max_heating_rate") rate = std::stod(v);
      else if (k == "stddev_x") sx = std::stod(v);
      else if (k == "stddev_y") sy = std::stod(v);
      else if (k == "center_x") cx = std::stod(v);
      else if (k == "center_y") cy = std::stod(v);
      else if (k == "ramp_time") ramp = std::stod(v);
      else if (k == "ms_electron_heating_fraction") msf = std::stod(v);
    }
  }
  bool needs_host_dt() const override { return true; }
  // the template depends on positions only: the reference builds it at its first preIterateModule
  // (:31-50), this build at set-up
  void setup(Domain& d) override {
    require(msf >= 0.0 && msf <= 1.0, SPRUCE_E_ARG,
            "Localized Heating MS electron heating fraction must be between 0 and 1");
    upload_vec(tmpl, injection_template(d, rate, sx, sy, cx, cy));
  }
  // the ramp (:56-59) and its derivative in time
  double ramp_factor(double now, double* deriv) const {
    const double t = now - start;
    double rf = 1.0, dr = 0.0;
    if (t < ramp && t <= 0.5 * dur) {
      rf = t / ramp;
      dr = 1.0 / ramp;
    } else if (t > dur - ramp && t > 0.5 * dur) {
      rf = (dur - t) / ramp;
      dr = -1.0 / ramp;
    }
    if (deriv) *deriv = dr;
    return rf;
  }
  bool source_rev(double now, SourceRev& r) const override {
    if (now < start || now > start + dur) return false;
    r.mode = SRC_ADD;
    r.t = tmpl.p;
    r.g = ramp_factor(now, &r.gp);
    return true;
  }
  void post(Domain& d) override {  // :52-66
    const double now = d.tstep;
    if (now < start || now > start + dur) return;
    const double a = d.host_dt * ramp_factor(now, nullptr);
    d.timed("source", [&] { launch_source(d.L.te.p, tmpl.p, nullptr, a, 0.0, nullptr, SRC_ADD, d.mesh, d.stream); });
    d.propagate_live();
    if (d.multispecies_on) {   // (:63-65): the template without the ramp factor
      if (msf < 1.0) launch_ms_source(d.ms(1), tmpl.p, d.host_dt, nullptr, 1.0 - msf, MS_SRC_TEMPLATE, d.mesh, d.stream);
      if (msf > 0.0) launch_ms_source(d.ms(0), tmpl.p, d.host_dt, nullptr, msf, MS_SRC_TEMPLATE, d.mesh, d.stream);
    }
  }
};

// MassInjection (massinjection.cpp)
struct MassInjModule : PostSourceModule {
  double start = 0.0, dur = 0.0, rate = 0.0, sx = 1.0, sy = 1.0, cx = 0.0, cy = 0.0;
  DBuf tmpl;
  explicit MassInjModule(const Opts& o) {  // :14-27
    name = "mass_injection";
    for (auto& kv : o) {
      const std::string& k = kv.first;
      if (k == "start_time") start = std::stod(kv.second);
      else if (k == "duration") dur = std::stod(kv.second);
      else if (k == "max_injection_rate") rate = std::stod(kv.second);
      else if (k == "stddev_x") sx = std::stod(kv.second);
      else if (k == "stddev_y") sy = std::stod(kv.second);
      else if (k == "center_x") cx = std::stod(kv.second);
      else if (k == "center_y") cy = std::stod(kv.second);
    }
  }
  bool needs_host_dt() const override { return true; }
  void setup(Domain& d) override { upload_vec(tmpl, injection_template(d, rate, sx, sy, cx, cy)); }  // :29-48
  bool source_rev(double now, SourceRev& r) const override {
    if (now < start || now > start + dur) return false;
    r.mode = SRC_ADD_SCALED;
    r.t = tmpl.p;
    return true;
  }
  void post(Domain& d) override {  // :49-53
    if (d.tstep < start || d.tstep > start + dur) return;
    d.timed("source", [&] {
      launch_source(d.L.rho.p, tmpl.p, nullptr, d.host_dt, d.kc.m, nullptr, SRC_ADD_SCALED, d.mesh, d.stream);
    });
    d.propagate_live();
  }
};

// MomentumInjection (momentuminjection.cpp)
struct MomInjModule : PostSourceModule {
  double start = 0.0, dur = 0.0, amax = 0.0, sx = 1.0, sy = 1.0, cx = 0.0, cy = 0.0, ang = 0.0, period = 1.0;
  double dir[2] = {0.0, 0.0};
  bool osc = false;
  DBuf tmpl[2];
  explicit MomInjModule(const Opts& o) {  // :17-35
    name = "momentum_injection";
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "start_time") start = std::stod(v);
      else if (k == "duration") dur = std::stod(v);
      else if (k == "max_accel") amax = std::stod(v);
      else if (k == "stddev_x") sx = std::stod(v);
      else if (k == "stddev_y") sy = std::stod(v);
      else if (k == "center_x") cx = std::stod(v);
      else if (k == "center_y") cy = std::stod(v);
      else if (k == "dir_x") dir[0] = std::stod(v);
      else if (k == "dir_y") dir[1] = std::stod(v);
      else if (k == "template_angle") ang = std::stod(v);
      else if (k == "oscillatory") osc = v == "true";
      else if (k == "oscillation_period") period = std::stod(v);
    }
  }
  bool needs_host_dt() const override { return true; }
  void setup(Domain& d) override {  // :37-68
    require(!(dir[0] == 0.0 && dir[1] == 0.0), SPRUCE_E_ARG,
            "Momentum Injection module must be given a nonzero acceleration direction");
    const double mag = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1]);
    dir[0] = dir[0] / mag;
    dir[1] = dir[1] / mag;
    for (int c = 0; c < 2; ++c) {
      const double m = dir[c];
      auto g = [&](double x, double y) { return gaussian_grid_rot(d.nx, d.ny, -1.0e-2 * m, m, sx, sy, x, y, ang); };
      std::vector<double> t = g(cx, cy);
      // the periodic images are folded in with min() for either sign of the component (:47-64)
      if (d.desc.bc[0] == SPRUCE_BC_PERIODIC) {
        vmin_into(t, g(cx + (double)d.nx, cy));
        vmin_into(t, g(cx - (double)d.nx, cy));
      }
      if (d.desc.bc[2] == SPRUCE_BC_PERIODIC) {
        vmin_into(t, g(cx, cy + (double)d.ny));
        vmin_into(t, g(cx, cy - (double)d.ny));
      }
      for (double& e : t) e = m > 0.0 ? smax(e, 0.0) : smin(e, 0.0);
      upload_vec(tmpl[c], t);
    }
  }
  bool source_rev(double now, SourceRev& r) const override {
    if (now < start || now > start + dur) return false;
    r.mode = SRC_ADD_WEIGHTED;
    r.t = tmpl[0].p;
    r.t2 = tmpl[1].p;
    r.g = (osc ? std::sin(2.0 * PI * (now - start) / period) : 1.0) * amax;   // as post() forms it
    r.gp = osc ? amax * (2.0 * PI / period) * std::cos(2.0 * PI * (now - start) / period) : 0.0;
    return true;
  }
  void post(Domain& d) override {  // :70-76
    if (d.tstep < start || d.tstep > start + dur) return;
    const double of = osc ? std::sin(2.0 * PI * (d.tstep - start) / period) : 1.0;
    const double a = d.host_dt * (of * amax);
    d.timed("source", [&] {
      launch_source(d.L.mx.p, tmpl[0].p, d.L.rho.p, a, 0.0, nullptr, SRC_ADD_WEIGHTED, d.mesh, d.stream);
      launch_source(d.L.my.p, tmpl[1].p, d.L.rho.p, a, 0.0, nullptr, SRC_ADD_WEIGHTED, d.mesh, d.stream);
    });
    d.propagate_live();
  }
};

// AmbientHeatingSink (ambientheatingsink.cpp)
struct HeatSinkModule : PostSourceModule {
  double rate = 0.0, exp_base = 0.0, exp_h = 1.0, cx = 0.0, hw = 1.0, msf = 0.5;   // msf: ambientheatingsink.hpp:28
  bool exp_mode = false;
  DBuf red;
  explicit HeatSinkModule(const Opts& o) {  // :12-25
    name = "ambient_heating_sink";
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "heating_rate") rate = std::stod(v);
      if (k == "exp_mode") exp_mode = v == "true";
      else if (k == "exp_base_heating_rate") exp_base = std::stod(v);
      else if (k == "exp_scale_height") exp_h = std::stod(v);
      else if (k == "center_x") cx = std::stod(v);
      else if (k == "half_width") hw = std::stod(v);
      else if (k == "ms_electron_heating_fraction") msf = std::stod(v);
    }
  }
  bool needs_host_dt() const override { return false; }
  // the reduction depends on positions only: built once on the host with the C library's exp (:27-33)
  void setup(Domain& d) override {
    const auto& x = d.host_grids.at("pos_x");
    const auto& y = d.host_grids.at("pos_y");
    std::vector<double> r(d.N);
    for (int i = 0; i < d.nx; ++i)
      for (int j = 0; j < d.ny; ++j) {
        const size_t c = (size_t)i * d.ny + j;
        const double mask = (i >= d.xl && i <= d.xu && j >= d.yl && j <= d.yu) ? 1.0 : 0.0;
        if (exp_mode) {
          const double w = (x[c] - cx) / hw;
          r[c] = mask * exp_base * std::exp(-1.0 * y[c] / exp_h) * smax(1.0 - w * w, 0.0);
        } else {
          r[c] = mask * rate;
        }
      }
    require(msf >= 0.0 && msf <= 1.0, SPRUCE_E_ARG,
            "Ambient Heating Sink MS electron heating fraction must be between 0 and 1");
    upload_vec(red, r);
  }
  bool source_rev(double, SourceRev& r) const override {
    r.mode = SRC_SUB;
    r.t = red.p;
    return true;
  }
  void post(Domain& d) override {  // :35-37
    d.timed("source", [&] { launch_source(d.L.te.p, red.p, nullptr, 0.0, 0.0, d.step.p, SRC_SUB, d.mesh, d.stream); });
    d.propagate_live();
    if (d.multispecies_on) {   // (:38-41)
      if (msf < 1.0) launch_ms_source(d.ms(1), red.p, 0.0, d.step.p, 1.0 - msf, MS_SRC_SINK, d.mesh, d.stream);
      if (msf > 0.0) launch_ms_source(d.ms(0), red.p, 0.0, d.step.p, msf, MS_SRC_SINK, d.mesh, d.stream);
    }
  }
};

// FieldHeating (fieldheating.cpp)
struct FieldHeatModule : GpuModule {
  FheatParams p{};
  bool inactive = false, to_file = false;
  DBuf heat;   // after a step: mask * (dt * heating), the energy the step added
  explicit FieldHeatModule(const Opts& o) {  // :15-28
    name = "field_heating";
    p.cc = C_LIGHT / (4.0 * PI);
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "output_to_file") to_file = v == "true";
      else if (k == "coeff") p.coeff = std::stod(v);
      else if (k == "current_pow") p.cpow = std::stod(v);
      else if (k == "b_pow") p.bpow = std::stod(v);
      else if (k == "n_pow") p.npow = std::stod(v);
      else if (k == "roc_pow") p.rpow = std::stod(v);
      else if (k == "inactive_mode") inactive = v == "true";
    }
  }
  bool needs_pre_scalars() const override { return false; }
  bool needs_host_dt() const override { return false; }   // the step size is read on the device
  bool needs_bhat() const override { return p.coeff != 0.0 && p.rpow != 0.0; }
  void setup(Domain& d) override { alloc_zero(heat, d.N, d.stream); }   // :30-32
  void file_output(OutList& o) override {                                  // :73-80
    if (to_file) o.emplace_back("field_heating", heat.p);
  }
  void pre_launch(Domain& d, double*) override {  // preIterateModule -> computeHeating (:34-54)
    d.timed("fheat", [&] { launch_fheat_pre(d.L.view(), d.stat(), d.bhx.p, d.bhy.p, p, d.mesh, heat.p, d.stream); });
  }
  void pre_finish(double, const double*) override {}
  void iterate(Domain& d, double) override {  // :56-61
    d.timed("fheat", [&] { launch_fheat_apply(d.L.te.p, heat.p, d.step.p, d.mesh, inactive ? 0 : 1, d.stream); });
    if (!inactive) d.propagate_live();
  }
};

// DivCleaning (divcleaning.cpp)
struct DivCleanModule : GpuModule {
  double eps = 0.1, ts = 1.0, h = 0.0;
  DBuf coeff, bx[2], by[2], dxb, dyb;
  explicit DivCleanModule(const Opts& o) {  // :12-20
    name = "div_cleaning";
    for (auto& kv : o) {
      if (kv.first == "epsilon") eps = std::stod(kv.second);
      if (kv.first == "time_scale") ts = std::stod(kv.second);
    }
  }
  bool needs_pre_scalars() const override { return false; }
  void pre_launch(Domain&, double*) override {}
  void pre_finish(double, const double*) override {}
  // coeff = (1 / (1/dx^2 + 1/dy^2)) / 2 / time_scale, on the host (:22-26)
  void setup(Domain& d) override {
    const auto& gx = d.host_grids.at("d_x");
    const auto& gy = d.host_grids.at("d_y");
    std::vector<double> c(d.N);
    for (size_t k = 0; k < d.N; ++k) c[k] = (1.0 / (1.0 / (gx[k] * gx[k]) + 1.0 / (gy[k] * gy[k]))) / 2. / ts;
    upload_vec(coeff, c);
    for (auto* b : {&bx[0], &bx[1], &by[0], &by[1], &dxb, &dyb}) b->alloc(d.N);
  }
  // the subcycle count is fixed by the step size (computeNumSubcycles, :48-50); the cleaning itself
  // runs after the integrator
  void iterate(Domain&, double dt) override {
    nsub = subcycle_count(dt, eps * ts);
    h = dt / ((double)nsub);
  }
  // while recording (the adjoint's replay of a step): the total fields at the start of every
  // subcycle (bx, by; 2 grids each)
  bool recording = false;
  std::vector<DBuf> tape;
  void post(Domain& d) override {  // :28-46
    int cur = 0;
    d.timed("divclean", [&] {
      launch_field_sum(d.L.bx.p, d.L.by.p, d.bex.p, d.bey.p, d.N, bx[0].p, by[0].p, d.stream);
      for (int s = 0; s < nsub; ++s) {
        if (recording) {
          while (tape.size() <= (size_t)s) tape.emplace_back();
          if (tape[s].n != 2 * d.N) tape[s].alloc(2 * d.N);
          hip_check(hipMemcpyAsync(tape[s].p, bx[cur].p, d.N * 8, hipMemcpyDeviceToDevice, d.stream), "record");
          hip_check(hipMemcpyAsync(tape[s].p + d.N, by[cur].p, d.N * 8, hipMemcpyDeviceToDevice, d.stream), "record");
        }
        launch_dc_inner(bx[cur].p, by[cur].p, d.mesh, dxb.p, dyb.p, d.stream);
        launch_dc_update(bx[cur].p, by[cur].p, dxb.p, dyb.p, d.L.bx.p, d.L.by.p, d.bex.p, d.bey.p, coeff.p, h, d.mesh,
                         bx[1 - cur].p, by[1 - cur].p, d.stream);
        cur = 1 - cur;
      }
    });
    d.propagate_live();
  }
};

// ParametricHeating: heating(y) = sum_k theta_k phi_k(y), the parametrised heating profile of the
// gradient workload (no reference module; injected as AmbientHeating::postIterateModule does,
// ambientheating.cpp:42-49).  phi_k are Gaussians in height centred on K points spread evenly over
// [min pos_y, max pos_y], of width basis_width times their spacing.  The adjoint returns dJ/dtheta.
struct ParamHeatModule : PostSourceModule {
  int K = 16;
  double width = 1.0;
  std::vector<double> theta;
  std::vector<double> basis_y;   // K x ny: phi_k at the height of row j
  DBuf heating, basis, g_theta;  // g_theta: dJ/dtheta of the last adjoint
  explicit ParamHeatModule(const Opts& o) {
    name = "parametric_heating";
    std::vector<std::pair<int, double>> th;
    for (auto& kv : o) {
      if (kv.first == "num_basis") K = std::stoi(kv.second);
      else if (kv.first == "basis_width") width = std::stod(kv.second);
      else if (kv.first.rfind("coeff_", 0) == 0) th.emplace_back(std::stoi(kv.first.substr(6)), std::stod(kv.second));
    }
    require(K >= 1 && K <= 64, SPRUCE_E_ARG, "parametric_heating: num_basis must be in [1, 64]");
    theta.assign(K, 0.0);
    for (auto& t : th) {
      require(t.first >= 0 && t.first < K, SPRUCE_E_ARG, "parametric_heating: coefficient index out of range");
      theta[t.first] = t.second;
    }
  }
  bool needs_host_dt() const override { return false; }
  void setup(Domain& d) override {
    const auto& y = d.host_grids.at("pos_y");
    // the basis (and dJ/dtheta) is a function of the row height: every column must share it
    for (int i = 1; i < d.nx; ++i)
      for (int j = 0; j < d.ny; ++j)
        require(y[(size_t)i * d.ny + j] == y[j], SPRUCE_E_ARG,
                "parametric_heating: pos_y must not vary along x (the heating profile is a function of height)");
    double lo = y[0], hi = y[0];
    for (double v : y) { lo = std::min(lo, v); hi = std::max(hi, v); }
    const double sp = K > 1 ? (hi - lo) / (double)(K - 1) : hi - lo;
    const double w = width * sp;
    std::vector<double> h(d.N);
    basis_y.assign((size_t)K * d.ny, 0.0);
    for (int i = 0; i < d.nx; ++i)
      for (int j = 0; j < d.ny; ++j) {
        const size_t c = (size_t)i * d.ny + j;
        const double mask = (i >= d.xl && i <= d.xu && j >= d.yl && j <= d.yu) ? 1.0 : 0.0;
        double hsum = 0.0;
        for (int q = 0; q < K; ++q) {
          const double u = (y[c] - (lo + (double)q * sp)) / w;
          const double phi = std::exp(-0.5 * (u * u));
          hsum += theta[q] * phi;
          if (i == 0) basis_y[(size_t)q * d.ny + j] = phi;   // pos_y varies along y only
        }
        h[c] = mask * hsum;
      }
    upload_vec(heating, h);
    upload_vec(basis, basis_y);
    g_theta.alloc(K);
  }
  void post(Domain& d) override {
    d.timed("heating", [&] { launch_add_step_times(d.L.te.p, heating.p, d.step.p, d.N, d.stream); });
    d.propagate_live();
  }
};

// AnomalousResistivity (anomalousresistivity.cpp)
struct ResistModule : GpuModule {
  std::string integ = "euler", tmode = "flood_fill", model = "time_scale";
  double safety = 1.0, frob_coeff = 1.0e50, ff_threshold = 1.0, ff_ramp = 1.0e-5, ff_argmin_radius = 5.0e9;
  double sigma = 3.0, ts = -1.0, ff_max_radius = -1.0, ff_min_current = -1.0;
  bool smoothing = true, grad_corr = false, to_file = false;
  std::vector<double> params;
  DBuf avg;                                             // output_to_file: the Joule heating rate of the last step
  int null_i = 0, null_j = 0, kr = 0;
  int nx = 0, ny = 0;
  std::vector<double> px, py;                           // host positions (the argmin box)
  DBuf dpx, dpy, kern, J, bx, by, bm, ratio, tmpl, tmp, diff, coeff, td, lap[3], work[4], d1[4], d2[4], d3[4], d4[4],
      mid[3], part, scal, argout;
  int* changed = nullptr;
  static constexpr double CC = C_LIGHT / (4.0 * PI);
  explicit ResistModule(const Opts& o) {  // :51-77
    name = "anomalous_resistivity";
    for (auto& kv : o) {
      const std::string& a = kv.first;
      const std::string& v = kv.second;
      if (a == "time_scale") ts = std::stod(v);
      else if (a == "output_to_file") to_file = v == "true";
      else if (a == "frobenius_metric_coeff") frob_coeff = std::stod(v);
      else if (a == "smoothing_sigma") sigma = std::stod(v);
      else if (a == "safety_factor") safety = std::stod(v);
      else if (a == "metric_smoothing") smoothing = v == "true";
      else if (a == "time_integrator") integ = v;
      else if (a == "template_mode") tmode = v;
      else if (a == "flood_fill_max_radius") ff_max_radius = std::stod(v);
      else if (a == "flood_fill_argmin_radius") ff_argmin_radius = std::stod(v);
      else if (a == "flood_fill_min_current") ff_min_current = std::stod(v);
      else if (a == "flood_fill_current_ramp_length") ff_ramp = std::stod(v);
      else if (a == "resistivity_model") model = v;
      else if (a == "gradient_correction") grad_corr = v == "true";
      else if (a == "resistivity_model_params") {
        params.clear();
        std::stringstream ss(v);
        std::string t;
        while (std::getline(ss, t, ',')) params.push_back(std::stod(t));
      } else if (a == "flood_fill_threshold") ff_threshold = std::stod(v);
    }
    if (integ.empty()) integ = "euler";
  }
  ~ResistModule() override { if (changed) (void)hipFree(changed); }
  bool needs_pre_scalars() const override { return false; }
  void pre_launch(Domain&, double*) override {}
  void pre_finish(double, const double*) override {}

  ArParams ar_params() const {
    ArParams p{};
    p.model = model == "time_scale" ? AR_TIME_SCALE : (model == "syntelis_19" ? AR_SYNTELIS : AR_YS94);
    p.ts = ts;
    if (p.model != AR_TIME_SCALE) { p.p0 = params[0]; p.p1 = params[1]; p.p2 = params[2]; }
    return p;
  }
  // argminLocalized (:202-225): box on the host from the positions, minimum on the device
  void argmin_local(Domain& d) {
    const int ci = null_i, cj = null_j;
    auto X = [&](int i, int j) { return px[(size_t)i * ny + j]; };
    auto Y = [&](int i, int j) { return py[(size_t)i * ny + j]; };
    const double r = ff_argmin_radius;
    int il = ci, iu = ci, jl = cj, ju = cj;
    for (; il >= d.xl; il--) if (std::abs(X(ci, cj) - X(il, cj)) > r) break;
    for (; iu <= d.xu; iu++) if (std::abs(X(ci, cj) - X(iu, cj)) > r) break;
    for (; jl >= d.yl; jl--) if (std::abs(Y(ci, cj) - Y(ci, jl)) > r) break;
    for (; ju <= d.yu; ju++) if (std::abs(Y(ci, cj) - Y(ci, ju)) > r) break;
    require(il >= 0 && jl >= 0 && iu < nx && ju < ny, SPRUCE_E_UNSUPPORTED,
            "anomalous_resistivity: the localized argmin box leaves the grid (undefined in the reference)");
    launch_ar_argmin_box(bm.p, dpx.p, dpy.p, ci, cj, r, il, iu, jl, ju, ny, argout.p, d.stream);
    double o[3];
    hip_check(hipMemcpyAsync(o, argout.p, sizeof o, hipMemcpyDeviceToHost, d.stream), "read argmin");
    hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    require(std::min(o[0], o[2]) < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the localized argmin calculation");
    if (o[1] >= 0.0 && o[0] < o[2]) {
      const long long c = (long long)o[1];
      null_i = (int)(c / ny);
      null_j = (int)(c % ny);
    }
  }
  // computeTemplate (:158-199) for the field be + bi and the current density Jc
  void compute_template(Domain& d, const double* bix, const double* biy, const double* Jc) {
    launch_ar_field(d.bex.p, d.bey.p, bix, biy, d.N, bx.p, by.p, bm.p, d.stream);
    if (tmode == "frobenius") {
      launch_ar_frob(bx.p, by.p, bm.p, frob_coeff, d.mesh, ratio.p, tmpl.p, d.stream);
    } else {
      argmin_local(d);
      const size_t start = (size_t)null_i * ny + null_j;
      launch_ar_flood_seed(bm.p, d.N, start, ff_threshold, tmpl.p, d.stream);
      for (;;) {                                        // sweeps until the region stops growing
        hip_check(hipMemsetAsync(changed, 0, sizeof(int), d.stream), "hipMemsetAsync");
        for (int k = 0; k < 8; ++k) launch_ar_flood_sweep(bm.p, ff_threshold, d.mesh, tmpl.p, changed, d.stream);
        int c = 0;
        hip_check(hipMemcpyAsync(&c, changed, sizeof(int), hipMemcpyDeviceToHost, d.stream), "read flood flag");
        hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
        if (!c) break;
      }
      if (ff_max_radius > 0.0 || ff_min_current > 0.0)
        launch_ar_masks(tmpl.p, Jc, dpx.p, dpy.p, start, ff_max_radius, ff_min_current, ff_ramp, d.mesh, d.stream);
    }
    if (smoothing) {
      launch_ar_smooth(tmpl.p, kern.p, kr, d.mesh, tmp.p, d.stream);
      tmpl.swap(tmp);
    }
    if (tmode == "frobenius") launch_ar_frob_final(tmpl.p, d.N, d.stream);
  }
  void compute_diffusion(Domain& d, const double* bix, const double* biy, const double* biz) {  // :124-156
    if (model != "time_scale") launch_ar_current(bix, biy, biz, CC, d.mesh, J.p, d.stream);
    launch_ar_diffusion(J.p, d.L.n.p, ar_params(), d.mesh, diff.p, d.stream);
  }
  // computeTimeDerivatives (:79-102) into out[4]
  void derivs(Domain& d, const double* bix, const double* biy, const double* biz, DBuf* out) {
    launch_ar_current(bix, biy, biz, CC, d.mesh, J.p, d.stream);
    compute_template(d, bix, biy, J.p);
    compute_diffusion(d, bix, biy, biz);                  // the same field: J is recomputed unchanged
    launch_ar_coeff(tmpl.p, diff.p, d.mesh, coeff.p, td.p, d.stream);
    ArRhsArgs a{};
    a.bix = bix; a.biy = biy; a.biz = biz; a.coeff = coeff.p; a.J = J.p;
    a.lbx = lap[0].p; a.lby = lap[1].p; a.lbz = lap[2].p;
    a.d0 = out[0].p; a.d1 = out[1].p; a.d2 = out[2].p; a.d3 = out[3].p;
    a.eta_factor = 4.0 * PI / C_LIGHT / C_LIGHT;
    a.grad_corr = grad_corr ? 1 : 0;
    a.mesh = d.mesh;
    launch_ar_rhs(a, d.stream);
  }
  void setup(Domain& d) override {  // :22-49
    require(model != "time_scale" || ts > 0.0, SPRUCE_E_ARG,
            "Anomalous Resistivity time_scale must be specified, and positive, if run in time_scale mode");
    require(integ == "euler" || integ == "rk2" || integ == "rk4", SPRUCE_E_ARG,
            "Invalid time integrator given for Anomalous Resistivity module");
    require(tmode == "frobenius" || tmode == "flood_fill", SPRUCE_E_ARG, "Invalid template mode given for Anomalous Resistivity module");
    require(model == "time_scale" || model == "syntelis_19" || model == "ys_94", SPRUCE_E_ARG,
            "Invalid resistivity model given for Anomalous Resistivity module");
    require(model == "time_scale" || params.size() == 3, SPRUCE_E_ARG, "the resistivity model requires three parameters");
    nx = d.nx;
    ny = d.ny;
    if (tmode == "flood_fill" && (ff_max_radius > 0.0 || ff_min_current > 0.0))
      require(nx <= ny, SPRUCE_E_UNSUPPORTED,
              "anomalous_resistivity masks need xdim <= ydim (the reference loops columns over the row count)");
    px = d.host_grids.at("pos_x");
    py = d.host_grids.at("pos_y");
    upload_vec(dpx, px);
    upload_vec(dpy, py);
    for (DBuf* b : {&J, &bx, &by, &bm, &ratio, &tmpl, &tmp, &coeff, &td}) b->alloc(d.N);
    alloc_zero(diff, d.N, d.stream);                    // diffusivity is zero until the first step (:20)
    if (to_file) alloc_zero(avg, d.N, d.stream);
    for (int k = 0; k < 3; ++k) { lap[k].alloc(d.N); mid[k].alloc(d.N); }
    for (int k = 0; k < 4; ++k) {
      work[k].alloc(d.N); d1[k].alloc(d.N);
      if (integ != "euler") d2[k].alloc(d.N);
      if (integ == "rk4") { d3[k].alloc(d.N); d4[k].alloc(d.N); }
    }
    part.alloc(2 * (size_t)ar_scale_blocks(d.mesh));
    scal.alloc(4);
    argout.alloc(3);
    hip_check(hipMalloc(&changed, sizeof(int)), "hipMalloc");
    launch_ar_laplacian(d.bex.p, d.mesh, lap[0].p, d.stream);
    launch_ar_laplacian(d.bey.p, d.mesh, lap[1].p, d.stream);
    launch_ar_laplacian(d.bez.p, d.mesh, lap[2].p, d.stream);
    // null_location = argmin of |b| in the plane over the interior (Grid::argmin, grid.cpp:86-100)
    launch_ar_field(d.bex.p, d.bey.p, d.L.bx.p, d.L.by.p, d.N, bx.p, by.p, bm.p, d.stream);
    std::vector<double> hb(d.N);
    hip_check(hipMemcpyAsync(hb.data(), bm.p, d.N * 8, hipMemcpyDeviceToHost, d.stream), "read |b|");
    hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    int bi = (int)(0.5 * nx), bj = (int)(0.5 * ny);
    double m = hb[(size_t)bi * ny + bj];
    for (int i = d.xl; i <= d.xu; ++i)
      for (int j = d.yl; j <= d.yu; ++j)
        if (hb[(size_t)i * ny + j] < m) { m = hb[(size_t)i * ny + j]; bi = i; bj = j; }
    require(m < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the argmin calculation");
    null_i = bi;
    null_j = bj;
    if (smoothing) {
      require(sigma > 0.0, SPRUCE_E_ARG, "smoothing_sigma must be a positive number");
      kr = (int)std::nearbyint(4.0 * sigma);
      const int ks = 2 * kr + 1;
      upload_vec(kern, gaussian_grid(ks, ks, 0.0, 1.0 / (2.0 * PI * sigma * sigma), sigma, sigma, (double)kr, (double)kr));
    }
    launch_ar_current(d.L.bx.p, d.L.by.p, d.L.bz.p, CC, d.mesh, J.p, d.stream);
    compute_template(d, d.L.bx.p, d.L.by.p, J.p);
    launch_ar_coeff(tmpl.p, diff.p, d.mesh, coeff.p, td.p, d.stream);   // td = template * diffusivity
    hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
  }
  // fileOutput (:320-329): template * diffusivity and the template of the last derivative
  // evaluation, and the Joule heating rate of the last step
  void file_output(OutList& o) override {
    if (!to_file) return;
    o.emplace_back("anomalous_diffusivity", td.p);
    o.emplace_back("anomalous_template", tmpl.p);
    o.emplace_back("joule_heating", avg.p);
  }
  void iterate(Domain& d, double dt) override {  // :104-122, 170-220
    const size_t bytes = d.N * 8;
    double* L4[4] = {d.L.bx.p, d.L.by.p, d.L.bz.p, d.L.te.p};
    for (int k = 0; k < 4; ++k)
      hip_check(hipMemcpyAsync(work[k].p, L4[k], bytes, hipMemcpyDeviceToDevice, d.stream), "copy field");
    d.timed("resist", [&] {
      launch_ar_current(work[0].p, work[1].p, work[2].p, CC, d.mesh, J.p, d.stream);
      compute_template(d, work[0].p, work[1].p, J.p);
      compute_diffusion(d, work[0].p, work[0].p, work[2].p);   // the reference passes bi_x for bi_y (:107)
      launch_ar_coeff(tmpl.p, diff.p, d.mesh, coeff.p, td.p, d.stream);
      // computeNumSubcycles (:105-122)
      // the time scale's minimum runs over the dt bounds (anomalousresistivity.cpp:191)
      launch_ar_scales(td.p, d.mesh, d.mesh.x.dlo, d.mesh.x.dhi, d.mesh.y.dlo, d.mesh.y.dhi, part.p, scal.p, d.stream);
      launch_mhd_dt(d.L.view(), d.stat(), d.kc, d.mesh, nullptr, d.partial.p, d.stream);
      launch_reduce_step(d.partial.p, d.dt_blocks, d.kc.eps, scal.p + 2, nullptr, 0, d.err_flag, d.stream);
    });
    double sc[3];
    hip_check(hipMemcpyAsync(sc, scal.p, sizeof sc, hipMemcpyDeviceToHost, d.stream), "read scales");
    hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    if (!(sc[0] > 0.0)) {
      nsub = 1;
    } else {
      if (model != "time_scale") {
        require(sc[1] < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (anomalous resistivity)");
        ts = sc[1];
      }
      const double q = sc[2] / (safety * ts);
      require(q < 1.0e8, SPRUCE_E_NUMERIC, "subcycle count out of range");
      nsub = (int)(1.0 + q);
    }
    const double h = dt / ((double)nsub);
    auto axpy = [&](DBuf* base, double sfac, DBuf* k1, DBuf* out, int nf) {
      ArAxpyArgs a{};
      for (int q = 0; q < nf; ++q) { a.base[q] = base[q].p; a.k1[q] = k1[q].p; a.out[q] = out[q].p; }
      a.s = sfac; a.nf = nf; a.rk4 = 0; a.n = d.N;
      launch_ar_axpy(a, d.stream);
    };
    d.timed("resist", [&] {
      for (int s = 0; s < nsub; ++s) {
        if (integ == "euler") {
          derivs(d, work[0].p, work[1].p, work[2].p, d1);
          axpy(work, h, d1, work, 4);
        } else if (integ == "rk2") {
          derivs(d, work[0].p, work[1].p, work[2].p, d1);
          axpy(work, 0.5 * h, d1, mid, 3);
          derivs(d, mid[0].p, mid[1].p, mid[2].p, d2);
          axpy(work, h, d2, work, 4);
        } else {
          derivs(d, work[0].p, work[1].p, work[2].p, d1);
          axpy(work, 0.5 * h, d1, mid, 3);
          derivs(d, mid[0].p, mid[1].p, mid[2].p, d2);
          axpy(work, 0.5 * h, d2, mid, 3);
          derivs(d, mid[0].p, mid[1].p, mid[2].p, d3);
          axpy(work, h, d3, mid, 3);
          derivs(d, mid[0].p, mid[1].p, mid[2].p, d4);
          ArAxpyArgs a{};
          for (int q = 0; q < 4; ++q) {
            a.base[q] = work[q].p; a.k1[q] = d1[q].p; a.k2[q] = d2[q].p; a.k3[q] = d3[q].p; a.k4[q] = d4[q].p;
            a.out[q] = work[q].p;
          }
          a.s = h; a.nf = 4; a.rk4 = 1; a.n = d.N;
          launch_ar_axpy(a, d.stream);
        }
      }
    });
    if (to_file) launch_out_rate(work[3].p, d.L.te.p, dt, d.N, avg.p, d.stream);          // (:167)
    if (d.multispecies_on) launch_ms_change(d.ms(2), work[3].p, d.L.te.p, 1.0, false, d.N, d.stream);  // (:168-170)
    for (int k = 0; k < 4; ++k)
      hip_check(hipMemcpyAsync(L4[k], work[k].p, bytes, hipMemcpyDeviceToDevice, d.stream), "copy field");
    d.propagate_live();
  }
};

// BoundaryOutflow (boundaryoutflow.cpp)
struct OutflowModule : PostSourceModule {
  OutflowParams p{};
  double fall = 0.0, feather = 0.0;
  std::string side = "y_bound_2", shape = "exp";
  DBuf tmpl, part, sc;
  explicit OutflowModule(const Opts& o) {  // :16-31
    name = "boundary_outflow";
    p.dtime = 1.0;
    for (auto& kv : o) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      if (k == "max_accel") p.amax = std::stod(v);
      else if (k == "falloff_length") fall = std::stod(v);
      else if (k == "boundary") side = v;
      else if (k == "falloff_shape") shape = v;
      else if (k == "feather_length") feather = std::stod(v);
      else if (k == "field_aligned_mode") p.aligned = v == "true" ? 1 : 0;
      else if (k == "dynamic_mode") p.dynamic = v == "true" ? 1 : 0;
      else if (k == "dynamic_time") p.dtime = std::stod(v);
      else if (k == "dynamic_target_speed") p.target = std::stod(v);
    }
  }
  bool needs_host_dt() const override { return false; }  // the step size is read on the device
  bool xside() const { return p.side == OF_X1 || p.side == OF_X2; }
  // the interior, extended to the dt bound on the outflow side (:80-86, :144-149)
  void bounds(const Domain& d, int& xl, int& xu, int& yl, int& yu) const {
    xl = d.xl; xu = d.xu; yl = d.yl; yu = d.yu;
    if (p.side == OF_X1) xl = d.mesh.x.dlo;
    else if (p.side == OF_X2) xu = d.mesh.x.dhi;
    else if (p.side == OF_Y2) yu = d.mesh.y.dhi;
    else yl = d.mesh.y.dlo;
  }
  // constructBoundaryAccel (:78-138): depends on positions only, built once with the C library's exp
  std::vector<double> build(const Domain& d) const {
    const auto& x = d.host_grids.at("pos_x");
    const auto& y = d.host_grids.at("pos_y");
    auto lo = [](const std::vector<double>& a) { double m = a[0]; for (double e : a) m = (e < m) ? e : m; return m; };
    auto hi = [](const std::vector<double>& a) { double m = a[0]; for (double e : a) m = (m < e) ? e : m; return m; };
    const double xmin = lo(x), xmax = hi(x), ymin = lo(y), ymax = hi(y), len = fall, fe = feather;
    int bxl, bxu, byl, byu;
    bounds(d, bxl, bxu, byl, byu);
    std::vector<double> r(d.N, 0.0);
    for (int i = 0; i < d.nx; ++i)
      for (int j = 0; j < d.ny; ++j) {
        const size_t c = (size_t)i * d.ny + j;
        const double X = x[c], Y = y[c];
        const bool in = i >= bxl && i <= bxu && j >= byl && j <= byu;
        double t = 0.0;
        if (shape == "exp") {
          if (p.side == OF_X1) t = std::exp(-2.3 * (X - xmin) / len);
          else if (p.side == OF_X2) t = std::exp(2.3 * (X - xmax) / len);
          else if (p.side == OF_Y2) t = std::exp(2.3 * (Y - ymax) / len);
          else t = std::exp(-2.3 * (Y - ymin) / len);
        } else if (shape == "gaussian") {
          const double u = p.side == OF_X1 ? (X - xmin) / len : p.side == OF_X2 ? (-X + xmax) / len
                         : p.side == OF_Y2 ? (-Y + ymax) / len : (Y - ymin) / len;
          t = std::exp(-2.3 * (u * u));
        } else {  // flat: 1 within the falloff length of the boundary, on the interior
          const double e = p.side == OF_X1 ? xmin : p.side == OF_X2 ? xmax : p.side == OF_Y2 ? ymax : ymin;
          if (in && std::fabs((xside() ? X : Y) - e) <= len) t = 1.0;
        }
        if (fe > 0.0) {  // feathering along the boundary (:128-135)
          const double q = xside() ? Y : X, qmax = xside() ? ymax : xmax, qmin = xside() ? ymin : xmin;
          const double a = smax(q - (qmax - 2.0 * fe), 0.0) / fe;
          const double b = smin(q - (qmin + 2.0 * fe), 0.0) / fe;
          t = t * smax(std::exp(-2.3 * (a * a)) * std::exp(-2.3 * (b * b)) - 0.01, 0.0);
        }
        r[c] = (in ? 1.0 : 0.0) * t;
      }
    return r;
  }
  // the cells computeMeanOutflow scans (:142-212): the interior, less the feathered ends along the
  // boundary and cut off at the falloff length from it
  void window(const Domain& d) {
    const auto& x = d.host_grids.at("pos_x");
    const auto& y = d.host_grids.at("pos_y");
    auto X = [&](int i) { return x[(size_t)i * d.ny]; };
    auto Y = [&](int j) { return y[(size_t)j]; };
    int xl, xu, yl, yu;
    bounds(d, xl, xu, yl, yu);
    if (xside()) {
      for (int j = yl; j <= yu; ++j) if (Y(j) - Y(yl) >= feather) { yl = j; break; }
      for (int j = yu; j >= yl; --j) if (Y(yu) - Y(j) >= feather) { yu = j; break; }
      if (p.side == OF_X1) { for (int i = xl; i <= xu; ++i) if (X(i) - X(xl) >= fall) { xu = i; break; } }
      else { for (int i = xu; i >= xl; --i) if (X(xu) - X(i) >= fall) { xl = i; break; } }
    } else {
      for (int i = xl; i <= xu; ++i) if (X(i) - X(xl) >= feather) { xl = i; break; }
      for (int i = xu; i >= xl; --i) if (X(xu) - X(i) >= feather) { xu = i; break; }
      if (p.side == OF_Y1) { for (int j = yl; j <= yu; ++j) if (Y(j) - Y(yl) >= fall) { yu = j; break; } }
      else { for (int j = yu; j >= yl; --j) if (Y(yu) - Y(j) >= fall) { yl = j; break; } }
    }
    p.wxl = xl; p.wxu = xu; p.wyl = yl; p.wyu = yu;
  }
  void setup(Domain& d) override {  // :33-37
    require(shape == "exp" || shape == "gaussian" || shape == "flat", SPRUCE_E_ARG,
            "BoundaryOutflow shape must be exp or gaussian or flat");
    if (side == "x_bound_1") p.side = OF_X1;
    else if (side == "x_bound_2") p.side = OF_X2;
    else if (side == "y_bound_1") p.side = OF_Y1;
    else if (side == "y_bound_2") p.side = OF_Y2;
    else throw Error(SPRUCE_E_ARG, "BoundaryOutflow boundary config must be {x,y}_bound_{1,2}");
    upload_vec(tmpl, build(d));
    window(d);
    require(p.wxl <= p.wxu && p.wyl <= p.wyu, SPRUCE_E_ARG, "boundary_outflow: empty interior");
    part.alloc(2 * (size_t)outflow_blocks(p));
    sc.alloc(2);   // mean outflow, acceleration (0 until the first step)
    d.timed("outflow", [&] { launch_outflow_max(d.L.view(), d.stat(), p, d.mesh, part.p, sc.p, 0, d.stream); });
  }
  void post(Domain& d) override {  // :39-67
    d.timed("outflow", [&] {
      launch_outflow_max(d.L.view(), d.stat(), p, d.mesh, part.p, sc.p, 1, d.stream);
      launch_outflow_apply(d.L.view(), d.stat(), tmpl.p, sc.p, d.step.p, p, d.mesh, d.stream);
    });
    d.propagate_live();
  }
  std::vector<double> report(Domain& d) override {  // commandLineMessage (:69-74)
    std::vector<double> v(2);
    hip_check(hipMemcpyAsync(v.data(), sc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, d.stream), "read outflow");
    hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    return v;
  }
};

// ---------------------------------------------------------------------------------------------
// Discrete adjoint of the step
// ---------------------------------------------------------------------------------------------

void Domain::replay_step(int n, const MhdSet& start, StepTape& tp, const std::vector<double>& times) {
  copy_set(start, L.view(), N, stream);
  hip_check(hipMemcpyAsync(step.p, history.p + n, sizeof(double), hipMemcpyDeviceToDevice, stream), "copy step");
  tstep = times[n];
  auto set_recording = [&](bool on) {
    for (auto& m : mods) {
      if (auto* tc = dynamic_cast<TcModule*>(m.get())) tc->recording = on;
      if (auto* vm = dynamic_cast<ViscModule*>(m.get())) vm->recording = on;
      if (auto* pv = dynamic_cast<PvModule*>(m.get())) pv->recording = on;
      if (auto* dc = dynamic_cast<DivCleanModule*>(m.get())) dc->recording = on;
    }
  };
  set_recording(true);
  rec = &tp;
  advance(n);
  rec = nullptr;
  set_recording(false);
}

std::vector<double> Domain::param_grad() {
  std::vector<double> out;
  for (auto& m : mods)
    if (auto* ph = dynamic_cast<ParamHeatModule*>(m.get())) {
      std::vector<double> g(ph->K);
      hip_check(hipMemcpy(g.data(), ph->g_theta.p, ph->K * sizeof(double), hipMemcpyDeviceToHost), "read dJ/dtheta");
      out.insert(out.end(), g.begin(), g.end());
    }
  return out;
}

void Domain::adjoint(int nsteps, int ckpt, const double* seed, double* grad, double* dt_out) {
  require(ready, SPRUCE_E_STATE, "setup must precede adjoint");
  // however the call ends (a failure inside a step included), later steps record nothing: the
  // tapes and records this call points the domain and its modules at are its own
  struct RecordingReset {
    Domain* d;
    ~RecordingReset() {
      d->rec = nullptr;
      for (auto& m : d->mods) {
        if (auto* tc = dynamic_cast<TcModule*>(m.get())) {
          tc->recording = false;
          tc->reuse = false;
          tc->slot = -1;
        }
        if (auto* vm = dynamic_cast<ViscModule*>(m.get())) vm->recording = false;
        if (auto* pv = dynamic_cast<PvModule*>(m.get())) pv->recording = false;
        if (auto* dc = dynamic_cast<DivCleanModule*>(m.get())) dc->recording = false;
      }
    }
  } recording_reset{this};
  const int integ = desc.time_integrator;
  require(integ == SPRUCE_EULER || integ == SPRUCE_RK2 || integ == SPRUCE_RK4, SPRUCE_E_UNSUPPORTED,
          "the adjoint covers euler, rk2 and rk4");
  const int* bc = desc.bc;
  for (int k = 0; k < 4; ++k)
    require(bc[k] == SPRUCE_BC_PERIODIC || bc[k] == SPRUCE_BC_OPEN || bc[k] == SPRUCE_BC_FIXED ||
                bc[k] == SPRUCE_BC_REFLECT || bc[k] == SPRUCE_BC_OPEN_UCNP, SPRUCE_E_UNSUPPORTED,
            "the adjoint covers periodic, open, fixed, reflect and open_ucnp boundaries");
  require(!((bc[0] == SPRUCE_BC_OPEN || bc[1] == SPRUCE_BC_OPEN) && bc[2] != SPRUCE_BC_PERIODIC) &&
              !((bc[2] == SPRUCE_BC_OPEN || bc[3] == SPRUCE_BC_OPEN) && bc[0] != SPRUCE_BC_PERIODIC),
          SPRUCE_E_UNSUPPORTED, "the adjoint of an open boundary needs the other axis periodic");
  // modules: conduction (no saturation) and radiative losses act in iterateModule; ambient and
  // parametrised heating and the source modules after the integrator
  std::vector<int> it_mods, post_mods;
  std::vector<ViscModule*> visc_mods;   // right-hand-side terms of every stage
  bool any_sat = false, any_pv = false, any_fheat = false, any_dc = false;
  rec_post.assign(mods.size(), false);
  for (size_t k = 0; k < mods.size(); ++k) {
    GpuModule* m = mods[k].get();
    // (an inactive module computes its update without applying it: the step does not depend on it)
    if (auto* tc = dynamic_cast<TcModule*>(m)) {
      any_sat = any_sat || (tc->p.saturation && !tc->inactive);
      it_mods.push_back((int)k);
    } else if (dynamic_cast<RlModule*>(m)) {
      it_mods.push_back((int)k);
    } else if (auto* pv = dynamic_cast<PvModule*>(m)) {
      it_mods.push_back((int)k);
      any_pv = true;
      (void)pv;
    } else if (dynamic_cast<DivCleanModule*>(m)) {
      it_mods.push_back((int)k);   // its iterate only fixes the subcycle count
      post_mods.push_back((int)k);
      rec_post[k] = true;
      any_dc = true;
    } else if (auto* fm = dynamic_cast<FieldHeatModule*>(m)) {
      it_mods.push_back((int)k);
      any_fheat = any_fheat || !fm->inactive;
    } else if (auto* vm = dynamic_cast<ViscModule*>(m)) {
      visc_mods.push_back(vm);                     // its right-hand-side terms
      if (vm->any_hyper) it_mods.push_back((int)k);   // its subcycled hyperviscous terms
    } else if (dynamic_cast<ParamHeatModule*>(m) || dynamic_cast<HeatModule*>(m) ||
               dynamic_cast<LocalHeatModule*>(m) || dynamic_cast<MassInjModule*>(m) ||
               dynamic_cast<MomInjModule*>(m) || dynamic_cast<HeatSinkModule*>(m) ||
               dynamic_cast<OutflowModule*>(m)) {
      post_mods.push_back((int)k);
      rec_post[k] = true;
    } else {
      throw Error(SPRUCE_E_UNSUPPORTED, "the adjoint does not cover module " + m->name);
    }
  }
  require(nsteps >= 1 && ckpt >= 1, SPRUCE_E_ARG, "nsteps and checkpoint_interval must be positive");
  require(seed && grad, SPRUCE_E_ARG, "null argument");
  const bool with_mods = !mods.empty();
  const bool rk4 = integ == SPRUCE_RK4;
  // modules and RK4 replay each step of a segment keeping its intermediate states; the plain
  // Euler/RK2 step keeps its stages for the whole segment instead
  const bool replay = with_mods || rk4;
  const int nseg = (nsteps + ckpt - 1) / ckpt;
  const int sl = std::min(ckpt, nsteps);
  const bool rk2 = integ == SPRUCE_RK2;
  const int nb = adj_blocks(mesh);
  const int nit = (int)it_mods.size(), npost = (int)post_mods.size();
  // One allocation: checkpoints (nseg sets), the recomputed segment (sl + 1 sets); without modules
  // its RK2 midpoints (sl sets) and stage derivatives (8 grids per step each), with modules the
  // snapshots of one replayed step; the three adjoint vectors, face coefficients, module scratch
  // (dJ/dT, the six frozen conduction fields, column sums) and the reduction partials.
  const size_t SET = 10 * N, EV = 8 * N, CO = (size_t)ADJ_COEFS * N;
  const size_t seg_store = replay ? (size_t)(nit + npost + 2 + (rk4 ? 3 : 0)) * SET + (rk4 ? 4 : 2) * EV
                                  : (rk2 ? (size_t)sl * SET : 0) + (size_t)sl * EV * (rk2 ? 2 : 1);
  const bool visc = !visc_mods.empty();
  const size_t modscratch = replay ? 10 * N + (size_t)ny + 2 * (size_t)nb + 2 + (any_sat ? 8 * N : 0) : 0;
  const size_t rk4scratch = rk4 ? 4 * EV + 2 * (size_t)nb : 0;
  bool any_hyper = false;
  for (ViscModule* vm : visc_mods) any_hyper = any_hyper || vm->any_hyper;
  const size_t viscscratch = visc ? (rk4 ? 0 : EV) + 3 * N + (size_t)adj_mod_blocks(mesh) + 1 + (any_hyper ? 5 * N : 0) : 0;
  // physical viscosity: the forward derivatives of an evaluation set (7), the scatter accumulators
  // (velocity derivatives 6, T^2.5, b_hat 3, velocities 3, T) and an RK2 subcycle's two adjoint sets (4 + 4)
  const size_t pvscratch = any_pv ? 29 * N : 0;
  // field heating: the adjoint of the step's start state its heating reads, and of that state's b_hat
  const size_t fhscratch = any_fheat ? 14 * N : 0;
  const size_t dcscratch = any_dc ? 4 * N : 0;   // divergence cleaning's reverse passes
  const size_t total = (size_t)nseg * SET + (size_t)(sl + 1) * SET + seg_store + 3 * EV + 2 * CO + 4 * (size_t)nb +
                       modscratch + rk4scratch + viscscratch + pvscratch + fhscratch + dcscratch;
  DBuf slab;
  slab.alloc(total);
  adj_bytes = (int64_t)(total * sizeof(double));
  double* cur = slab.p;
  auto take = [&](size_t n) { double* p = cur; cur += n; return p; };
  std::vector<MhdSet> cps(nseg), seg(sl + 1), mid;
  for (auto& c : cps) c = set_at(take(SET), N);
  for (auto& c : seg) c = set_at(take(SET), N);
  std::vector<double*> k1s, k2s;
  StepTape tape_one;   // the snapshots of one replayed step
  StepTape& tape = tape_one;
  if (replay) {
    tape.after_it.resize(nit);
    tape.after_post.resize(npost);
    for (auto& c : tape.after_it) c = set_at(take(SET), N);
    for (auto& c : tape.after_post) c = set_at(take(SET), N);
    tape.W = set_at(take(SET), N);
    tape.A3 = set_at(take(SET), N);
    tape.k1 = take(EV);
    tape.k2 = take(EV);
    if (rk4) {
      tape.W2 = set_at(take(SET), N);
      tape.W3 = set_at(take(SET), N);
      tape.Lm = set_at(take(SET), N);
      tape.k3 = take(EV);
      tape.kf = take(EV);
    }
  } else {
    mid.resize(rk2 ? sl : 0);
    for (auto& c : mid) c = set_at(take(SET), N);
    k1s.resize(sl);
    k2s.resize(rk2 ? sl : 0);
    for (auto& k : k1s) k = take(EV);
    for (auto& k : k2s) k = take(EV);
  }
  double* lam = take(EV);
  double* lamN = take(EV);
  double* lamW = take(EV);
  double* cx = take(CO);
  double* cy = take(CO);
  double* part = take(nb);
  double* part2 = take(nb);
  double* pmin = take(nb);
  int* pidx = reinterpret_cast<int*>(take(nb));
  double *lamT = nullptr, *lamfz = nullptr, *colsum = nullptr, *partm = nullptr, *gh = nullptr;
  double *partt = nullptr, *gtime = nullptr;   // dJ/d(simulation time) of the time-dependent sources
  double *lacc = nullptr, *lx = nullptr, *lamS = nullptr;   // conduction RK2/RK4 subcycle reverse
  if (replay) {
    lamT = take(N);
    lamfz = take(6 * N);
    lacc = take(N);
    lx = take(N);
    lamS = take(N);
    colsum = take(ny);
    partm = take(nb);
    gh = take(1);
    partt = take(nb);
    gtime = take(1);
  }
  // flux saturation: the forward coefficient and raw flux at the evaluated temperatures, and the
  // stage reverses' outputs for them
  TcSatRev satr;
  double *sat_coef = nullptr, *sat_rx = nullptr, *sat_ry = nullptr;
  if (replay && any_sat) {
    sat_coef = take(N);
    sat_rx = take(N);
    sat_ry = take(N);
    satr.coef = sat_coef; satr.rx = sat_rx; satr.ry = sat_ry;
    satr.lcoef = take(N); satr.wx = take(N); satr.wy = take(N); satr.lfx = take(N); satr.lfy = take(N);
  }
  // RK4: the adjoint of the stages' common base, the final combination's h*mu, one stage's
  // mult*h*mu, a second stage adjoint, and the partials of the two middle stages
  double *lamM = nullptr, *lkf = nullptr, *lka = nullptr, *lamX = nullptr, *part3 = nullptr, *part4 = nullptr;
  if (rk4) {
    lamM = take(EV);
    lkf = take(EV);
    lka = take(EV);
    lamX = take(EV);
    part3 = take(nb);
    part4 = take(nb);
  }
  // artificial viscosity: the stage derivatives' adjoint (the RK4 buffer otherwise), the Laplacian's
  // scatter, the dt grid's adjoint, the global minimum's block sums and its adjoint
  double *vlq = nullptr, *vlf = nullptr, *vldt = nullptr, *vpart = nullptr, *vg = nullptr;
  double* hv_init = nullptr;            // hyperviscous subcycles: the adjoint of the subcycle's start
  double* hv_d[4] = {nullptr, nullptr, nullptr, nullptr};   // and of its evaluations
  if (visc) {
    if (!rk4) lka = take(EV);
    vlq = take(N);
    vlf = take(N);
    launch_fill(vlf, 0.0, N, stream);
    vldt = take(N);
    vpart = take((size_t)adj_mod_blocks(mesh));
    vg = take(1);
    if (any_hyper) {
      hv_init = take(N);
      for (auto& q : hv_d) q = take(N);
    }
    launch_fill(vlq, 0.0, N, stream);
    launch_fill(vldt, 0.0, N, stream);
    launch_fill(vg, 0.0, 1, stream);
  }
  // one term's reverse at S for the seed sa*ka + kb of component evol
  auto visc_term_rev = [&](const ViscTermArgs& a, int evol, const double* ka, double sa, const double* kb,
                           double* lamS) {
    launch_adj_visc_term(a, evol, ka, sa, kb, vlq, vlf, lamS, vldt, a.dtgrid ? nullptr : vpart, stream);
    if (!a.dtgrid) launch_adj_accum(vpart, adj_mod_blocks(mesh), 1.0, vg, stream);
    if (a.grad) {
      launch_adj_visc_f(a, vlf, lamS, vldt, a.dtgrid ? nullptr : vpart, stream);
      if (!a.dtgrid) launch_adj_accum(vpart, adj_mod_blocks(mesh), 1.0, vg, stream);
    }
    launch_adj_visc_vars(a, vlq, lamS, stream);
  };
  // the timescales' adjoint (dt grid of S, per cell, and its minimum) taken to S
  auto visc_dt_rev = [&](const MhdSet& S, double* lamS) {
    launch_adj_dt_grid(S, stat(), kc, mesh, vldt, lamS, stream);
    launch_adj_dt(S, stat(), kc, mesh, 1.0, nullptr, nullptr, vg, pmin, pidx, lamS, stream);
    launch_fill(vg, 0.0, 1, stream);
  };
  double* dcs = nullptr;
  if (any_dc) {
    dcs = take(4 * N);
    launch_fill(dcs, 0.0, 4 * N, stream);
  }
  double *fh_lam = nullptr, *fh_fz = nullptr;
  if (any_fheat) {
    fh_lam = take(8 * N);
    fh_fz = take(6 * N);
    launch_fill(fh_lam, 0.0, 8 * N, stream);
  }
  double* pvs = nullptr;
  if (any_pv) {
    pvs = take(29 * N);
    launch_fill(pvs, 0.0, 29 * N, stream);
  }
  auto visc_rev = [&](const MhdSet& S, const double* ka, double sa, const double* kb, double* lamS) {
    for (ViscModule* vm : visc_mods)
      for (auto& t : vm->terms) {
        if (t.hyper()) continue;
        const ViscTermArgs a = vm->term_args(*this, t, S);
        timed("adj_visc", [&] { visc_term_rev(a, t.evol_idx, ka, sa, kb, lamS); });
      }
  };
  if ((int)history.n < nsteps + 1) history.alloc(nsteps + 1);
  hip_check(hipMemcpyAsync(history.p, step.p, sizeof(double), hipMemcpyDeviceToDevice, stream), "copy step");
  sub_hist.assign((size_t)nsteps * std::max<size_t>(1, mods.size()), -1);
  const double time0 = time;
  tstep = time;
  // forward sweep, exactly the steps of run()
  for (int n = 0; n < nsteps; ++n) {
    if (n % ckpt == 0) copy_set(L.view(), cps[n / ckpt], N, stream);
    advance(n);
  }
  sync_and_check();
  std::vector<double> h(nsteps), times(nsteps + 1);
  hip_check(hipMemcpy(h.data(), history.p, nsteps * sizeof(double), hipMemcpyDeviceToHost), "read dt history");
  times[0] = time0;
  for (int n = 0; n < nsteps; ++n) {
    times[n + 1] = times[n] + h[n];
    time += h[n];
    iter++;
    if (dt_out) dt_out[n] = h[n];
  }
  // the handle keeps the state after the steps: the replays of the backward sweep overwrite the
  // live grids, so a copy is restored at the end
  DBuf keep;
  MhdSet kept{};
  std::vector<std::pair<OutflowModule*, DBuf>> kept_sc;   // the outflow scalars the last step left
  if (replay) {
    keep.alloc(SET);
    kept = set_at(keep.p, N);
    copy_set(L.view(), kept, N, stream);
    adj_bytes += (int64_t)(SET * sizeof(double));
    for (auto& m : mods)
      if (auto* om = dynamic_cast<OutflowModule*>(m.get())) {
        kept_sc.emplace_back(om, DBuf());
        kept_sc.back().second.alloc(2);
        hip_check(hipMemcpyAsync(kept_sc.back().second.p, om->sc.p, 2 * sizeof(double), hipMemcpyDeviceToDevice,
                                 stream), "keep outflow scalars");
      }
  }
  // Conduction records of whole segments: the segment recompute keeps every step's subcycle
  // results, so the replay of a step skips its subcycles.  Slot i holds step n0 + i of the segment
  // being reversed; it is used when the records of the largest counts the forward took fit in 80%
  // of the free device memory.
  std::vector<TcModule*> slot_tcs;
  if (replay) {
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    size_t need = 0;
    std::vector<std::pair<TcModule*, size_t>> cand;
    for (size_t k = 0; k < mods.size(); ++k)
      if (auto* tc = dynamic_cast<TcModule*>(mods[k].get())) {
        tc->slot = -1;                                  // (an earlier call that failed may have left one)
        tc->reuse = false;
        tc->slot_pool = DBuf();
        tc->slot_off.assign(sl, 0);
        tc->slot_cap.assign(sl, 0);
        size_t recs = 0;
        for (int i = 0; i < sl; ++i) {
          int mx = 0;
          for (int sg = 0; sg < nseg; ++sg) {
            const int n = sg * ckpt + i;
            if (n < nsteps) mx = std::max(mx, (int)sub_hist[(size_t)n * mods.size() + k]);
          }
          tc->slot_off[i] = recs;
          tc->slot_cap[i] = (size_t)mx * tc->rec_per_sub();
          recs += tc->slot_cap[i];
        }
        cand.emplace_back(tc, recs);
        need += recs * N * sizeof(double);
      }
    if (adj_tc_slots && !cand.empty() && (double)need <= 0.8 * (double)free_b)
      for (auto& c : cand) {
        if (c.second) c.first->slot_pool.alloc(c.second * N);
        c.first->slot_n.assign(sl, -1);
        slot_tcs.push_back(c.first);
        adj_bytes += (int64_t)(c.second * N * sizeof(double));
      }
  }
  // Whole-segment tapes: with conduction records kept per segment and only modules whose reverse
  // needs nothing of a step beyond its tape, its step size and subcycle counts (conduction, radiative
  // losses, the heating and source modules after the integrator), the segment recompute also keeps
  // every step's StepTape and no step is replayed; b_hat (and its derivatives) are formed again from
  // the step's start state.  Used where the tapes fit in 80% of the free device memory.
  DBuf tape_pool;
  std::vector<StepTape> tapes;
  bool seg_tapes = false;
  if (replay && adj_tc_slots) {
    bool ok = true, any_tc = false;
    for (auto& m : mods) {
      GpuModule* g = m.get();
      if (dynamic_cast<TcModule*>(g)) any_tc = true;
      else if (auto* vm = dynamic_cast<ViscModule*>(g)) ok = ok && !vm->any_hyper;   // its dt grid is formed again
      else if (!(dynamic_cast<RlModule*>(g) || dynamic_cast<FieldHeatModule*>(g) || dynamic_cast<OutflowModule*>(g) ||
                 dynamic_cast<ParamHeatModule*>(g) || dynamic_cast<HeatModule*>(g) ||
                 dynamic_cast<LocalHeatModule*>(g) || dynamic_cast<MassInjModule*>(g) ||
                 dynamic_cast<MomInjModule*>(g) || dynamic_cast<HeatSinkModule*>(g)))
        ok = false;
    }
    if (ok && (!any_tc || !slot_tcs.empty())) {
      const size_t per = (size_t)(nit + npost + 2 + (rk4 ? 3 : 0)) * SET + (rk4 ? 4 : 2) * EV;
      size_t free_b = 0, total_b = 0;
      hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
      if ((double)(per * sl * sizeof(double)) <= 0.8 * (double)free_b) {
        tape_pool.alloc(per * sl);
        adj_bytes += (int64_t)(per * sl * sizeof(double));
        tapes.resize(sl);
        double* q = tape_pool.p;
        auto tk = [&](size_t n) { double* r = q; q += n; return r; };
        for (auto& t : tapes) {
          t.after_it.resize(nit);
          t.after_post.resize(npost);
          for (auto& c : t.after_it) c = set_at(tk(SET), N);
          for (auto& c : t.after_post) c = set_at(tk(SET), N);
          t.W = set_at(tk(SET), N);
          t.A3 = set_at(tk(SET), N);
          t.k1 = tk(EV);
          t.k2 = tk(EV);
          if (rk4) {
            t.W2 = set_at(tk(SET), N);
            t.W3 = set_at(tk(SET), N);
            t.Lm = set_at(tk(SET), N);
            t.k3 = tk(EV);
            t.kf = tk(EV);
          }
        }
        seg_tapes = true;
      }
    }
  }
  for (auto& m : mods)
    if (auto* ph = dynamic_cast<ParamHeatModule*>(m.get())) launch_fill(ph->g_theta.p, 0.0, ph->K, stream);
  if (lamT) launch_fill(lamT, 0.0, N, stream);
  if (gtime) launch_fill(gtime, 0.0, 1, stream);
  hip_check(hipMemcpyAsync(lam, seed, 8 * N * 8, hipMemcpyHostToDevice, stream), "upload seed");
  const MhdStatic st = stat();
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int n0 = sg * ckpt, n1 = std::min(nsteps, n0 + ckpt);
    copy_set(cps[sg], seg[0], N, stream);
    if (!replay) {
      // recompute the segment: states, k1, W, k2 (and the final ghost fill of each step)
      for (int n = n0; n < n1; ++n) {
        const int i = n - n0;
        if (!rk2) {
          stage_view(seg[i], seg[i], seg[i + 1], 1.0, STAGE_STORE_K, history.p + n, k1s[i]);
        } else {
          stage_view(seg[i], seg[i], mid[i], 0.5, STAGE_STORE_K, history.p + n, k1s[i]);
          ghost_set(mid[i], kinds_ucnp, GHOST_NORMAL);
          stage_view(mid[i], seg[i], seg[i + 1], 1.0, STAGE_STORE_K, history.p + n, k2s[i]);
        }
        launch_mhd_ghost(seg[i + 1], kc, mesh, kinds_all, s1, s2, GHOST_NORMAL, stream);
      }
    } else {
      copy_set(seg[0], L.view(), N, stream);
      hip_check(hipMemcpyAsync(step.p, history.p + n0, sizeof(double), hipMemcpyDeviceToDevice, stream), "copy step");
      tstep = times[n0];
      for (int n = n0; n < n1; ++n) {
        for (TcModule* t : slot_tcs) { t->slot = n - n0; t->recording = true; }
        if (seg_tapes) rec = &tapes[n - n0];
        advance(n);
        rec = nullptr;
        for (TcModule* t : slot_tcs) { t->recording = false; t->slot = -1; }
        copy_set(L.view(), seg[n - n0 + 1], N, stream);
      }
    }
    for (int n = n1 - 1; n >= n0; --n) {
      const int i = n - n0;
      const MhdSet& Ln = seg[i];
      const double* hp = history.p + n;
      if (!replay) {
        adj_ghost(seg[i + 1], lam);
        // lamN and lamW are written first by the stage-cell kernels (assign), so they need no clear
        if (rk2) {
          timed("adj_stage", [&] {
            launch_adj_stage_cells(Ln, k2s[i], hp, 1.0, mid[i], st, kc, mesh, lam, lamN, 1, lamW, 1, cx, cy, part,
                                   stream);
          });
          timed("adj_faces", [&] { launch_adj_faces(mid[i], st, kc, mesh, cx, cy, lamW, stream); });
          adj_ghost(mid[i], lamW, kinds_ucnp);             // the midpoint's open_ucnp ghost fill
          timed("adj_stage", [&] {
            launch_adj_stage_cells(Ln, k1s[i], hp, 0.5, Ln, st, kc, mesh, lamW, lamN, 0, lamN, 0, cx, cy, part2,
                                   stream);
          });
        } else {
          timed("adj_stage", [&] {
            launch_adj_stage_cells(Ln, k1s[i], hp, 1.0, Ln, st, kc, mesh, lam, lamN, 1, lamN, 0, cx, cy, part, stream);
          });
        }
        timed("adj_faces", [&] { launch_adj_faces(Ln, st, kc, mesh, cx, cy, lamN, stream); });
        timed("adj_dt", [&] {
          launch_adj_dt(Ln, st, kc, mesh, kc.eps, part, rk2 ? part2 : nullptr, nullptr, pmin, pidx, lamN, stream);
        });
        std::swap(lam, lamN);
        continue;
      }
      // with modules: replay the step keeping its intermediate states, then reverse it
      StepTape& tape = seg_tapes ? tapes[i] : tape_one;
      if (seg_tapes) {
        // the module state the reverse reads, as the step set it: subcycle counts, their step
        // sizes (iterate's dt / nsub), b_hat of the step's start state and its derivatives
        bool bh = false;
        for (size_t k = 0; k < mods.size(); ++k) {
          GpuModule* m = mods[k].get();
          bh = bh || m->needs_bhat();
          const int ns = sub_hist[(size_t)n * mods.size() + k];
          if (auto* tc = dynamic_cast<TcModule*>(m)) {
            tc->nsub = ns;
            tc->h_last = h[n] / (double)ns;
            tc->slot = i;
          } else if (auto* rl = dynamic_cast<RlModule*>(m)) {
            rl->nsub = ns;
            rl->h_last = h[n] / (double)ns;
          }
        }
        if (bh) launch_bhat(Ln, st, mesh, bhx.p, bhy.p, stream);
        for (TcModule* t : slot_tcs)
          launch_bhat_derivs(bhx.p, bhy.p, mesh, t->dbxx.p, t->dbxy.p, t->dbyx.p, t->dbyy.p, stream);
        // artificial viscosity's right-hand-side terms read the dt grid of the state the integrator
        // starts from and its minimum (ViscModule::refresh before the integrator)
        const MhdSet& Av = nit ? tape.after_it[nit - 1] : Ln;
        for (ViscModule* vm : visc_mods) {
          launch_mhd_dt(Av, st, kc, mesh, vm->dtg.p, vm->part.p, stream);
          launch_reduce_step(vm->part.p, dt_blocks, 1.0, vm->dtmin.p, nullptr, 0, err_flag, stream);
        }
      } else {
        for (TcModule* t : slot_tcs) { t->slot = i; t->reuse = true; }   // the reverse reads slot i too
        replay_step(n, Ln, tape, times);
        for (TcModule* t : slot_tcs) t->reuse = false;
      }
      // dJ/dh of this step starts with dJ/d(time) of every later step: the simulation time a
      // time-dependent source reads at step m is the initial time plus the step sizes before m
      hip_check(hipMemcpyAsync(gh, gtime, sizeof(double), hipMemcpyDeviceToDevice, stream), "copy time adjoint");
      for (int p = npost - 1; p >= 0; --p) {              // post-iterate modules (heating, sources)
        const MhdSet& post = tape.after_post[p];
        const MhdSet& pre = p ? tape.after_post[p - 1] : tape.A3;
        GpuModule* m = mods[post_mods[p]].get();
        auto* ph = dynamic_cast<ParamHeatModule*>(m);
        if (auto* om = dynamic_cast<OutflowModule*>(m)) {     // boundaryoutflow.cpp:39-67
          adj_ghost(post, lam);
          // without a replay, the window's largest outflow speed, its cell and the acceleration are
          // formed again from the state the module acted on, as the step formed them
          if (seg_tapes) launch_outflow_max(pre, st, om->p, mesh, om->part.p, om->sc.p, 1, stream);
          timed("adj_outflow", [&] {
            launch_adj_minmask(pre.rho, pre.te, nullptr, nullptr, kc, mesh, lam, nullptr, stream);
            const int onb = outflow_blocks(om->p);
            launch_adj_outflow(pre, st, om->tmpl.p, om->sc.p, hp, om->p, mesh, om->part.p, om->part.p + onb, onb, lam,
                               partt, partm, stream);
            launch_adj_accum(partm, nb, 1.0, gh, stream);
          });
          continue;
        }
        if (auto* dc = dynamic_cast<DivCleanModule*>(m)) {   // divcleaning.cpp:28-46, subcycles last first
          adj_ghost(post, lam);
          timed("adj_divclean", [&] {
            launch_adj_minmask(pre.rho, pre.te, nullptr, nullptr, kc, mesh, lam, nullptr, stream);
            for (int sc = dc->nsub - 1; sc >= 0; --sc) {
              const double* bxs = dc->tape[(size_t)sc].p;
              const double* bys = bxs + N;
              launch_dc_inner(bxs, bys, mesh, dc->dxb.p, dc->dyb.p, stream);
              launch_adj_divclean(bxs, bys, dc->dxb.p, dc->dyb.p, dc->coeff.p, dc->h, mesh, lam + 5 * N, lam + 6 * N,
                                  dcs, partm, stream);
              launch_adj_accum(partm, nb, 1.0 / (double)dc->nsub, gh, stream);
            }
          });
          continue;
        }
        if (!ph && !dynamic_cast<HeatModule*>(m)) {
          SourceRev r;
          if (!static_cast<PostSourceModule*>(m)->source_rev(times[n], r)) continue;   // the module did nothing
          adj_ghost(post, lam);
          timed("adj_source", [&] {
            launch_adj_source(pre, r.mode, r.t, r.t2, hp, r.g, r.gp, kc, mesh, lam, partm, partt, stream);
            launch_adj_accum(partm, nb, 1.0, gh, stream);
            if (r.gp != 0.0) launch_adj_accum(partt, nb, 1.0, gtime, stream);
          });
          continue;
        }
        adj_ghost(post, lam);
        const double* heat = ph ? ph->heating.p : static_cast<HeatModule*>(m)->heating.p;
        timed("adj_heating", [&] {
          launch_adj_minmask(pre.rho, pre.te, heat, hp, kc, mesh, lam, partm, stream);
          launch_adj_accum(partm, nb, 1.0, gh, stream);
          if (ph) launch_adj_heat_params(lam + 4 * N, ph->basis.p, ph->K, hp, mesh, colsum, ph->g_theta.p, stream);
        });
      }
      adj_ghost(tape.A3, lam);                             // the integrator's final ghost fill
      const MhdSet& A2 = nit ? tape.after_it[nit - 1] : Ln;
      if (rk4) {
        // final: L = F(Lm + h kf), kf = (k1 + k4)/6 + (k2 + k3)/3 with k4 = k(W3); lkf = h mu
        timed("adj_stage", [&] {
          launch_adj_update(tape.Lm, tape.kf, hp, 1.0, kc, mesh, lam, lamM, 1, lkf, part, stream);
          launch_adj_rhs_cells(tape.W3, st, kc, mesh, lkf, 1.0 / 6.0, nullptr, lamW, 1, cx, cy, stream);
        });
        visc_rev(tape.W3, lkf, 1.0 / 6.0, nullptr, lamW);
        timed("adj_faces", [&] { launch_adj_faces(tape.W3, st, kc, mesh, cx, cy, lamW, stream); });
        adj_ghost(tape.W3, lamW, kinds_ucnp);              // each stage state's open_ucnp ghost fill
        // W3 = F(Lm + h k3), k3 = k(W2): lambda_k3 = h mu3 + lkf/3
        timed("adj_stage", [&] {
          launch_adj_update(tape.Lm, tape.k3, hp, 1.0, kc, mesh, lamW, lamM, 0, lka, part2, stream);
          launch_adj_rhs_cells(tape.W2, st, kc, mesh, lkf, 1.0 / 3.0, lka, lamX, 1, cx, cy, stream);
        });
        visc_rev(tape.W2, lkf, 1.0 / 3.0, lka, lamX);
        timed("adj_faces", [&] { launch_adj_faces(tape.W2, st, kc, mesh, cx, cy, lamX, stream); });
        adj_ghost(tape.W2, lamX, kinds_ucnp);
        // W2 = F(Lm + h/2 k2), k2 = k(W): lambda_k2 = h/2 mu2 + lkf/3
        timed("adj_stage", [&] {
          launch_adj_update(tape.Lm, tape.k2, hp, 0.5, kc, mesh, lamX, lamM, 0, lka, part3, stream);
          launch_adj_rhs_cells(tape.W, st, kc, mesh, lkf, 1.0 / 3.0, lka, lamW, 1, cx, cy, stream);
        });
        visc_rev(tape.W, lkf, 1.0 / 3.0, lka, lamW);
        timed("adj_faces", [&] { launch_adj_faces(tape.W, st, kc, mesh, cx, cy, lamW, stream); });
        adj_ghost(tape.W, lamW, kinds_ucnp);
        // W = F(A2 + h/2 k1), k1 = k(A2): lambda_k1 = h/2 mu1 + lkf/6
        timed("adj_stage", [&] {
          launch_adj_update(A2, tape.k1, hp, 0.5, kc, mesh, lamW, lamN, 1, lka, part4, stream);
          launch_adj_rhs_cells(A2, st, kc, mesh, lkf, 1.0 / 6.0, lka, lamN, 0, cx, cy, stream);
        });
        visc_rev(A2, lkf, 1.0 / 6.0, lka, lamN);
        // (its faces follow below, as for the other integrators); Lm is A2 with its ghost cells refilled from A2's interior without floors
        // (GHOST_LIVE_RAW, evolution.cpp:126-152 writing the live grids): its ghost adjoint goes to
        // that interior, the rest is A2's own
        launch_adj_ghost_axis(A2, kc, mesh, kinds_live, s1, s2, 1, lamM, stream, 1);
        launch_adj_ghost_axis(A2, kc, mesh, kinds_live, s1, s2, 0, lamM, stream, 1);
        launch_add(lamN, lamM, 8 * N, stream);
        launch_adj_accum(part3, nb, 1.0, gh, stream);
        launch_adj_accum(part4, nb, 1.0, gh, stream);
      } else {
        // one stage reverse: T = F(B + mult*h*k(S)) with lam_T in, lam_B and lam_S out; with viscosity
        // terms the stage derivative's adjoint is kept for them
        auto stage_rev = [&](const double* kv, double mult, const MhdSet& S, const double* lam_T, double* lam_B,
                             int assign_B, double* lam_S, int assign_S, double* prt) {
          timed("adj_stage", [&] {
            if (!visc) {
              launch_adj_stage_cells(A2, kv, hp, mult, S, st, kc, mesh, lam_T, lam_B, assign_B, lam_S, assign_S, cx,
                                     cy, prt, stream);
            } else {
              launch_adj_update(A2, kv, hp, mult, kc, mesh, lam_T, lam_B, assign_B, lka, prt, stream);
              launch_adj_rhs_cells(S, st, kc, mesh, lka, 1.0, nullptr, lam_S, assign_S, cx, cy, stream);
            }
          });
          if (visc) visc_rev(S, lka, 1.0, nullptr, lam_S);
        };
        if (rk2) {
          stage_rev(tape.k2, 1.0, tape.W, lam, lamN, 1, lamW, 1, part);
          timed("adj_faces", [&] { launch_adj_faces(tape.W, st, kc, mesh, cx, cy, lamW, stream); });
          adj_ghost(tape.W, lamW, kinds_ucnp);
          stage_rev(tape.k1, 0.5, A2, lamW, lamN, 0, lamN, 0, part2);
        } else {
          stage_rev(tape.k1, 1.0, A2, lam, lamN, 1, lamN, 0, part);
        }
      }
      timed("adj_faces", [&] { launch_adj_faces(A2, st, kc, mesh, cx, cy, lamN, stream); });
      if (visc) {   // the viscosity coefficient's timescale: the dt grid of A2 (local) or its minimum
        timed("adj_visc", [&] { visc_dt_rev(A2, lamN); });
      }
      std::swap(lam, lamN);
      for (int k = nit - 1; k >= 0; --k) {                 // iterating modules, last first
        const MhdSet& post = tape.after_it[k];
        const MhdSet& pre = k ? tape.after_it[k - 1] : Ln;
        GpuModule* m = mods[it_mods[k]].get();
        if (dynamic_cast<DivCleanModule*>(m)) continue;   // its iterate leaves the state alone
        {                                                 // inactive modules leave it alone too
          auto* tci = dynamic_cast<TcModule*>(m);
          auto* rli = dynamic_cast<RlModule*>(m);
          auto* fhi = dynamic_cast<FieldHeatModule*>(m);
          if ((tci && tci->inactive) || (rli && rli->inactive) || (fhi && fhi->inactive)) continue;
        }
        adj_ghost(post, lam);
        if (auto* tc = dynamic_cast<TcModule*>(m)) {
          const int ns = tc->nsub;
          const double* te_last = ns ? tc->rec_out(ns - 1) : pre.te;
          const double hs = tc->h_last, k0 = tc->p.k0;
          double* lte = lam + 4 * N;
          // with flux saturation every evaluation at temperatures `ev` first recomputes the forward
          // coefficient and raw flux there, and its stage reverse is followed by their reverse
          const bool sat = tc->p.saturation != 0;
          const TcSatRev sr = sat ? satr : TcSatRev();
          auto sat_pre = [&](const double* ev) {
            if (sat) launch_tc_sat(ev, pre.n, pre.rho, bhx.p, bhy.p, tc->p, kc, mesh, sat_rx, sat_ry, sat_coef, stream);
          };
          auto sat_post = [&](const double* ev) {
            if (sat) launch_adj_tc_satrev(ev, pre.n, pre.rho, bhx.p, bhy.p, tc->p, kc, mesh, sr, lamT, lamfz, lam, stream);
          };
          // one stage reverse: dE at `ev`, clamp base te_s (or none), seed extra
          // the step-size partials of every subcycle kernel are summed per block, reduced once
          int acc = 0;
          auto stage_rev = [&](const double* ev, double mult, const double* cb, const double* lout, int assign,
                               const double* sadd, double coef) {
            sat_pre(ev);
            launch_adj_tc_stage(ev, pre.n, bhx.p, bhy.p, tc->dbxx.p, tc->dbxy.p, tc->dbyx.p, tc->dbyy.p, hs, mult, k0,
                                kc, mesh, cb, lout, lacc, assign, sadd, coef, lamT, lamfz, partm, acc, stream, sr);
            sat_post(ev);
            acc = 1;
          };
          // dJ/dT at the temperatures of `ev` onto dJ/dte (into `into`, cleared first unless it
          // accumulates) and dJ/drho
          auto temp_rev = [&](const double* ev, double* into, bool clear) {
            if (clear) launch_fill(into, 0.0, N, stream);
            launch_adj_tc_temp(ev, pre.n, kc, N, lamT, into, lam, stream);
          };
          // Euler subcycles without saturation: the strip form, alternating the adjoint of te between
          // lte and lamT (zero outside this reverse)
          const int srows = adj_tc_rows == 0 ? 8 : adj_tc_rows;
          const int sblocks = adj_tc_rows >= 0 ? adj_tc_strip_blocks(mesh, srows) : 0;
          const bool strips = tc->integ == "euler" && !sat && sblocks > 0 && sblocks <= nb;
          timed("adj_tc", [&] {
            launch_adj_minmask(pre.rho, te_last, nullptr, nullptr, kc, mesh, lam, nullptr, stream);
            launch_fill(lamfz, 0.0, 6 * N, stream);
            if (strips) {
              double *lin = lte, *lout = lamT;
              for (int q = ns - 1; q >= 0; --q) {
                const double* te_s = q ? tc->rec_out(q - 1) : pre.te;
                launch_adj_tc_strip(te_s, pre.n, bhx.p, bhy.p, hs, k0, kc, mesh, lin, lout, lam, lamfz, lamfz + N,
                                    partm, acc, srows, stream);
                acc = 1;
                std::swap(lin, lout);
              }
              if (lin != lte)
                hip_check(hipMemcpyAsync(lte, lin, N * 8, hipMemcpyDeviceToDevice, stream), "copy conduction adjoint");
              if (ns) launch_fill(lamT, 0.0, N, stream);
              if (acc) launch_adj_accum(partm, sblocks, 1.0 / (double)ns, gh, stream);
              launch_adj_bhat(Ln, st, mesh, lamfz, lam, stream);
              return;
            }
            for (int q = ns - 1; q >= 0; --q) {
              const double* te_s = q ? tc->rec_out(q - 1) : pre.te;
              if (tc->integ == "euler") {
                sat_pre(te_s);
                launch_adj_tc_sub(te_s, pre.n, bhx.p, bhy.p, tc->dbxx.p, tc->dbxy.p, tc->dbyx.p, tc->dbyy.p, hs, k0,
                                  kc, mesh, lte, lamT, lamfz, partm, acc, stream, sr);
                sat_post(te_s);
                acc = 1;
                launch_adj_tc_temp(te_s, pre.n, kc, N, lamT, lte, lam, stream);
                continue;
              }
              if (tc->integ == "rk2") {                 // thermalconduction.cpp:67-78
                const double* half = tc->rec_at(q, 0);
                stage_rev(half, 1.0, te_s, lte, 1, nullptr, 0.0);
                temp_rev(half, lx, true);
                stage_rev(te_s, 0.5, te_s, lx, 0, nullptr, 0.0);
              } else {                                  // thermalconduction.cpp:79-98
                const double *i1 = tc->rec_at(q, 0), *i2 = tc->rec_at(q, 1), *i3 = tc->rec_at(q, 2);
                sat_pre(i3);
                launch_adj_tc_final4(te_s, tc->rec_at(q, 3), tc->rec_at(q, 4), tc->rec_at(q, 5), i3, pre.n, bhx.p,
                                     bhy.p, tc->dbxx.p, tc->dbxy.p, tc->dbyx.p, tc->dbyy.p, hs, k0, kc, mesh, lte, lacc,
                                     lamS, lamT, lamfz, partm, acc, stream, sr);
                sat_post(i3);
                acc = 1;
                temp_rev(i3, lx, true);
                stage_rev(i2, 1.0, te_s, lx, 0, lamS, 2.0);
                temp_rev(i2, lx, true);
                stage_rev(i1, 0.5, te_s, lx, 0, lamS, 2.0);
                temp_rev(i1, lx, true);
                stage_rev(te_s, 0.5, te_s, lx, 0, lamS, 1.0);
              }
              temp_rev(te_s, lacc, false);
              hip_check(hipMemcpyAsync(lte, lacc, N * 8, hipMemcpyDeviceToDevice, stream), "copy conduction adjoint");
            }
            if (acc) launch_adj_accum(partm, adj_tc_blocks(mesh), 1.0 / (double)ns, gh, stream);
            launch_adj_bhat(Ln, st, mesh, lamfz, lam, stream);
          });
        } else if (auto* fm = dynamic_cast<FieldHeatModule*>(m)) {
          // field heating (fieldheating.cpp:34-61): its heating reads the step's start state Ln, whose
          // adjoint collects in fh_lam until the modules are reversed
          timed("adj_fheat", [&] {
            launch_fill(fh_fz, 0.0, 6 * N, stream);
            launch_adj_fheat(Ln, pre, st, bhx.p, bhy.p, fm->p, hp, kc, mesh, lam, fh_lam, fh_fz, partm, stream);
            launch_adj_accum(partm, nb, 1.0, gh, stream);
            launch_adj_bhat(Ln, st, mesh, fh_fz, fh_lam, stream);
          });
        } else if (auto* pv = dynamic_cast<PvModule*>(m)) {
          // physical viscosity (physicalviscosity.cpp:148-250): its subcycles reversed, last first;
          // the thermal energy and momenta of lam carry the subcycle states' adjoint
          if (!pv->inactive && pv->nsub > 0) timed("adj_pvisc", [&] {
            const bool e = pv->integ == "euler";
            const PvFrozen fz{pv->bhx.p, pv->bhy.p, pv->bhz.p, pv->coef.p, pre.n, pre.rho};
            double* q = pvs;
            PvDerivs pd{q, q + N, q + 2 * N, q + 3 * N, q + 4 * N, q + 5 * N, q + 6 * N};
            q += 7 * N;
            PvLam Lm{};
            for (int k = 0; k < 6; ++k) Lm.l_d[k] = q + (size_t)k * N;
            Lm.l_t25 = q + 6 * N;
            for (int k = 0; k < 3; ++k) Lm.l_bh[k] = q + (size_t)(7 + k) * N;
            for (int k = 0; k < 3; ++k) Lm.l_v[k] = q + (size_t)(10 + k) * N;
            Lm.l_T = q + 13 * N;
            double* lH = q + 14 * N;     // te, mx, my, mz of the half step
            double* lB = q + 18 * N;     // te, mx, my, mz of the subcycle start (RK2)
            Lm.partial = partm;
            double* lte = lam + 4 * N;
            double* lm3[3] = {lam + N, lam + 2 * N, lam + 3 * N};
            PvArgs a{};
            a.d = pd; a.f = fz; a.k = kc; a.mesh = mesh; a.h = pv->h_last;
            a.heating_on = pv->heating_on ? 1 : 0; a.force_on = pv->force_on ? 1 : 0; a.apply = 1;
            a.coeff_zero = pv->coeff == 0.0 ? 1 : 0; a.grad_corr = pv->grad_corr ? 1 : 0;
            int acc = 0;
            // one update reversed: output adjoint (lo_te, lo_m) into the base's (mode 0/1/2), then the
            // evaluation's derivatives and derived variables onto (out_te, out_m) (assign or add)
            auto update_rev = [&](const PvSet& E, const PvSet& B, int mode, const double* lo_te, double* const lo_m[3],
                                  double* lb_te, double* const lb_m[3], int base_mode, double* out_te,
                                  double* const out_m[3], int assign) {
              launch_pv_derivs(E, mesh, pd, stream);
              a.eval = E; a.base = B; a.mode = mode;
              Lm.lo_te = lo_te; Lm.lb_te = lb_te; Lm.base_mode = base_mode; Lm.accumulate = acc;
              for (int k = 0; k < 3; ++k) { Lm.lo_m[k] = lo_m[k]; Lm.lb_m[k] = lb_m[k]; }
              launch_adj_pv_apply(a, Lm, stream);
              acc = 1;
              launch_adj_pv_derivs(E, mesh, Lm, stream);
              launch_adj_pv_derived(E, fz, kc, N, Lm, out_te, out_m[0], out_m[1], out_m[2], assign, lam, stream);
            };
            double* lHm[3] = {lH + N, lH + 2 * N, lH + 3 * N};
            double* lBm[3] = {lB + N, lB + 2 * N, lB + 3 * N};
            for (int sc = pv->nsub - 1; sc >= 0; --sc) {
              if (e) {
                const PvSet E = PvModule::set_of(pv->tape[(size_t)sc].p, N);
                update_rev(E, E, PV_FULL, lte, lm3, lte, lm3, 0, lte, lm3, 0);
              } else {
                const PvSet E = PvModule::set_of(pv->tape[2 * (size_t)sc].p, N);
                const PvSet H = PvModule::set_of(pv->tape[2 * (size_t)sc + 1].p, N);
                update_rev(H, E, PV_FULL, lte, lm3, lB, lBm, 1, lH, lHm, 1);
                update_rev(E, E, PV_HALF, lH, lHm, lB, lBm, 2, lB, lBm, 0);
                hip_check(hipMemcpyAsync(lte, lB, N * 8, hipMemcpyDeviceToDevice, stream), "copy adjoint");
                for (int k = 0; k < 3; ++k)
                  hip_check(hipMemcpyAsync(lm3[k], lBm[k], N * 8, hipMemcpyDeviceToDevice, stream), "copy adjoint");
              }
            }
            launch_adj_pv_bhat(pre, st, N, Lm, lam, stream);
            launch_adj_accum(partm, nb, 1.0 / (double)pv->nsub, gh, stream);
          });
        } else if (auto* vm = dynamic_cast<ViscModule*>(m)) {
          // hyperviscous subcycles (viscosity.cpp:125-180), last term and last subcycle first: each
          // update g = init + mask*(c*h)*D and its propagate reversed (launch_adj_hv_update), then
          // the evaluation that produced D's newest derivative at the state it read, with that
          // state's timescales recomputed
          const int K = vm->hv_integ == "euler" ? 1 : (vm->hv_integ == "rk2" ? 2 : 4);
          const size_t per = vm->hv_integ == "rk4" ? 9 : (size_t)(2 * K);
          const double cs[4] = {K == 1 ? 1.0 : 0.5, K == 2 ? 1.0 : 0.5, 1.0, 1.0};
          timed("adj_visc", [&] {
            int slot = (int)vm->hv_rec.size();
            for (int ti = (int)vm->terms.size() - 1; ti >= 0; --ti) {
              auto& t = vm->terms[ti];
              if (!t.hyper()) continue;
              const auto& hr = vm->hv_rec[--slot];
              const int g = t.evol_idx;
              // the state after this term's subcycles: the next term's first recorded state, or the module's output
              const MhdSet out_state =
                  slot + 1 < (int)vm->hv_rec.size() ? set_at(vm->tape[vm->hv_rec[slot + 1].start].p, N) : post;
              int acc = 0;
              for (int sc = hr.n - 1; sc >= 0; --sc) {
                const size_t b = hr.start + (size_t)sc * per;
                auto S_at = [&](int k) { return set_at(vm->tape[b + 2 * k].p, N); };   // read by evaluation k+1
                const MhdSet Sq = S_at(0);
                const double* init = evolved_of(Sq, g);
                launch_fill(hv_init, 0.0, N, stream);
                for (int q = 0; q < K; ++q) launch_fill(hv_d[q], 0.0, N, stream);
                for (int k = K; k >= 1; --k) {
                  const MhdSet after = k < K ? S_at(k) : (sc + 1 < hr.n ? set_at(vm->tape[b + per].p, N) : out_state);
                  const MhdSet before = S_at(k - 1);
                  adj_ghost(after, lam);
                  const bool comb = K == 4 && k == 4;        // RK4: (d1 + 2 d2 + 2 d3 + d4) / 6
                  const double* D = comb ? vm->tape[b + 8].p : vm->tape[b + 2 * (k - 1) + 1].p;
                  double* tg[4] = {hv_d[0], hv_d[1], hv_d[2], hv_d[3]};
                  const double w4[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};
                  double* one[4] = {hv_d[k - 1], nullptr, nullptr, nullptr};
                  const double w1[4] = {1.0, 0.0, 0.0, 0.0};
                  const double c = cs[k - 1];
                  launch_adj_hv_update(before, g, init, D, c, c == 1.0 ? hr.h : 0.5 * hr.h, kc, mesh, lam, hv_init,
                                       comb ? 4 : 1, comb ? tg : one, comb ? w4 : w1, partm, acc, stream);
                  acc = 1;
                  // evaluation k at `before`: its timescales, then the term's reverse for hv_d[k-1]
                  launch_mhd_dt(before, st, kc, mesh, vm->dtg.p, vm->part.p, stream);
                  launch_reduce_step(vm->part.p, dt_blocks, 1.0, vm->dtmin.p, nullptr, 0, err_flag, stream);
                  visc_term_rev(vm->term_args(*this, t, before), 0, hv_d[k - 1], 1.0, nullptr, lam);
                  visc_dt_rev(before, lam);
                }
                launch_add(lam + (size_t)g * N, hv_init, N, stream);   // init is the subcycle start's g
              }
              if (acc) launch_adj_accum(partm, nb, 1.0 / (double)hr.n, gh, stream);
            }
          });
        } else {
          auto* rl = static_cast<RlModule*>(m);
          const int ig = rl->integ == "euler" ? 0 : (rl->integ == "rk2" ? 1 : 2);
          timed("adj_rl", [&] {
            launch_adj_rl(pre, st, rl->p, kc, mesh, rl->nsub, rl->h_last, ig, lam, partm, stream);
            if (rl->nsub > 0) launch_adj_accum(partm, nb, 1.0 / (double)rl->nsub, gh, stream);
          });
        }
      }
      if (any_fheat) {   // the field heating's adjoint of the step's start state
        launch_add(lam, fh_lam, 8 * N, stream);
        launch_fill(fh_lam, 0.0, 8 * N, stream);
      }
      timed("adj_dt", [&] {
        launch_adj_dt(Ln, st, kc, mesh, kc.eps, part, (rk2 || rk4) ? part2 : nullptr, gh, pmin, pidx, lam, stream);
      });
    }
  }
  hip_check(hipMemcpyAsync(grad, lam, 8 * N * 8, hipMemcpyDeviceToHost, stream), "download gradient");
  if (replay) {                                         // the state after the steps and its step size
    copy_set(kept, L.view(), N, stream);
    for (auto& o : kept_sc)
      hip_check(hipMemcpyAsync(o.first->sc.p, o.second.p, 2 * sizeof(double), hipMemcpyDeviceToDevice, stream),
                "restore outflow scalars");
    finish_step(-1);
    for (auto& m : mods)                                  // conduction's subcycle record
      if (auto* tc = dynamic_cast<TcModule*>(m.get())) adj_bytes += (int64_t)(tc->tape.size() * N * sizeof(double));
  }
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  for (TcModule* t : slot_tcs) {                        // the segment records are released
    t->slot = -1;
    t->slot_pool = DBuf();
    t->slot_n.clear();
  }
  tstep = time;
  if (clock.enabled) clock.collect();
}

void Domain::add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>& opts) {
  require(!ready, SPRUCE_E_STATE, "modules must be added before setup");
  static const char* known[] = {"ambient_heating", "radiative_losses", "thermal_conduction", "localized_heating",
                                "field_heating", "sg_filtering", "coulomb_explosion", "eic_thermalization",
                                "artificial_viscosity", "global_temperature", "anomalous_resistivity",
                                "momentum_injection", "tracer_particles", "mass_injection", "ambient_heating_sink",
                                "physical_viscosity", "div_cleaning", "boundary_outflow", "parametric_heating"};
  bool ok = false;
  for (auto* k : known) ok = ok || name == k;
  require(ok, SPRUCE_E_ARG, "Module name was not recognized");
  if (name == "thermal_conduction") mods.push_back(std::make_unique<TcModule>(opts));
  else if (name == "radiative_losses") mods.push_back(std::make_unique<RlModule>(opts));
  else if (name == "artificial_viscosity") mods.push_back(std::make_unique<ViscModule>(opts));
  else if (name == "ambient_heating") mods.push_back(std::make_unique<HeatModule>(opts));
  else if (name == "physical_viscosity") mods.push_back(std::make_unique<PvModule>(opts));
  else if (name == "localized_heating") mods.push_back(std::make_unique<LocalHeatModule>(opts));
  else if (name == "mass_injection") mods.push_back(std::make_unique<MassInjModule>(opts));
  else if (name == "momentum_injection") mods.push_back(std::make_unique<MomInjModule>(opts));
  else if (name == "ambient_heating_sink") mods.push_back(std::make_unique<HeatSinkModule>(opts));
  else if (name == "field_heating") mods.push_back(std::make_unique<FieldHeatModule>(opts));
  else if (name == "div_cleaning") mods.push_back(std::make_unique<DivCleanModule>(opts));
  else if (name == "parametric_heating") mods.push_back(std::make_unique<ParamHeatModule>(opts));
  else if (name == "anomalous_resistivity") mods.push_back(std::make_unique<ResistModule>(opts));
  else if (name == "boundary_outflow") mods.push_back(std::make_unique<OutflowModule>(opts));
  else throw Error(SPRUCE_E_UNSUPPORTED, "module " + name + " is not available in this build");
  modules.emplace_back(name, opts);
}

// Device arithmetic probe kernel.
__global__ void k_arith_probe(const double* a, const double* b, int64_t n, double* dv, double* sq, double* qd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i], y = b[i];
  dv[i] = x / y;
  sq[i] = sqrt(x);
  const double ry = 1.0 / y;
  qd[i] = qdiv(x, y, ry);
}

__global__ void k_pow_probe(const double* x, int64_t n, int which, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  out[i] = which == 0 ? pow_1p5(v) : which == 1 ? pow_2p5(v) : which == 2 ? pow_m1p5(v) : pow_third(v);
}

}  // namespace spruce

using namespace spruce;

struct spruce_gpu_domain {
  std::unique_ptr<DomainBase> d;
};

struct spruce_gpu_batch {
  std::unique_ptr<TfBatchRun> b;
};

#define API_BEGIN try {
#define API_END(ret)                                        \
  }                                                         \
  catch (const Error& e) {                                  \
    g_err = e.what();                                       \
    return e.code;                                          \
  }                                                         \
  catch (const std::exception& e) {                         \
    g_err = e.what();                                       \
    return SPRUCE_E_ARG;                                    \
  }                                                         \
  return ret;

static std::vector<std::string> strs(int32_t n, const char* const* p) {
  std::vector<std::string> v;
  for (int32_t i = 0; i < n; ++i) v.emplace_back(p[i] ? p[i] : "");
  return v;
}

static DomainBase& dom(spruce_gpu_domain* h) {
  if (!h || !h->d) throw Error(SPRUCE_E_ARG, "null domain handle");
  return *h->d;
}

static TfBatchRun& bat(spruce_gpu_batch* h) {
  if (!h || !h->b) throw Error(SPRUCE_E_ARG, "null batch handle");
  require(!h->b->broken, SPRUCE_E_STATE, "a member domain of this batch was destroyed");
  return *h->b;
}

extern "C" {

const char* spruce_gpu_last_error(void) { return g_err.c_str(); }

spruce_gpu_domain* spruce_gpu_create(const spruce_domain_desc* desc, const char* eqset, int32_t nopts,
                                     const char* const* keys, const char* const* vals) {
  try {
    if (!desc || !eqset) throw Error(SPRUCE_E_ARG, "null argument");
    auto h = std::make_unique<spruce_gpu_domain>();
    const std::string eq = eqset;
    if (eq == "ideal_2F") {
      h->d = std::make_unique<TfDomain>(*desc, strs(nopts, keys), strs(nopts, vals));
    } else if (eq == "ideal_mhd_2E") {
      h->d = std::make_unique<E2Domain>(*desc, strs(nopts, keys), strs(nopts, vals));
    } else if (eq == "ideal_mhd_cons") {
      h->d = std::make_unique<ConsDomain>(*desc);   // parseEquationSetConfigs accepts any key (idealmhdcons.cpp:10-11)
    } else {
      auto d = std::make_unique<Domain>(*desc, eqset);
      d->parse_eqset(strs(nopts, keys), strs(nopts, vals));
      h->d = std::move(d);
    }
    return h.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

int spruce_gpu_add_module(spruce_gpu_domain* h, const char* name, int32_t nopts, const char* const* keys,
                          const char* const* vals) {
  API_BEGIN
  std::vector<std::pair<std::string, std::string>> o;
  auto k = strs(nopts, keys), v = strs(nopts, vals);
  for (size_t i = 0; i < k.size(); ++i) o.emplace_back(k[i], v[i]);
  dom(h).add_module(name ? name : "", o);
  API_END(SPRUCE_OK)
}

int spruce_gpu_set_grid(spruce_gpu_domain* h, const char* name, const double* host, int64_t count) {
  API_BEGIN
  require(name && host, SPRUCE_E_ARG, "null argument");
  dom(h).set_grid(name, host, count);
  API_END(SPRUCE_OK)
}

int spruce_gpu_setup(spruce_gpu_domain* h) {
  API_BEGIN
  dom(h).setup();
  API_END(SPRUCE_OK)
}

int spruce_gpu_step(spruce_gpu_domain* h, int32_t nsteps, double* dt_out, int32_t* subcycles_out) {
  API_BEGIN
  dom(h).run(nsteps, dt_out, subcycles_out);
  API_END(SPRUCE_OK)
}

int spruce_gpu_get_grid(spruce_gpu_domain* h, const char* name, double* host, int64_t count) {
  API_BEGIN
  require(name && host, SPRUCE_E_ARG, "null argument");
  dom(h).get_grid(name, host, count);
  API_END(SPRUCE_OK)
}

int spruce_gpu_time(spruce_gpu_domain* h, double* t, int32_t* it) {
  API_BEGIN
  DomainBase& d = dom(h);
  if (t) *t = d.time;
  if (it) *it = d.iter;
  API_END(SPRUCE_OK)
}

int spruce_gpu_rhs(spruce_gpu_domain* h, double* host_out, int64_t count) {
  API_BEGIN
  require(host_out, SPRUCE_E_ARG, "null argument");
  dom(h).rhs(host_out, count);
  API_END(SPRUCE_OK)
}

int spruce_gpu_put_grid(spruce_gpu_domain* h, const char* name, const double* host, int64_t count) {
  API_BEGIN
  require(name && host, SPRUCE_E_ARG, "null argument");
  dom(h).put_grid(name, host, count);
  API_END(SPRUCE_OK)
}

int spruce_gpu_ghost_fill(spruce_gpu_domain* h) {
  API_BEGIN
  dom(h).ghost_fill();
  API_END(SPRUCE_OK)
}

int spruce_gpu_propagate(spruce_gpu_domain* h, double* step_out) {
  API_BEGIN
  const double v = dom(h).propagate();
  if (step_out) *step_out = v;
  API_END(SPRUCE_OK)
}

int spruce_gpu_rhs_noalloc(spruce_gpu_domain* h, double* host_out, int64_t count) {
  API_BEGIN
  require(host_out, SPRUCE_E_ARG, "null argument");
  dom(h).rhs_noalloc(host_out, count);
  API_END(SPRUCE_OK)
}

int spruce_gpu_dt_min(spruce_gpu_domain* h, double* step_out) {
  API_BEGIN
  require(step_out, SPRUCE_E_ARG, "null argument");
  *step_out = dom(h).dt_min();
  API_END(SPRUCE_OK)
}

int spruce_gpu_num_evolved(spruce_gpu_domain* h) {
  API_BEGIN
  return dom(h).num_evolved();
  API_END(SPRUCE_OK)
}

int spruce_gpu_stream(spruce_gpu_domain* h, void** s) {
  API_BEGIN
  require(s, SPRUCE_E_ARG, "null argument");
  *s = (void*)dom(h).stream;
  API_END(SPRUCE_OK)
}

int spruce_gpu_device_grid(spruce_gpu_domain* h, const char* name, double** p) {
  API_BEGIN
  require(name && p, SPRUCE_E_ARG, "null argument");
  *p = dom(h).device_grid(name);
  API_END(SPRUCE_OK)
}

int spruce_gpu_set_option(spruce_gpu_domain* h, const char* key, const char* value) {
  API_BEGIN
  require(key && value, SPRUCE_E_ARG, "null argument");
  dom(h).set_option(key, value);
  API_END(SPRUCE_OK)
}

int spruce_gpu_set_profiling(spruce_gpu_domain* h, int32_t enable) {
  API_BEGIN
  DomainBase& d = dom(h);
  d.clock.enabled = enable != 0;
  d.clock.acc.clear();
  API_END(SPRUCE_OK)
}

int spruce_gpu_kernel_times(spruce_gpu_domain* h, char* names, int32_t names_len, double* ms_avg, int32_t* launches,
                            int32_t max_kernels) {
  API_BEGIN
  return dom(h).clock.report(names, names_len, ms_avg, launches, max_kernels);
  API_END(SPRUCE_OK)
}

int spruce_gpu_arith_probe(const double* a, const double* b, int64_t n, double* div_out, double* sqrt_out,
                           double* qdiv_out) {
  API_BEGIN
  require(a && b && div_out && sqrt_out && qdiv_out && n >= 0, SPRUCE_E_ARG, "bad argument");
  if (n == 0) return SPRUCE_OK;
  DBuf da, db, d1, d2, d3;
  da.alloc(n); db.alloc(n); d1.alloc(n); d2.alloc(n); d3.alloc(n);
  hip_check(hipMemcpy(da.p, a, n * 8, hipMemcpyHostToDevice), "upload");
  hip_check(hipMemcpy(db.p, b, n * 8, hipMemcpyHostToDevice), "upload");
  hipLaunchKernelGGL(k_arith_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, da.p, db.p, n, d1.p, d2.p, d3.p);
  hip_check(hipGetLastError(), "launch");
  hip_check(hipMemcpy(div_out, d1.p, n * 8, hipMemcpyDeviceToHost), "download");
  hip_check(hipMemcpy(sqrt_out, d2.p, n * 8, hipMemcpyDeviceToHost), "download");
  hip_check(hipMemcpy(qdiv_out, d3.p, n * 8, hipMemcpyDeviceToHost), "download");
  API_END(SPRUCE_OK)
}

int spruce_gpu_pow_probe(const double* x, int64_t n, int32_t which, double* out) {
  API_BEGIN
  require(x && out && n >= 0 && which >= 0 && which <= 3, SPRUCE_E_ARG, "bad argument");
  if (n == 0) return SPRUCE_OK;
  DBuf dx, dout;
  dx.alloc(n); dout.alloc(n);
  hip_check(hipMemcpy(dx.p, x, n * 8, hipMemcpyHostToDevice), "upload");
  hipLaunchKernelGGL(k_pow_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx.p, n, which, dout.p);
  hip_check(hipGetLastError(), "launch");
  hip_check(hipMemcpy(out, dout.p, n * 8, hipMemcpyDeviceToHost), "download");
  API_END(SPRUCE_OK)
}

void spruce_gpu_destroy(spruce_gpu_domain* h) { delete h; }

spruce_gpu_batch* spruce_gpu_batch_create(spruce_gpu_domain* const* domains, int32_t n) {
  try {
    require(domains && n > 0, SPRUCE_E_ARG, "a batch needs at least one domain");
    std::vector<TfDomain*> m;
    for (int32_t k = 0; k < n; ++k) {
      auto* t = dynamic_cast<TfDomain*>(&dom(domains[k]));
      require(t != nullptr, SPRUCE_E_UNSUPPORTED, "batches hold ideal_2F domains");
      require(t->ready, SPRUCE_E_STATE, "batch members must be set up");
      require(t->batch == nullptr, SPRUCE_E_STATE, "domain already belongs to a batch");
      for (auto* o : m) require(o != t, SPRUCE_E_ARG, "domain listed twice");
      m.push_back(t);
    }
    auto h = std::make_unique<spruce_gpu_batch>();
    h->b = std::make_unique<TfBatchRun>(m, m[0]->shape);
    for (auto* t : m) t->batch = h->b.get();
    return h.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

int spruce_gpu_batch_set_option(spruce_gpu_batch* h, const char* key, const char* value) {
  API_BEGIN
  require(key && value, SPRUCE_E_ARG, "null argument");
  TfBatchRun& b = bat(h);
  const std::string k = key, v = value;
  if (k == "graph") {
    require(v == "true" || v == "false", SPRUCE_E_ARG, "graph must be true or false");
    b.use_graph = v == "true";
  } else if (k == "block") {
    b.rebuild(parse_tf_shape(v));
  } else if (k == "persistent") {
    require(v == "true" || v == "false", SPRUCE_E_ARG, "persistent must be true or false");
    b.persistent = v == "true";
  } else {
    throw Error(SPRUCE_E_ARG, "unknown batch option " + k);
  }
  API_END(SPRUCE_OK)
}

int spruce_gpu_batch_autotune(spruce_gpu_batch* h, int32_t trial_steps, char* chosen, int32_t chosen_len,
                              double* ms_per_step) {
  API_BEGIN
  require(trial_steps > 0, SPRUCE_E_ARG, "trial_steps must be positive");
  std::string c;
  const double ms = bat(h).autotune(trial_steps, &c);
  if (ms_per_step) *ms_per_step = ms;
  if (chosen && chosen_len > 0) {
    std::strncpy(chosen, c.c_str(), chosen_len - 1);
    chosen[chosen_len - 1] = 0;
  }
  API_END(SPRUCE_OK)
}

int spruce_gpu_batch_step(spruce_gpu_batch* h, int32_t nsteps, double* dt_out) {
  API_BEGIN
  require(nsteps >= 0, SPRUCE_E_ARG, "nsteps must be non-negative");
  if (nsteps == 0) return SPRUCE_OK;
  TfBatchRun& b = bat(h);
  std::vector<std::vector<double>> dts;
  b.run(nsteps, b.use_graph, dt_out ? &dts : nullptr);
  if (dt_out)
    for (size_t k = 0; k < dts.size(); ++k)
      for (int32_t s = 0; s < nsteps; ++s) dt_out[k * nsteps + s] = dts[k][s];
  API_END(SPRUCE_OK)
}

int spruce_gpu_batch_set_profiling(spruce_gpu_batch* h, int32_t enable) {
  API_BEGIN
  TfBatchRun& b = bat(h);
  b.clock.enabled = enable != 0;
  b.clock.acc.clear();
  API_END(SPRUCE_OK)
}

int spruce_gpu_batch_kernel_times(spruce_gpu_batch* h, char* names, int32_t names_len, double* ms_avg,
                                  int32_t* launches, int32_t max_kernels) {
  API_BEGIN
  return bat(h).clock.report(names, names_len, ms_avg, launches, max_kernels);
  API_END(SPRUCE_OK)
}

int spruce_gpu_adjoint(spruce_gpu_domain* h, int32_t nsteps, int32_t checkpoint_interval, const double* seed,
                       double* grad_out, double* dt_out) {
  API_BEGIN
  dom(h).adjoint(nsteps, checkpoint_interval, seed, grad_out, dt_out);
  API_END(SPRUCE_OK)
}

int spruce_gpu_adjoint_state(spruce_gpu_domain* h, const double* grad_evolved, double* grad_state) {
  API_BEGIN
  dom(h).adjoint_state(grad_evolved, grad_state);
  API_END(SPRUCE_OK)
}

int spruce_gpu_adjoint_device_bytes(spruce_gpu_domain* h, int64_t* bytes) {
  API_BEGIN
  require(bytes != nullptr, SPRUCE_E_ARG, "null argument");
  *bytes = dom(h).adjoint_bytes();
  API_END(SPRUCE_OK)
}

int spruce_gpu_module_report(spruce_gpu_domain* h, int32_t module, double* out, int32_t count, int32_t* count_out) {
  API_BEGIN
  require(count_out || out, SPRUCE_E_ARG, "null argument");
  const std::vector<double> v = dom(h).module_report(module);
  if (count_out) *count_out = (int32_t)v.size();
  if (out) {
    require(count >= (int32_t)v.size(), SPRUCE_E_ARG, "out holds fewer values than the module reports");
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  }
  API_END(SPRUCE_OK)
}

int spruce_gpu_module_output(spruce_gpu_domain* h, int32_t module, int32_t k, char* name, int32_t name_len,
                             double* host, int64_t count, int32_t* count_out) {
  API_BEGIN
  require(count_out || name || host, SPRUCE_E_ARG, "null argument");
  DomainBase& d = dom(h);
  const int n = d.module_output(module, -1, nullptr, nullptr, 0);
  if (count_out) *count_out = n;
  if (name || host) {
    std::string nm;
    d.module_output(module, k, &nm, host, count);
    if (name) {
      require(name_len > (int32_t)nm.size(), SPRUCE_E_ARG, "name buffer too small for " + nm);
      std::memcpy(name, nm.c_str(), nm.size() + 1);
    }
  }
  API_END(SPRUCE_OK)
}

int spruce_gpu_multispecies(spruce_gpu_domain* h, int32_t which, double* host, int64_t count) {
  API_BEGIN
  require(host != nullptr, SPRUCE_E_ARG, "null argument");
  dom(h).multispecies(which, host, count);
  API_END(SPRUCE_OK)
}

int spruce_gpu_multispecies_reset(spruce_gpu_domain* h) {
  API_BEGIN
  dom(h).multispecies_reset();
  API_END(SPRUCE_OK)
}

int spruce_gpu_adjoint_params(spruce_gpu_domain* h, double* grad_out, int32_t count, int32_t* count_out) {
  API_BEGIN
  const std::vector<double> g = dom(h).param_grad();
  require(count_out || grad_out, SPRUCE_E_ARG, "null argument");
  if (count_out) *count_out = (int32_t)g.size();
  if (grad_out) {
    require(count >= (int32_t)g.size(), SPRUCE_E_ARG, "grad_out holds fewer values than the parameters");
    std::memcpy(grad_out, g.data(), g.size() * sizeof(double));
  }
  API_END(SPRUCE_OK)
}

void spruce_gpu_batch_destroy(spruce_gpu_batch* h) {
  if (!h) return;
  if (h->b && !h->b->broken)
    for (auto* t : h->b->members) t->batch = nullptr;
  delete h;
}

}  // extern "C"
