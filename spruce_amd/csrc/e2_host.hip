// Host drivers of ideal_mhd_2E (full time step) and ideal_mhd_cons (set-up, derived variables and
// step size; the reference cannot take a step of this set, see ConsDomain::run).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>

#include "arith.hpp"
#include "e2.hpp"
#include "e2_host.hpp"

namespace spruce {

namespace {

const char* PD_NAMES_E2[7] = {"d_x", "d_y", "pos_x", "pos_y", "be_x", "be_y", "be_z"};

bool listed(const std::string& n, const char* const* l, int k) {
  for (int i = 0; i < k; ++i) if (n == l[i]) return true;
  return false;
}

// Mesh, constants and open-boundary factors shared by the two single-fluid variants
// (plasmadomain.cpp:138-161; evolution.cpp:166-167).
void build_mesh(const spruce_domain_desc& desc, int nx, int ny, std::map<std::string, std::vector<double>>& hg,
                AxisBufs& ax, AxisBufs& ay, Mesh& mesh, MhdConst& kc, double* s1, double* s2) {
  const int* bc = desc.bc;
  require(nx > 2 * N_GHOST && ny > 2 * N_GHOST, SPRUCE_E_ARG, "Grid too small for ghost zones");
  for (int k = 0; k < 4; ++k) require(bc[k] != SPRUCE_BC_OPEN_MOC, SPRUCE_E_UNSUPPORTED, "open_moc is not available");
  require((bc[0] == 0) == (bc[1] == 0) && (bc[2] == 0) == (bc[3] == 0), SPRUCE_E_ARG,
          "an axis must be periodic on both sides or on neither");
  const auto& dxg = hg["d_x"];
  const auto& dyg = hg["d_y"];
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
  const int xl = bc[0] == 0 ? 0 : N_GHOST, xu = bc[1] == 0 ? nx - 1 : nx - N_GHOST - 1;
  const int yl = bc[2] == 0 ? 0 : N_GHOST, yu = bc[3] == 0 ? ny - 1 : ny - N_GHOST - 1;
  ax.build(dx, bc[0] == 0);
  ay.build(dy, bc[2] == 0);
  mesh.x = ax.view(nx, xl, xu, bc[0] == 0);
  mesh.y = ay.view(ny, yl, yu, bc[2] == 0);
  mesh.nx = nx;
  mesh.ny = ny;
  kc.m = desc.ion_mass; kc.rm = 1.0 / desc.ion_mass;
  kc.gamma = desc.adiabatic_index; kc.gm1 = desc.adiabatic_index - 1.0; kc.rgm1 = 1.0 / kc.gm1;
  kc.nmin = desc.density_min; kc.temin = desc.thermal_energy_min; kc.tmin = desc.temp_min;
  kc.eps = desc.epsilon; kc.obs = desc.open_boundary_strength;
  kc.r4pi = 1.0 / (4.0 * PI); kc.r6 = 1.0 / 6.0; kc.r3 = 1.0 / 3.0;
  require(std::isfinite(kc.m) && kc.m > 0 && std::isnormal(kc.rm), SPRUCE_E_ARG, "ion_mass must be positive");
  require(std::isfinite(kc.gm1) && kc.gm1 > 0 && std::isnormal(kc.rgm1), SPRUCE_E_ARG, "adiabatic_index must exceed 1");
  for (int k = 0; k < 4; ++k) {
    const std::vector<double>& d = k < 2 ? dx : dy;
    const int n = (int)d.size();
    const bool upper = k & 1;
    const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
    s2[k] = std::pow(desc.open_boundary_decay_base, 0.5 * (d[e2] + d[e3]) / d[e3]);
    s1[k] = std::pow(desc.open_boundary_decay_base, 0.5 * (d[e1] + d[e2]) / d[e3]);
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// ideal_mhd_2E
// ---------------------------------------------------------------------------------------------
const char* const E2Domain::STATE[9] = {"rho", "i_temp", "e_temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x",
                                        "grav_y"};
const char* const E2Domain::DERIVED[14] = {"i_temp", "e_temp", "i_press", "e_press", "press", "v_x", "v_y",
                                           "kinetic_energy", "b_x", "b_y", "b_mag", "b_hat_x", "b_hat_y", "dt"};

E2Domain::E2Domain(const spruce_domain_desc& d, const std::vector<std::string>& keys,
                   const std::vector<std::string>& vals)
    : desc(d) {
  require(d.xdim > 0 && d.ydim > 0, SPRUCE_E_ARG, "xdim and ydim must be positive");
  for (int k = 0; k < 4; ++k) require(d.bc[k] >= 0 && d.bc[k] <= 5, SPRUCE_E_ARG, "unknown boundary kind");
  require(d.time_integrator >= 0 && d.time_integrator <= 2, SPRUCE_E_ARG, "unknown time integrator");
  for (size_t i = 0; i < keys.size(); ++i) {  // IdealMHD2E::parseEquationSetConfigs (idealmhd2E.cpp:10-20)
    if (keys[i] == "global_viscosity") (void)std::stod(vals[i]);
    else if (keys[i] == "viscosity_opt") {}
    else throw Error(SPRUCE_E_ARG, keys[i] + " is not recognized for this equation set.");
  }
  nx = d.xdim; ny = d.ydim; N = (size_t)nx * ny;
  time = d.time;
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
}

E2Domain::~E2Domain() {
  if (err_flag) (void)hipFree(err_flag);
  if (stream) (void)hipStreamDestroy(stream);
}

void E2Domain::add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>&) {
  require(!ready, SPRUCE_E_STATE, "modules must be added before setup");
  // EICThermalization takes no options (eic_thermalization.cpp:7-10)
  if (name == "eic_thermalization") { ++eic; return; }
  throw Error(SPRUCE_E_UNSUPPORTED, "module " + name + " is not available with ideal_mhd_2E in this build");
}

void E2Domain::set_grid(const std::string& name, const double* p, int64_t count) {
  require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
  require(listed(name, PD_NAMES_E2, 7) || listed(name, STATE, 9), SPRUCE_E_ARG,
          "Variable name <" + name + "> not recognized");
  host_grids[name].assign(p, p + N);
}

void E2Domain::set_option(const std::string& k, const std::string& v) {
  require(k == "stage_kernel" || k == "strip_rows" || k == "block" || k == "multispecies_mode", SPRUCE_E_ARG,
          "unknown option " + k);
  if (k == "multispecies_mode") multispecies_on = v == "true";
}

E2Set E2Domain::Set::view() { return E2Set{rho.p, n.p, mx.p, my.p, tei.p, tee.p, bx.p, by.p, rho_pre.p}; }
void E2Domain::Set::alloc(size_t n_) {
  for (DBuf* b : {&rho, &n, &mx, &my, &tei, &tee, &bx, &by, &rho_pre}) b->alloc(n_);
}

void E2Domain::setup() {
  require(!ready, SPRUCE_E_STATE, "setup called twice");
  for (auto* s : PD_NAMES_E2) require(host_grids.count(s), SPRUCE_E_ARG, std::string("PlasmaDomain grid ") + s + " not set");
  for (auto* s : STATE) require(host_grids.count(s), SPRUCE_E_ARG, std::string("state variable ") + s + " not set");
  build_mesh(desc, nx, ny, host_grids, ax, ay, mesh, kc, s1, s2);
  eic_nu = 2.0 * M_ELECTRON / desc.ion_mass;
  for (int k = 0; k < 4; ++k) {
    kinds_all[k] = desc.bc[k];
    kinds_ucnp[k] = desc.bc[k] == SPRUCE_BC_OPEN_UCNP ? desc.bc[k] : 0;
    any_open = any_open || desc.bc[k] == SPRUCE_BC_OPEN;
  }
  L.alloc(N);
  W.alloc(N);
  if (desc.time_integrator == SPRUCE_RK4) {
    W2.alloc(N);
    for (auto& b : kbuf) b.alloc(7 * N);
  }
  auto up = [&](DBuf& b, const char* nm) {
    if (b.n != N) b.alloc(N);
    hip_check(hipMemcpy(b.p, host_grids[nm].data(), N * 8, hipMemcpyHostToDevice), "upload grid");
  };
  up(bex, "be_x"); up(bey, "be_y"); up(gx, "grav_x"); up(gy, "grav_y");
  up(L.rho, "rho"); up(L.mx, "mom_x"); up(L.my, "mom_y"); up(L.bx, "bi_x"); up(L.by, "bi_y");
  DBuf ti, te;
  up(ti, "i_temp"); up(te, "e_temp");
  dt_blocks = (int)(((ny + 63) / 64) * ((nx + 3) / 4));
  partial.alloc(dt_blocks);
  step.alloc(1);
  history.alloc(1);
  hip_check(hipMalloc(&err_flag, sizeof(int)), "hipMalloc");
  hip_check(hipMemset(err_flag, 0, sizeof(int)), "hipMemset");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  // populateVariablesFromState (equationset.cpp:96-104)
  launch_e2_populate(L.view(), ti.p, te.p, kc, nx, ny, stream);
  ghost(L, kinds_all);
  finish_step(-1);
  sync_and_check();
  ready = true;
}

void E2Domain::ghost(Set& G, const int* kinds) {
  if (!(kinds[0] || kinds[1] || kinds[2] || kinds[3])) return;
  timed("e2_ghost", [&] { launch_e2_ghost(G.view(), kc, mesh, kinds, s1, s2, stream); });
}

void E2Domain::stage(Set& S, Set& B, Set& T, double mult, int mode, int kslot) {
  E2StageArgs a{};
  a.S = S.view(); a.B = B.view(); a.T = T.view();
  a.st = E2Static{bex.p, bey.p, gx.p, gy.p};
  a.k = kc; a.mesh = mesh; a.step = step.p; a.step_mult = mult; a.mode = mode;
  a.write_pre = 1;
  a.eic = eic;
  a.eic_nu = eic_nu;
  if (mode == STAGE_STORE_K) a.kbuf = kbuf[kslot].p;
  if (mode == STAGE_RK4_FINAL) { a.kbuf = kbuf[0].p; a.k2buf = kbuf[1].p; a.k3buf = kbuf[2].p; }
  timed("e2_stage", [&] { launch_e2_stage(a, stream); });
}

void E2Domain::finish_step(int idx) {
  timed("e2_dt", [&] { launch_e2_dt(L.view(), E2Static{bex.p, bey.p, gx.p, gy.p}, kc, mesh, partial.p, stream); });
  launch_reduce_step(partial.p, dt_blocks, kc.eps, step.p, idx >= 0 ? history.p : nullptr, idx, err_flag, stream);
}

void E2Domain::sync_and_check() {
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  int e = 0;
  hip_check(hipMemcpy(&e, err_flag, sizeof(int), hipMemcpyDeviceToHost), "read error flag");
  if (e) {
    hip_check(hipMemset(err_flag, 0, sizeof(int)), "hipMemset");
    throw Error(SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (no finite dt)");
  }
}

// PlasmaDomain::advanceTime (evolution.cpp:59-124).  The live-grid ghost writes of the
// intermediate propagates rewrite ghost values that the final propagate overwrites before any read,
// so only open_ucnp, which writes the stage grids, acts between stages.
void E2Domain::advance(int s) {
  const int integ = desc.time_integrator;
  if (integ == SPRUCE_EULER) {
    stage(L, L, W, 1.0, STAGE_PLAIN, 0);
    L.swap(W);
  } else if (integ == SPRUCE_RK2) {
    stage(L, L, W, 0.5, STAGE_PLAIN, 0);
    ghost(W, kinds_ucnp);
    stage(W, L, L, 1.0, STAGE_PLAIN, 0);
  } else {
    stage(L, L, W, 0.5, STAGE_STORE_K, 0);
    ghost(W, kinds_ucnp);
    stage(W, L, W2, 0.5, STAGE_STORE_K, 1);
    ghost(W2, kinds_ucnp);
    stage(W2, L, W, 1.0, STAGE_STORE_K, 2);
    ghost(W, kinds_ucnp);
    stage(W, L, L, 1.0, STAGE_RK4_FINAL, 0);
  }
  ghost(L, kinds_all);
  finish_step(s + 1);
}

void E2Domain::run(int nsteps, double* dt_out, int32_t* sub_out) {
  require(ready, SPRUCE_E_STATE, "setup must precede step");
  require(nsteps >= 0, SPRUCE_E_ARG, "nsteps must be non-negative");
  if (sub_out)   // no module of this equation set subcycles
    for (int k = 0; k < nsteps * eic; ++k) sub_out[k] = -1;
  if (nsteps == 0) return;
  if ((int)history.n < nsteps + 1) history.alloc(nsteps + 1);
  hip_check(hipMemcpyAsync(history.p, step.p, sizeof(double), hipMemcpyDeviceToDevice, stream), "copy step");
  for (int s = 0; s < nsteps; ++s) advance(s);
  sync_and_check();
  if (clock.enabled) clock.collect();
  std::vector<double> h(nsteps);
  hip_check(hipMemcpy(h.data(), history.p, nsteps * sizeof(double), hipMemcpyDeviceToHost), "read dt history");
  for (int s = 0; s < nsteps; ++s) {
    time += h[s];
    iter++;
    if (dt_out) dt_out[s] = h[s];
  }
}

void E2Domain::get_grid(const std::string& name, double* out, int64_t count) {
  require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
  if (listed(name, PD_NAMES_E2, 7)) {
    require(host_grids.count(name), SPRUCE_E_ARG, "grid " + name + " not set");
    std::memcpy(out, host_grids[name].data(), N * 8);
    return;
  }
  require(ready, SPRUCE_E_STATE, "setup must precede get_grid");
  const double* src = nullptr;
  if (name == "rho") src = L.rho.p;
  else if (name == "n") src = L.n.p;
  else if (name == "mom_x") src = L.mx.p;
  else if (name == "mom_y") src = L.my.p;
  else if (name == "i_thermal_energy") src = L.tei.p;
  else if (name == "e_thermal_energy") src = L.tee.p;
  else if (name == "bi_x") src = L.bx.p;
  else if (name == "bi_y") src = L.by.p;
  else if (name == "grav_x") src = gx.p;
  else if (name == "grav_y") src = gy.p;
  else {
    int which = -1;
    for (int k = 0; k < 14; ++k) if (name == DERIVED[k]) which = k;
    require(which >= 0, SPRUCE_E_ARG, "Variable name <" + name + "> not recognized");
    if (scratch.n != N) scratch.alloc(N);
    launch_e2_derived(L.view(), E2Static{bex.p, bey.p, gx.p, gy.p}, kc, mesh, which, scratch.p, stream);
    src = scratch.p;
  }
  hip_check(hipMemcpyAsync(out, src, N * 8, hipMemcpyDeviceToHost, stream), "download grid");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

double* E2Domain::device_grid(const std::string& name) {
  require(ready, SPRUCE_E_STATE, "setup must precede device_grid");
  if (name == "rho") return L.rho.p;
  if (name == "mom_x") return L.mx.p;
  if (name == "mom_y") return L.my.p;
  if (name == "i_thermal_energy") return L.tei.p;
  if (name == "e_thermal_energy") return L.tee.p;
  if (name == "bi_x") return L.bx.p;
  if (name == "bi_y") return L.by.p;
  throw Error(SPRUCE_E_ARG, "no device-resident grid named " + name);
}

void E2Domain::rhs(double* out, int64_t count) {
  require(ready, SPRUCE_E_STATE, "setup must precede rhs");
  require(count == (int64_t)(7 * N), SPRUCE_E_ARG, "rhs output must hold nevolved*xdim*ydim doubles");
  DBuf tmp;
  tmp.alloc(7 * N);
  launch_e2_rhs(L.view(), E2Static{bex.p, bey.p, gx.p, gy.p}, kc, mesh, tmp.p, stream, eic, eic_nu);
  hip_check(hipMemcpyAsync(out, tmp.p, 7 * N * 8, hipMemcpyDeviceToHost, stream), "download rhs");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

double E2Domain::dt_min() {
  double v = 0.0;
  hip_check(hipMemcpy(&v, step.p, 8, hipMemcpyDeviceToHost), "read step");
  return v;
}

// ---------------------------------------------------------------------------------------------
// ideal_mhd_cons
// ---------------------------------------------------------------------------------------------
const char* const ConsDomain::NAMES[23] = {"rho", "temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x", "grav_y",
                                           "n", "v_x", "v_y", "kinetic_energy", "b_x", "b_y", "b_mag", "mag_energy",
                                           "b_hat_x", "b_hat_y", "thermal_energy", "press", "press_tot", "energy",
                                           "dt"};

ConsDomain::ConsDomain(const spruce_domain_desc& d) : desc(d) {
  require(d.xdim > 0 && d.ydim > 0, SPRUCE_E_ARG, "xdim and ydim must be positive");
  for (int k = 0; k < 4; ++k) require(d.bc[k] >= 0 && d.bc[k] <= 5, SPRUCE_E_ARG, "unknown boundary kind");
  nx = d.xdim; ny = d.ydim; N = (size_t)nx * ny;
  time = d.time;
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
}

ConsDomain::~ConsDomain() {
  if (err_flag) (void)hipFree(err_flag);
  if (stream) (void)hipStreamDestroy(stream);
}

void ConsDomain::add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>&) {
  throw Error(SPRUCE_E_UNSUPPORTED, "module " + name + " is not available with ideal_mhd_cons in this build");
}

void ConsDomain::set_grid(const std::string& name, const double* p, int64_t count) {
  require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
  require(listed(name, PD_NAMES_E2, 7) || listed(name, NAMES, 23), SPRUCE_E_ARG,
          "Variable name <" + name + "> not recognized");
  host_grids[name].assign(p, p + N);
}

void ConsDomain::set_option(const std::string& k, const std::string& v) {
  require(k == "stage_kernel" || k == "strip_rows" || k == "block" || k == "multispecies_mode", SPRUCE_E_ARG,
          "unknown option " + k);
  if (k == "multispecies_mode") multispecies_on = v == "true";
}

void ConsDomain::setup() {
  require(!ready, SPRUCE_E_STATE, "setup called twice");
  for (auto* s : PD_NAMES_E2) require(host_grids.count(s), SPRUCE_E_ARG, std::string("PlasmaDomain grid ") + s + " not set");
  for (int k = 0; k < 8; ++k)
    require(host_grids.count(NAMES[k]), SPRUCE_E_ARG, std::string("state variable ") + NAMES[k] + " not set");
  // press_tot is never computed by the equation set (idealmhdcons.cpp:96-128), so the reference
  // asserts at set-up unless the .state file provides it (equationset.cpp:89-90).
  require(host_grids.count("press_tot"), SPRUCE_E_ARG,
          "All variables must be initialized by <populateVariablesFromState> (press_tot)");
  build_mesh(desc, nx, ny, host_grids, ax, ay, mesh, kc, s1, s2);
  for (auto& b : g) b.alloc(N);
  auto up = [&](DBuf& b, const char* nm) {
    if (b.n != N) b.alloc(N);
    hip_check(hipMemcpy(b.p, host_grids[nm].data(), N * 8, hipMemcpyHostToDevice), "upload grid");
  };
  for (int k = 0; k < 8; ++k) up(g[k], NAMES[k]);
  up(g[20], "press_tot");
  up(bex, "be_x"); up(bey, "be_y");
  dt_blocks = 1;
  step.alloc(1);
  hip_check(hipMalloc(&err_flag, sizeof(int)), "hipMalloc");
  hip_check(hipMemset(err_flag, 0, sizeof(int)), "hipMemset");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  ConsGrids v = view();
  launch_cons_populate(v, kc, nx, ny, stream);
  launch_cons_ghost(v, kc, mesh, desc.bc, s1, s2, stream);
  launch_cons_derived(v, kc, mesh, stream);
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  // step size epsilon * min(dt over the dt bounds) (evolution.cpp:62; grid.cpp:71-82)
  std::vector<double> dt(N);
  hip_check(hipMemcpy(dt.data(), g[22].p, N * 8, hipMemcpyDeviceToHost), "read dt");
  double m = DBL_MAX;
  for (int i = mesh.x.lo; i <= mesh.x.hi; ++i)
    for (int j = mesh.y.lo; j <= mesh.y.hi; ++j) { const double x = dt[(size_t)i * ny + j]; m = (x < m) ? x : m; }
  require(m < DBL_MAX, SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (no finite dt)");
  step_h = kc.eps * m;
  ready = true;
}

ConsGrids ConsDomain::view() {
  ConsGrids v{};
  double** f[21] = {&v.rho, &v.temp, &v.mx, &v.my, &v.bx, &v.by, nullptr, nullptr, &v.n, &v.vx, &v.vy, &v.ke,
                    &v.b_x, &v.b_y, &v.bmag, &v.mag, &v.bhx, &v.bhy, &v.te, &v.press, nullptr};
  for (int k = 0; k < 21; ++k) if (f[k]) *f[k] = g[k].p;
  v.energy = g[21].p;
  v.dt = g[22].p;
  v.rho_pre = g[23].p;
  v.bex = bex.p;
  v.bey = bey.p;
  return v;
}

// IdealMHDCons::computeTimeDerivativesDerived returns 5 derivative grids for its 6 evolved variables
// (idealmhdcons.cpp:48-49), and the reference aborts at the first time derivative: there is no step
// of this equation set to reproduce.
void ConsDomain::run(int, double*, int32_t*) {
  throw Error(SPRUCE_E_UNSUPPORTED,
              "ideal_mhd_cons cannot be stepped: the reference time derivative returns 5 grids for 6 evolved "
              "variables (idealmhdcons.cpp:48-49) and aborts");
}
void ConsDomain::rhs(double*, int64_t) { run(0, nullptr, nullptr); }

void ConsDomain::get_grid(const std::string& name, double* out, int64_t count) {
  require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
  if (listed(name, PD_NAMES_E2, 7)) {
    require(host_grids.count(name), SPRUCE_E_ARG, "grid " + name + " not set");
    std::memcpy(out, host_grids[name].data(), N * 8);
    return;
  }
  require(ready, SPRUCE_E_STATE, "setup must precede get_grid");
  int k = -1;
  for (int q = 0; q < 23; ++q) if (name == NAMES[q]) k = q;
  require(k >= 0, SPRUCE_E_ARG, "Variable name <" + name + "> not recognized");
  hip_check(hipMemcpy(out, g[k].p, N * 8, hipMemcpyDeviceToHost), "download grid");
}

double* ConsDomain::device_grid(const std::string& name) {
  for (int q = 0; q < 23; ++q) if (name == NAMES[q]) return g[q].p;
  throw Error(SPRUCE_E_ARG, "no device-resident grid named " + name);
}

}  // namespace spruce
