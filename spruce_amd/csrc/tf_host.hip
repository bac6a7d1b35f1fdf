// Host driver of the two-fluid time step: TfDomain (one PlasmaDomain with the ideal_2F equation set)
// and TfBatchRun (many such domains advanced by the same launches).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "arith.hpp"
#include "tf_host.hpp"

namespace spruce {

namespace {

const char* PD_NAMES_TF[7] = {"d_x", "d_y", "pos_x", "pos_y", "be_x", "be_y", "be_z"};
// ideal2F.hpp:36-38 state variables and the evolved order of twofluid.hpp
const char* TF_STATE[16] = {"i_rho", "e_rho", "i_mom_x", "i_mom_y", "e_mom_x", "e_mom_y", "i_temp", "e_temp",
                            "bi_x", "bi_y", "bi_z", "E_x", "E_y", "E_z", "grav_x", "grav_y"};
const char* TF_EVOLVED[TF_NE] = {"i_rho", "e_rho", "i_mom_x", "i_mom_y", "e_mom_x", "e_mom_y", "i_thermal_energy",
                                 "e_thermal_energy", "E_x", "E_y", "E_z", "bi_x", "bi_y", "bi_z"};
// variables computed by k_tf_derived, in its `which` order
const char* TF_DERIVED[33] = {"i_temp", "e_temp", "i_n", "e_n", "i_v_x", "i_v_y", "e_v_x", "e_v_y", "j_x", "j_y",
                              "i_press", "e_press", "press", "rho", "rho_c", "n", "dn", "dt", "dt_i", "b_x", "b_y",
                              "b_z", "b_mag", "b_mag_xy", "b_hat_x", "b_hat_y", "curlE_z", "divE", "divB", "i_dPdx",
                              "e_dPdx", "grav_x", "grav_y"};

bool in_list(const std::string& n, const char* const* l, int k) {
  for (int i = 0; i < k; ++i) if (n == l[i]) return true;
  return false;
}

}  // namespace

TfDomain::TfDomain(const spruce_domain_desc& d, const std::vector<std::string>& keys,
                   const std::vector<std::string>& vals)
    : desc(d) {
  require(d.xdim > 0 && d.ydim > 0, SPRUCE_E_ARG, "xdim and ydim must be positive");
  for (int k = 0; k < 4; ++k) require(d.bc[k] >= 0 && d.bc[k] <= 5, SPRUCE_E_ARG, "unknown boundary kind");
  require(d.time_integrator >= 0 && d.time_integrator <= 2, SPRUCE_E_ARG, "unknown time integrator");
  for (size_t i = 0; i < keys.size(); ++i) {  // Ideal2F::parseEquationSetConfigs (ideal2F.cpp:10-22)
    if (keys[i] == "use_sub_cycling") use_sub_cycling = vals[i] == "true";
    else if (keys[i] == "epsilon_courant" || keys[i] == "verbose_2F") {}
    else if (keys[i] == "remove_curl_terms") remove_curl = vals[i] == "true";
    else throw Error(SPRUCE_E_ARG, keys[i] + " is not recognized for this equation set.");
  }
  nx = d.xdim;
  ny = d.ydim;
  N = (size_t)nx * ny;
  time = d.time;
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
}

TfDomain::~TfDomain() {
  if (batch) batch->broken = true;
  self.reset();
  if (counter) (void)hipFree(counter);
  if (stream) (void)hipStreamDestroy(stream);
}

void TfDomain::set_grid(const std::string& name, const double* p, int64_t count) {
  require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
  require(in_list(name, PD_NAMES_TF, 7) || in_list(name, TF_STATE, 16), SPRUCE_E_ARG,
          "Variable name <" + name + "> not recognized");
  host_grids[name].assign(p, p + N);
}

TfShape parse_tf_shape(const std::string& v) {
  TfShape sh;
  if (v.rfind("strip", 0) == 0) {
    const int r = std::stoi(v.substr(5));
    require(r >= 4 && r <= 64 && r % 4 == 0, SPRUCE_E_ARG, "strip rows must be a multiple of 4 in [4, 64]");
    sh.R = r;
    return sh;
  }
  const auto x = v.find('x');
  require(x != std::string::npos, SPRUCE_E_ARG, "block must be <columns>x<rows> or strip<rows>");
  const int bx = std::stoi(v.substr(0, x)), by = std::stoi(v.substr(x + 1));
  require(bx > 0 && by > 0 && bx * by <= 256 && (bx * by) % 64 == 0, SPRUCE_E_ARG,
          "block must hold 64, 128, 192 or 256 threads");
  sh.bx = bx;
  sh.by = by;
  return sh;
}

void TfDomain::set_option(const std::string& k, const std::string& v) {
  if (k == "block") {
    shape = parse_tf_shape(v);
    if (self) self->rebuild(shape);
  } else if (k == "stage_kernel" || k == "strip_rows") {
    // IdealMHD options; the two-fluid path has a single stage kernel
  } else if (k == "multispecies_mode") {
    multispecies_on = v == "true";
  } else {
    throw Error(SPRUCE_E_ARG, "unknown option " + k);
  }
}

void TfDomain::fill_descriptor() {
  TfDom& D = dom;
  for (int v = 0; v < TF_NE; ++v) { D.L.v[v] = L[v].p; D.W.v[v] = W[v].p; D.W2.v[v] = W2[v].p; }
  for (int k = 0; k < 3; ++k) D.k[k] = K[k].p;
  D.bex = bex.p; D.bey = bey.p; D.gx = gx.p; D.gy = gy.p;
  D.kx3 = kx3.p; D.kxv = kxv.p; D.ky3 = ky3.p; D.kyv = kyv.p;
  D.step = step.p; D.history = history.p; D.hist_cap = (int)history.n; D.counter = counter; D.partial = partial.p;
}

void TfDomain::setup() {
  require(!ready, SPRUCE_E_STATE, "setup called twice");
  for (auto* s : PD_NAMES_TF) require(host_grids.count(s), SPRUCE_E_ARG, std::string("PlasmaDomain grid ") + s + " not set");
  for (auto* s : TF_STATE) require(host_grids.count(s), SPRUCE_E_ARG, std::string("state variable ") + s + " not set");
  // The reference asserts at the first time derivative when the Maxwell subcycling is requested
  // (ideal2F.cpp:73-95 leaves 1x1 field derivatives), so that combination is refused here.
  require(!use_sub_cycling || remove_curl, SPRUCE_E_UNSUPPORTED,
          "ideal_2F with use_sub_cycling = true has no field derivatives (the reference asserts); set it false");
  const int* bc = desc.bc;
  require(nx > 2 * N_GHOST && ny > 2 * N_GHOST, SPRUCE_E_ARG, "Grid too small for ghost zones");
  for (int k = 0; k < 4; ++k) require(bc[k] != SPRUCE_BC_OPEN_MOC, SPRUCE_E_UNSUPPORTED, "open_moc is not available");
  require((bc[0] == 0) == (bc[1] == 0) && (bc[2] == 0) == (bc[3] == 0), SPRUCE_E_ARG,
          "an axis must be periodic on both sides or on neither");
  const int xl = bc[0] == 0 ? 0 : N_GHOST, xu = bc[1] == 0 ? nx - 1 : nx - N_GHOST - 1;
  const int yl = bc[2] == 0 ? 0 : N_GHOST, yu = bc[3] == 0 ? ny - 1 : ny - N_GHOST - 1;
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
  ax.build(dx, bc[0] == 0);
  ay.build(dy, bc[2] == 0);
  TfDom& D = dom;
  D.mesh.x = ax.view(nx, xl, xu, bc[0] == 0);
  D.mesh.y = ay.view(ny, yl, yu, bc[2] == 0);
  D.mesh.nx = nx;
  D.mesh.ny = ny;
  for (int k = 0; k < 4; ++k) D.bc[k] = bc[k];
  TfConst& c = D.c;
  c.mi = desc.ion_mass; c.rmi = 1.0 / desc.ion_mass;
  c.me = M_ELECTRON; c.rme = 1.0 / M_ELECTRON;
  c.gamma = desc.adiabatic_index; c.gm1 = desc.adiabatic_index - 1.0; c.rgm1 = 1.0 / c.gm1;
  c.nmin = desc.density_min; c.temin = desc.thermal_energy_min; c.eps = desc.epsilon;
  c.obs = desc.open_boundary_strength;
  c.rC = 1.0 / C_LIGHT; c.r6 = 1.0 / 6.0; c.r3 = 1.0 / 3.0;
  c.curl = remove_curl ? 0 : 1;
  c.em_dt = (!use_sub_cycling && !remove_curl) ? 1 : 0;          // ideal2F.cpp:195
  for (int k = 0; k < 4; ++k) {
    const std::vector<double>& d = k < 2 ? dx : dy;
    const int n = (int)d.size();
    const bool upper = k & 1;
    const int e1 = upper ? n - 1 : 0, e2 = upper ? n - 2 : 1, e3 = upper ? n - 3 : 2;
    D.s2[k] = std::pow(desc.open_boundary_decay_base, 0.5 * (d[e2] + d[e3]) / d[e3]);  // evolution.cpp:166-167
    D.s1[k] = std::pow(desc.open_boundary_decay_base, 0.5 * (d[e1] + d[e2]) / d[e3]);
  }
  // Langmuir-wave geometry: k = 2 pi / (2 d) and 3 k^2 (ideal2F.cpp:175-182)
  auto kgeom = [](const std::vector<double>& d, DBuf& k3, DBuf& kv) {
    std::vector<double> a(d.size()), b(d.size());
    for (size_t i = 0; i < d.size(); ++i) {
      const double k = 2. * PI / (d[i] * 2);
      b[i] = k;
      a[i] = 3 * (k * k);
    }
    k3.alloc(d.size()); kv.alloc(d.size());
    hip_check(hipMemcpy(k3.p, a.data(), a.size() * 8, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(kv.p, b.data(), b.size() * 8, hipMemcpyHostToDevice), "upload");
  };
  kgeom(dx, kx3, kxv);
  kgeom(dy, ky3, kyv);

  auto up = [&](DBuf& b, const std::string& name) {
    b.alloc(N);
    hip_check(hipMemcpy(b.p, host_grids[name].data(), N * 8, hipMemcpyHostToDevice), "upload grid");
  };
  for (int v = 0; v < TF_NE; ++v) {
    W[v].alloc(N);
    if (desc.time_integrator == SPRUCE_RK4) W2[v].alloc(N);
    const std::string nm = TF_EVOLVED[v];
    if (nm == "i_thermal_energy" || nm == "e_thermal_energy") L[v].alloc(N);
    else up(L[v], nm);
  }
  if (desc.time_integrator == SPRUCE_RK4) for (auto& k : K) k.alloc(TF_NE * N);
  up(bex, "be_x"); up(bey, "be_y"); up(gx, "grav_x"); up(gy, "grav_y");
  up(itemp, "i_temp"); up(etemp, "e_temp");
  step.alloc(1);
  history.alloc(2);
  hip_check(hipMalloc(&counter, sizeof(int)), "hipMalloc");
  hip_check(hipMemset(counter, 0, sizeof(int)), "hipMemset");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  fill_descriptor();
  self = std::make_unique<TfBatchRun>(std::vector<TfDomain*>{this}, shape);
  // populateVariablesFromState (equationset.cpp:96-104)
  const double* hp[2] = {itemp.p, etemp.p};
  double** dp = nullptr;
  hip_check(hipMalloc(&dp, 2 * sizeof(double*)), "hipMalloc");
  hip_check(hipMemcpy(dp, hp, 2 * sizeof(double*), hipMemcpyHostToDevice), "upload");
  launch_tf_populate(self->b, (const double* const*)dp, (const double* const*)(dp + 1), stream);
  launch_tf_ghost(self->b, TF_SEL_L, 0x3F, self->maxn, stream);
  launch_tf_dt(self->b, stream);  // counter is 0: the first step size lands in history[0]
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  (void)hipFree(dp);
  double h0 = 0.0;
  hip_check(hipMemcpy(&h0, history.p, 8, hipMemcpyDeviceToHost), "read");
  require(!std::isnan(h0), SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (no finite dt)");
  ready = true;
}

void TfDomain::add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>&) {
  throw Error(SPRUCE_E_UNSUPPORTED, "module " + name + " is not available with ideal_2F in this build");
}

void TfDomain::run(int nsteps, double* dt_out, int32_t*) {
  require(ready, SPRUCE_E_STATE, "setup must precede step");
  require(!batch, SPRUCE_E_STATE, "domain is part of a batch; step the batch");
  if (nsteps <= 0) return;
  std::vector<std::vector<double>> dts;
  self->run(nsteps, false, &dts);
  for (int s = 0; s < nsteps; ++s) if (dt_out) dt_out[s] = dts[0][s];
}

double TfDomain::dt_min() {
  double v = 0.0;
  hip_check(hipMemcpy(&v, step.p, 8, hipMemcpyDeviceToHost), "read step");
  return v;
}

double* TfDomain::device_grid(const std::string& name) {
  for (int v = 0; v < TF_NE; ++v) if (name == TF_EVOLVED[v]) return L[v].p;
  throw Error(SPRUCE_E_ARG, "no device-resident grid named " + name);
}

void TfDomain::get_grid(const std::string& name, double* out, int64_t count) {
  require(count == (int64_t)N, SPRUCE_E_ARG, "grid " + name + ": count must equal xdim*ydim");
  if (in_list(name, PD_NAMES_TF, 7)) {
    require(host_grids.count(name), SPRUCE_E_ARG, "grid " + name + " not set");
    std::memcpy(out, host_grids[name].data(), N * 8);
    return;
  }
  require(ready, SPRUCE_E_STATE, "setup must precede get_grid");
  const double* src = nullptr;
  for (int v = 0; v < TF_NE; ++v) if (name == TF_EVOLVED[v]) src = L[v].p;
  if (!src) {
    int which = -1;
    for (int k = 0; k < 33; ++k) if (name == TF_DERIVED[k]) which = k;
    require(which >= 0, SPRUCE_E_ARG, "Variable name <" + name + "> not recognized");
    if (scratch.n < N) scratch.alloc(N);
    launch_tf_derived(self->b, 0, which, scratch.p, nx, ny, stream);
    src = scratch.p;
  }
  hip_check(hipMemcpyAsync(out, src, N * 8, hipMemcpyDeviceToHost, stream), "download grid");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void TfDomain::rhs(double* out, int64_t count) {
  require(ready, SPRUCE_E_STATE, "setup must precede rhs");
  require(count == (int64_t)(TF_NE * N), SPRUCE_E_ARG, "rhs output must hold nevolved*xdim*ydim doubles");
  DBuf tmp;
  tmp.alloc(TF_NE * N);
  launch_tf_rhs(self->b, 0, tmp.p, nx, ny, stream);
  hip_check(hipMemcpyAsync(out, tmp.p, TF_NE * N * 8, hipMemcpyDeviceToHost, stream), "download rhs");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------------------------

TfBatchRun::TfBatchRun(const std::vector<TfDomain*>& m, TfShape sh) : members(m) {
  require(!members.empty(), SPRUCE_E_ARG, "empty batch");
  integrator = members[0]->desc.time_integrator;
  for (auto* d : members) {
    require(d->desc.time_integrator == integrator, SPRUCE_E_ARG, "batched domains must share the time integrator");
    for (int k = 0; k < 4; ++k) {
      any_ucnp = any_ucnp || d->desc.bc[k] == SPRUCE_BC_OPEN_UCNP;
      any_other = any_other || (d->desc.bc[k] != 0 && d->desc.bc[k] != SPRUCE_BC_OPEN_UCNP);
    }
    maxn = std::max(maxn, std::max(d->nx, d->ny));
  }
  rebuild(sh);
}

TfBatchRun::~TfBatchRun() {
  destroy_graph();
  if (d_doms) (void)hipFree(d_doms);
  if (d_tile_dom) (void)hipFree(d_tile_dom);
  if (d_strip_dom) (void)hipFree(d_strip_dom);
}

void TfBatchRun::destroy_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
}

void TfBatchRun::rebuild(TfShape sh) {
  destroy_graph();
  shape = sh;
  const int bx = sh.bx, by = sh.by;
  std::vector<TfDom> doms;
  std::vector<int> tile_dom, strip_dom;
  int base = 0, sbase = 0;
  for (size_t k = 0; k < members.size(); ++k) {
    TfDomain* d = members[k];
    TfDom D = d->dom;
    D.tiles_x = (d->ny + bx - 1) / bx;
    D.tiles_y = (d->nx + by - 1) / by;
    D.tile_base = base;
    const int nt = D.tiles_x * D.tiles_y;
    if ((int)d->partial.n < nt) { d->partial.alloc(nt); d->dom.partial = d->partial.p; }
    D.partial = d->partial.p;
    D.history = d->history.p;
    D.hist_cap = (int)d->history.n;
    d->dom.tiles_x = D.tiles_x; d->dom.tiles_y = D.tiles_y;
    for (int t = 0; t < nt; ++t) tile_dom.push_back((int)k);
    base += nt;
    D.strip_base = sbase;
    D.strips_x = sh.R > 0 ? (d->nx + sh.R - 1) / sh.R : 0;
    const int ns = sh.R > 0 ? D.strips_x * ((d->ny + 59) / 60) : 0;   // 60 output columns per strip
    for (int t = 0; t < ns; ++t) strip_dom.push_back((int)k);
    sbase += ns;
    doms.push_back(D);
  }
  uploaded = doms;
  if (d_doms) (void)hipFree(d_doms);
  if (d_tile_dom) (void)hipFree(d_tile_dom);
  if (d_strip_dom) (void)hipFree(d_strip_dom);
  d_strip_dom = nullptr;
  hip_check(hipMalloc(&d_doms, doms.size() * sizeof(TfDom)), "hipMalloc");
  hip_check(hipMalloc(&d_tile_dom, tile_dom.size() * sizeof(int)), "hipMalloc");
  hip_check(hipMemcpy(d_doms, doms.data(), doms.size() * sizeof(TfDom), hipMemcpyHostToDevice), "upload");
  hip_check(hipMemcpy(d_tile_dom, tile_dom.data(), tile_dom.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
  if (!strip_dom.empty()) {
    hip_check(hipMalloc(&d_strip_dom, strip_dom.size() * sizeof(int)), "hipMalloc");
    hip_check(hipMemcpy(d_strip_dom, strip_dom.data(), strip_dom.size() * sizeof(int), hipMemcpyHostToDevice),
              "upload");
  }
  b.doms = d_doms;
  b.tile_dom = d_tile_dom;
  b.ndom = (int)members.size();
  b.ntiles = base;
  b.bx = bx;
  b.by = by;
  b.strip_dom = d_strip_dom;
  b.nstrip_tiles = sbase;
  b.strip_rows = sh.R;
}

// One PlasmaDomain::advanceTime of every domain (evolution.cpp:59-124); no host interaction, so a
// HIP graph can capture it.
void TfBatchRun::enqueue_step(hipStream_t s) {
  const int ucnp = 1 << SPRUCE_BC_OPEN_UCNP;
  auto stage = [&](int S, int B, int T, double mult, int mode, int kslot) {
    clock.timed(s, "tf_stage", [&] { launch_tf_stage(b, S, B, T, mult, mode, kslot, s); });
  };
  auto ghost = [&](int G, int kinds) {
    clock.timed(s, "tf_ghost", [&] { launch_tf_ghost(b, G, kinds, maxn, s); });
  };
  if (integrator == SPRUCE_EULER) {
    stage(TF_SEL_L, TF_SEL_L, TF_SEL_W, 1.0, TF_STAGE_PLAIN, 0);
    clock.timed(s, "tf_copy", [&] { launch_tf_copy(b, TF_SEL_W, TF_SEL_L, s); });
  } else if (integrator == SPRUCE_RK2) {
    stage(TF_SEL_L, TF_SEL_L, TF_SEL_W, 0.5, TF_STAGE_PLAIN, 0);
    if (any_ucnp) ghost(TF_SEL_W, ucnp);
    stage(TF_SEL_W, TF_SEL_L, TF_SEL_L, 1.0, TF_STAGE_PLAIN, 0);
  } else {
    // The live-grid ghost writes of the intermediate propagates (evolution.cpp:168-281) rewrite
    // the values the previous final propagate left there (the two-fluid derived step changes no
    // evolved variable), so only open_ucnp, which writes the stage grids, acts here.
    stage(TF_SEL_L, TF_SEL_L, TF_SEL_W, 0.5, TF_STAGE_STORE_K, 0);
    if (any_ucnp) ghost(TF_SEL_W, ucnp);
    stage(TF_SEL_W, TF_SEL_L, TF_SEL_W2, 0.5, TF_STAGE_STORE_K, 1);
    if (any_ucnp) ghost(TF_SEL_W2, ucnp);
    stage(TF_SEL_W2, TF_SEL_L, TF_SEL_W, 1.0, TF_STAGE_STORE_K, 2);
    if (any_ucnp) ghost(TF_SEL_W, ucnp);
    stage(TF_SEL_W, TF_SEL_L, TF_SEL_L, 1.0, TF_STAGE_RK4_FINAL, 0);
  }
  if (any_ucnp || any_other) ghost(TF_SEL_L, 0x3F);
  clock.timed(s, "tf_dt", [&] { launch_tf_dt(b, s); });
}

void TfBatchRun::run(int nsteps, bool use_graph, std::vector<std::vector<double>>* dts) {
  hipStream_t s = members[0]->stream;
  bool realloc = false;
  for (size_t k = 0; k < members.size(); ++k) {
    TfDomain* d = members[k];
    if ((int)d->history.n < nsteps + 1) {
      d->history.alloc(nsteps + 1);
      d->dom.history = d->history.p;
      d->dom.hist_cap = (int)d->history.n;
    }
    // another batch holding this domain may have replaced its history or partial buffers
    realloc = realloc || uploaded[k].history != d->history.p || uploaded[k].hist_cap != (int)d->history.n ||
              uploaded[k].partial != d->partial.p;
  }
  if (realloc) rebuild(shape);
  if (persistent && !(shape.R > 0 && shape.bx == 64 && shape.by == 4)) {
    TfShape sh;                                      // the persistent form runs strips, 64x4 cell tiles
    sh.R = shape.R > 0 ? shape.R : 4;
    rebuild(sh);
  }
  launch_tf_reset_counters(b, s);
  if (persistent) {
    if (persistent_blocks == 0) persistent_blocks = tf_persistent_blocks();
    // every phase loops over its work items with the grid's stride, so any grid size is correct;
    // a multiple of 8 keeps each block's strip tiles on one XCD band
    const int work = std::max(b.nstrip_tiles, std::max(b.ntiles, b.ndom));
    int blocks = std::min(persistent_blocks, ((work + 7) / 8) * 8);
    blocks = std::max(8, blocks - blocks % 8);
    blocks = std::min(blocks, persistent_blocks);
    const int ucnp_mask = any_ucnp ? 1 << SPRUCE_BC_OPEN_UCNP : 0, final_mask = (any_ucnp || any_other) ? 0x3F : 0;
    clock.timed(s, "tf_persistent", [&] {
      launch_tf_persistent(b, nsteps, integrator, ucnp_mask, final_mask, maxn, blocks, s);
    });
  } else if (use_graph && !clock.enabled) {
    if (!gexec) {
      hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal), "hipStreamBeginCapture");
      enqueue_step(s);
      hip_check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    for (int k = 0; k < nsteps; ++k) hip_check(hipGraphLaunch(gexec, s), "hipGraphLaunch");
  } else {
    for (int k = 0; k < nsteps; ++k) enqueue_step(s);
  }
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (clock.enabled) clock.collect();
  check_history();
  if (dts) dts->clear();
  for (auto* d : members) {
    std::vector<double> h(nsteps + 1);
    hip_check(hipMemcpy(h.data(), d->history.p, (nsteps + 1) * 8, hipMemcpyDeviceToHost), "read dt history");
    for (int k = 1; k <= nsteps; ++k)
      require(!std::isnan(h[k]), SPRUCE_E_NUMERIC, "Something went wrong in the min calculation (no finite dt)");
    for (int k = 0; k < nsteps; ++k) { d->time += h[k]; d->iter++; }
    if (dts) dts->emplace_back(h.begin(), h.begin() + nsteps);
  }
}

// The step-size reduction stores a step size only while its counter is below the history's
// capacity; a run that stepped further than the history holds is an error of this class.
void TfBatchRun::check_history() {
  for (size_t k = 0; k < members.size(); ++k) {
    int cnt = 0;
    hip_check(hipMemcpy(&cnt, members[k]->counter, sizeof(int), hipMemcpyDeviceToHost), "read step counter");
    const int cap = std::min(uploaded[k].hist_cap, (int)members[k]->history.n);
    require(cnt <= cap, SPRUCE_E_STATE,
            "step-size history overflow: " + std::to_string(cnt) + " steps, capacity " + std::to_string(cap));
  }
}

// Times each candidate block shape on a copy of the batch state and keeps the fastest; the state
// is restored afterwards, so tuning does not advance the simulation.
double TfBatchRun::autotune(int trial_steps, std::string* chosen) {
  hipStream_t s = members[0]->stream;
  struct Save { std::vector<DBuf*> src; std::vector<std::unique_ptr<DBuf>> copy; double t; int it; };
  std::vector<Save> saves(members.size());
  for (size_t k = 0; k < members.size(); ++k) {
    TfDomain* d = members[k];
    for (int v = 0; v < TF_NE; ++v) saves[k].src.push_back(&d->L[v]);
    saves[k].src.push_back(&d->step);
    for (DBuf* b0 : saves[k].src) {
      auto c = std::make_unique<DBuf>();
      c->alloc(b0->n);
      hip_check(hipMemcpy(c->p, b0->p, b0->n * 8, hipMemcpyDeviceToDevice), "save");
      saves[k].copy.push_back(std::move(c));
    }
    saves[k].t = d->time;
    saves[k].it = d->iter;
  }
  std::vector<TfShape> shapes;
  for (auto& bxy : std::vector<std::pair<int, int>>{{64, 1}, {64, 2}, {64, 4}, {128, 1}, {128, 2}, {32, 2}, {32, 4},
                                                    {32, 8}, {16, 4}, {16, 8}}) {
    TfShape t;
    t.bx = bxy.first;
    t.by = bxy.second;
    shapes.push_back(t);
  }
  for (int r : {4, 8, 16}) {
    TfShape t;
    t.R = r;
    shapes.push_back(t);
  }
  double best = 1e300;
  TfShape bestsh = shape;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  // every timed trial stores trial_steps step sizes after the reset below
  for (auto* d : members) {
    if ((int)d->history.n < trial_steps + 1) {
      d->history.alloc(trial_steps + 1);
      d->dom.history = d->history.p;
      d->dom.hist_cap = (int)d->history.n;
    }
  }
  for (auto& sh : shapes) {
    rebuild(sh);
    run(1, false, nullptr);  // warm
    launch_tf_reset_counters(b, s);
    hip_check(hipEventRecord(e0, s), "record");
    for (int k = 0; k < trial_steps; ++k) enqueue_step(s);
    hip_check(hipEventRecord(e1, s), "record");
    hip_check(hipEventSynchronize(e1), "sync");
    check_history();
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (ms < best) { best = ms; bestsh = sh; }
    for (size_t k = 0; k < members.size(); ++k)
      for (size_t q = 0; q < saves[k].src.size(); ++q)
        hip_check(hipMemcpy(saves[k].src[q]->p, saves[k].copy[q]->p, saves[k].src[q]->n * 8, hipMemcpyDeviceToDevice),
                  "restore");
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (size_t k = 0; k < members.size(); ++k) { members[k]->time = saves[k].t; members[k]->iter = saves[k].it; }
  rebuild(bestsh);
  for (auto* d : members) d->shape = bestsh;
  if (chosen) *chosen = bestsh.str();
  return best / trial_steps;
}

}  // namespace spruce
