// spruce_run: the reference's command-line program on the GPU library.
//
// Same command line as the reference executable (main.cpp:16-90, utils.cpp:37-50):
//   spruce_run -m input -o <out dir> [-c <file>.config] [-s <file>.state] [-d <duration>] [-r <runtime>]
//   spruce_run -m continue -o <previous out dir> -d <duration> [-r <runtime>]
// It reads the same .config and .state files (fileio.cpp:14-122, equationset.cpp:38-85,
// module.cpp:11-35), advances the same PlasmaDomain::run loop (evolution.cpp:8-57) through
// libspruce_gpu.so, and writes the same mhd.out, init.state and end.state files with the same stream
// formatting (fileio.cpp:124-256, grid.cpp:412-427, plasmadomain.hpp:248-256), so with bit-identical
// grids the files are byte-identical.
//
// Differences, all on paths where the reference asserts or hangs: invalid input is reported on stderr
// with exit status 1 instead of an abort; an empty or comment line inside an equation-set block is
// skipped (the reference loops forever there, equationset.cpp:72-73).  When the run reaches its time
// or iteration limit the reference ends on a failed assert "to end recursive scripts"
// (evolution.cpp:54-56); this program prints the same message and exits with status 134, the status
// a shell reports for that abort.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "spruce_gpu.h"

namespace fs = std::filesystem;

namespace {

struct Fail : std::runtime_error {
  explicit Fail(const std::string& m) : std::runtime_error(m) {}
};
void check(bool ok, const std::string& msg) {
  if (!ok) throw Fail(msg);
}
void gpu(int rc, const std::string& what) {
  if (rc != SPRUCE_OK) throw Fail(what + ": " + spruce_gpu_last_error());
}

// utils.cpp:8-16
void clear_ws(std::string& s) {
  s.erase(std::remove_if(s.begin(), s.end(),
                         [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\b' ||
                                             c == '\r' || c == '\f' || c == '\a'; }),
          s.end());
}
// utils.cpp:20-25
std::istream& cleaned_line(std::istream& is, std::string& s) {
  std::getline(is, s);
  clear_ws(s);
  return is;
}
// utils.cpp:28-36: empty elements dropped
std::vector<std::string> split(const std::string& s, char d) {
  std::vector<std::string> r;
  std::istringstream ss(s);
  std::string e;
  while (std::getline(ss, e, d))
    if (!e.empty()) r.push_back(e);
  return r;
}
// utils.cpp:37-50
std::string cli_arg(int argc, char** argv, const std::string& s, const std::string& l) {
  std::string r;
  for (int i = 1; i < argc; i += 2) {
    check(argv[i][0] == '-' && i + 1 < argc, "Arguments must be given as flag followed by non-flag");
    if (s == argv[i] || l == argv[i]) {
      check(r.empty(), "Command line flags cannot be used more than once");
      r = argv[i + 1];
    }
  }
  return r;
}

// Grid::format (grid.cpp:412-427): precision -1 gives digits10 + 1 significant digits
std::string format_grid(const std::vector<double>& g, int nx, int ny, int precision) {
  std::ostringstream out;
  if (precision == -1) out.precision(std::numeric_limits<double>::digits10 + 1);
  else out.precision(precision);
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) out << g[(size_t)i * ny + j] << ',';
    out.seekp(-1, std::ios_base::end);
    out << '\n';
  }
  out.seekp(-1, std::ios_base::end);
  out << '\n';
  return out.str();
}

// PlasmaDomain::num2str (plasmadomain.hpp:248-256) as called with its default argument: the second
// precision() call sets precision -1, so the stream prints with its default 6 significant digits
template <typename T>
std::string num2str(T num) {
  std::ostringstream ss;
  ss.precision(std::numeric_limits<double>::digits10 + 1);
  ss.precision(-1);
  ss << num;
  return ss.str();
}

// Variable names of each equation set in index order, and its state variables
// (idealmhd.hpp:18-29, ideal2F.hpp:23-42, idealmhd2E.hpp:18-28, idealmhdcons.hpp:18-28).
struct EqNames {
  std::vector<std::string> vars, state;
};
const std::map<std::string, EqNames>& eq_names() {
  static const std::map<std::string, EqNames> m = {
      {"ideal_mhd",
       {{"rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z", "grav_x", "grav_y", "n", "press",
         "thermal_energy", "v_x", "v_y", "v_z", "kinetic_energy", "b_x", "b_y", "b_z", "b_mag", "b_hat_x", "b_hat_y",
         "b_hat_z", "dt"},
        {"rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z", "grav_x", "grav_y"}}},
      {"ideal_2F",
       {{"i_rho", "e_rho", "i_mom_x", "i_mom_y", "e_mom_x", "e_mom_y", "i_temp", "e_temp", "bi_x", "bi_y", "bi_z",
         "E_x", "E_y", "E_z", "grav_x", "grav_y", "i_n", "e_n", "i_v_x", "i_v_y", "e_v_x", "e_v_y", "j_x", "j_y",
         "i_press", "e_press", "press", "i_thermal_energy", "e_thermal_energy", "rho", "rho_c", "n", "dn", "dt",
         "dt_i", "b_x", "b_y", "b_z", "b_mag", "b_mag_xy", "b_hat_x", "b_hat_y", "curlE_z", "divE", "divB",
         "i_dPdx", "e_dPdx"},
        {"i_rho", "e_rho", "i_mom_x", "i_mom_y", "e_mom_x", "e_mom_y", "i_temp", "e_temp", "bi_x", "bi_y", "bi_z",
         "E_x", "E_y", "E_z", "grav_x", "grav_y"}}},
      {"ideal_mhd_2E",
       {{"rho", "i_temp", "e_temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x", "grav_y", "n", "i_press", "e_press",
         "press", "i_thermal_energy", "e_thermal_energy", "v_x", "v_y", "kinetic_energy", "b_x", "b_y", "b_mag",
         "b_hat_x", "b_hat_y", "dt"},
        {"rho", "i_temp", "e_temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x", "grav_y"}}},
      {"ideal_mhd_cons",
       {{"rho", "temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x", "grav_y", "n", "v_x", "v_y", "kinetic_energy",
         "b_x", "b_y", "b_mag", "mag_energy", "b_hat_x", "b_hat_y", "thermal_energy", "press", "press_tot", "energy",
         "dt"},
        {"rho", "temp", "mom_x", "mom_y", "bi_x", "bi_y", "grav_x", "grav_y"}}}};
  return m;
}

// PlasmaDomain grids (plasmadomain.hpp:36)
const std::vector<std::string> PD_GRIDS = {"d_x", "d_y", "pos_x", "pos_y", "be_x", "be_y", "be_z"};
// module names the reference instantiates (modulehandler.hpp:34-38)
const std::vector<std::string> MODULE_NAMES = {
    "radiative_losses", "thermal_conduction", "ambient_heating", "anomalous_resistivity", "momentum_injection",
    "localized_heating", "field_heating", "tracer_particles", "sg_filtering", "coulomb_explosion",
    "eic_thermalization", "artificial_viscosity", "global_temperature", "mass_injection", "ambient_heating_sink",
    "physical_viscosity", "div_cleaning", "boundary_outflow",
    "parametric_heating"};  // this build's parametrised heating profile (no reference module)

struct Module {
  std::string name;
  std::vector<std::pair<std::string, std::string>> opts;
  // the last setting of a key wins, as the modules' parsers assign in order
  std::string opt(const std::string& k) const {
    std::string v;
    for (auto& kv : opts)
      if (kv.first == k) v = kv.second;
    return v;
  }
  bool has(const std::string& k) const {
    for (auto& kv : opts)
      if (kv.first == k) return true;
    return false;
  }
};

struct Config {
  std::string eqset;
  std::vector<std::pair<std::string, std::string>> eq_opts;
  std::vector<Module> modules;
  std::vector<std::string> output_flags;
  std::map<std::string, std::string> set;   // PlasmaDomain settings as read
};

// the lines of a {...} block after an activation line (equationset.cpp:64-85, module.cpp:11-35)
std::vector<std::pair<std::string, std::string>> read_block(std::istream& in, const std::string& what) {
  std::vector<std::pair<std::string, std::string>> kv;
  std::string line;
  std::getline(in, line);
  check(!line.empty() && line[0] == '{', what + " activation must be followed by a line starting with {");
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] == '}') return kv;        // tested before whitespace removal
    clear_ws(line);
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string l, r;
    std::getline(ss, l, '=');
    std::getline(ss, r, '#');
    kv.emplace_back(l, r);
  }
  throw Fail(what + " block is not closed with }");
}

void skip_block(std::istream& in, const std::string& what) {
  std::string line;
  std::getline(in, line);
  clear_ws(line);
  check(!line.empty() && line[0] == '{', what + " deactivation must be followed by curly brackets");
  while (std::getline(in, line)) {
    clear_ws(line);
    if (!line.empty() && line[0] == '}') return;           // equationset.cpp:49-54, modulehandler.cpp:84-89
  }
  throw Fail(what + " block is not closed with }");
}

// PlasmaDomain::readConfigFile (fileio.cpp:84-122)
Config read_config(const fs::path& p) {
  static const std::vector<std::string> names = {
      "x_bound_1", "x_bound_2", "y_bound_1", "y_bound_2", "epsilon", "density_min", "temp_min",
      "thermal_energy_min", "max_iterations", "iter_output_interval", "time_output_interval", "output_flags",
      "xdim", "ydim", "open_boundary_strength", "std_out_interval", "write_interval", "open_boundary_decay_base",
      "x_origin", "y_origin", "time_integrator", "duration", "sg_opt", "write_precision", "multispecies_mode"};
  std::ifstream in(p);
  check(in.good(), "cannot open config file " + p.string());
  Config c;
  std::string line;
  while (std::getline(in, line)) {
    clear_ws(line);
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string lhs, rhs;
    std::getline(ss, lhs, '=');
    std::getline(ss, rhs, '#');
    if (eq_names().count(lhs)) {
      if (rhs == "true") {
        c.eqset = lhs;
        c.eq_opts = read_block(in, "equation set " + lhs);
      } else {
        skip_block(in, "equation set " + lhs);
      }
      continue;
    }
    if (std::find(MODULE_NAMES.begin(), MODULE_NAMES.end(), lhs) != MODULE_NAMES.end()) {
      check(!c.eqset.empty(), "<m_eqs> must be instantiated before instantiating modules");
      if (rhs == "true") c.modules.push_back(Module{lhs, read_block(in, "module " + lhs)});
      else skip_block(in, "module " + lhs);
      continue;
    }
    check(std::find(names.begin(), names.end(), lhs) != names.end(), lhs + " is not a valid config name.");
    const auto v = split(rhs, ',');
    if (lhs == "output_flags") {
      check(!c.eqset.empty(), "equation_set must be defined in config file before output_flags");
      if (v.size() == 1) c.output_flags.push_back(rhs);
      else for (auto& x : v) c.output_flags.push_back(x);
      continue;
    }
    check(v.size() <= 1, "Only output_flags support list specifier");
    if (v.size() == 1) c.set[lhs] = rhs;
  }
  check(!c.eqset.empty(), "the config file activates no equation set");
  return c;
}

struct State {
  int nx = 0, ny = 0;
  double ion_mass = 0, gamma = 0, t = 0;
  std::vector<std::string> comments;
  std::vector<std::pair<std::string, std::vector<double>>> grids;
};

// PlasmaDomain::readStateFile (fileio.cpp:14-80)
State read_state(const fs::path& p, bool continue_mode) {
  std::ifstream in(p);
  check(in.good(), "cannot open state file " + p.string());
  State s;
  std::string line, el;
  std::getline(in, line);
  while (line.empty() || line[0] == '#') {
    if (!line.empty() && line[0] == '#') s.comments.push_back(line);
    check((bool)std::getline(in, line), "state file ends before its header");
  }
  clear_ws(line);
  check(line == "xdim,ydim", "state file: expected xdim,ydim");
  cleaned_line(in, line);
  {
    std::istringstream ss(line);
    std::getline(ss, el, ',');
    s.nx = std::stoi(el);
    std::getline(ss, el, ',');
    s.ny = std::stoi(el);
  }
  cleaned_line(in, line);
  check(line == "ion_mass", "state file: expected ion_mass");
  cleaned_line(in, line);
  s.ion_mass = std::stod(line);
  cleaned_line(in, line);
  check(line == "adiabatic_index", "state file: expected adiabatic_index");
  cleaned_line(in, line);
  s.gamma = std::stod(line);
  cleaned_line(in, line);
  {
    std::istringstream ss(line);
    std::getline(ss, el, '=');
    check(el == "t", "state file: expected t=");
    std::getline(ss, el);
    s.t = continue_mode ? std::stod(el) : 0.0;
  }
  while (cleaned_line(in, line)) {
    std::vector<double> g((size_t)s.nx * s.ny, 0.0);
    for (int i = 0; i < s.nx; ++i) {
      std::string row;
      cleaned_line(in, row);
      const int c0 = row.empty() ? 0 : row[0];
      check(std::isdigit(c0) || c0 == '-' || c0 == '.' || c0 == 'n' || c0 == 'i',
            "Encountered non-numerical row in .state file sooner than expected");
      std::istringstream ss(row);
      int j = 0;
      while (std::getline(ss, el, ',')) {
        check(j < s.ny, "Row in .state file is too long (greater than ydim)");
        g[(size_t)i * s.ny + j] = std::stod(el);
        ++j;
      }
    }
    const int nxt = in.peek();
    check(!std::isdigit(nxt) && nxt != '-', "Encountered more rows in a .state file grid than expected");
    s.grids.emplace_back(line, std::move(g));
  }
  return s;
}

double num(const Config& c, const std::string& k, double dflt, bool required = false) {
  auto it = c.set.find(k);
  if (it == c.set.end()) {
    check(!required, "config setting " + k + " is required");
    return dflt;
  }
  return std::stod(it->second);
}
int bc_code(const std::string& s) {
  static const std::vector<std::string> n = {"periodic", "open", "fixed", "reflect", "open_moc", "open_ucnp"};
  auto it = std::find(n.begin(), n.end(), s);
  check(it != n.end(), "invalid boundary condition " + s);
  return (int)(it - n.begin());
}

class Runner {
 public:
  Runner(const fs::path& out, const fs::path& cfg_path, const fs::path& state_path, bool continue_mode,
         bool overwrite_init)
      : out_(out), continue_(continue_mode) {
    // PlasmaDomain::PlasmaDomain (plasmadomain.cpp:13-75): copy the config into the output directory
    // (the progress lines are the reference's, printed as its VERBOSE build prints them)
    const fs::path dst = out_ / cfg_path.filename();
    if (!(fs::exists(dst) && fs::equivalent(cfg_path, dst))) {
      std::cout << "Copying " << cfg_path.string() << " into " << out_.string() << "...\n";
      if (fs::exists(dst)) fs::remove(dst);
      fs::copy(cfg_path, dst, fs::copy_options::overwrite_existing);
    } else {
      std::cout << cfg_path.string() << " already located in output directory.\n";
    }
    std::cout << "Reading config file...\n";
    cfg_ = read_config(cfg_path);
    std::cout << "Reading state file...\n";
    st_ = read_state(state_path, continue_mode);
    std::cout << "Validating input data...\n";
    multispecies_ = cfg_.set.count("multispecies_mode") && cfg_.set.at("multispecies_mode") == "true";
    max_iterations_ = (int)num(cfg_, "max_iterations", 100);
    iter_out_ = (int)num(cfg_, "iter_output_interval", 1);
    time_out_ = num(cfg_, "time_output_interval", -1.0);
    std_out_ = (int)num(cfg_, "std_out_interval", 1);
    write_precision_ = (int)num(cfg_, "write_precision", 4);
    write_interval_ = (int)num(cfg_, "write_interval", 1);
    duration_ = num(cfg_, "duration", -1.0);
    names_ = eq_names().at(cfg_.eqset);
    for (auto& f : cfg_.output_flags)
      check(std::find(names_.vars.begin(), names_.vars.end(), f) != names_.vars.end(),
            "output flag " + f + " is not a variable of " + cfg_.eqset);

    spruce_domain_desc d{};
    d.xdim = st_.nx;
    d.ydim = st_.ny;
    const char* bcs[4] = {"x_bound_1", "x_bound_2", "y_bound_1", "y_bound_2"};
    for (int k = 0; k < 4; ++k) {
      check(cfg_.set.count(bcs[k]) > 0, std::string("config setting ") + bcs[k] + " is required");
      d.bc[k] = bc_code(cfg_.set.at(bcs[k]));
    }
    std::string integ = cfg_.set.count("time_integrator") ? cfg_.set.at("time_integrator") : "euler";
    check(integ == "euler" || integ == "rk2" || integ == "rk4", "invalid time integrator " + integ);
    d.time_integrator = integ == "euler" ? SPRUCE_EULER : (integ == "rk2" ? SPRUCE_RK2 : SPRUCE_RK4);
    d.epsilon = num(cfg_, "epsilon", 0, true);
    d.density_min = num(cfg_, "density_min", 0, true);
    d.temp_min = num(cfg_, "temp_min", 0, true);
    d.thermal_energy_min = num(cfg_, "thermal_energy_min", 0, true);
    d.open_boundary_strength = num(cfg_, "open_boundary_strength", 0.0);
    d.open_boundary_decay_base = num(cfg_, "open_boundary_decay_base", 1.0);
    d.ion_mass = st_.ion_mass;
    d.adiabatic_index = st_.gamma;
    d.time = st_.t;
    std::vector<const char*> k, v;
    for (auto& kv : cfg_.eq_opts) { k.push_back(kv.first.c_str()); v.push_back(kv.second.c_str()); }
    dom_ = spruce_gpu_create(&d, cfg_.eqset.c_str(), (int32_t)k.size(), k.data(), v.data());
    check(dom_ != nullptr, std::string("spruce_gpu_create: ") + spruce_gpu_last_error());
    for (auto& m : cfg_.modules) {
      std::vector<const char*> mk, mv;
      for (auto& kv : m.opts) { mk.push_back(kv.first.c_str()); mv.push_back(kv.second.c_str()); }
      gpu(spruce_gpu_add_module(dom_, m.name.c_str(), (int32_t)mk.size(), mk.data(), mv.data()),
          "module " + m.name);
    }
    if (multispecies_) gpu(spruce_gpu_set_option(dom_, "multispecies_mode", "true"), "multispecies_mode");
    for (auto& g : st_.grids) {
      gpu(spruce_gpu_set_grid(dom_, g.first.c_str(), g.second.data(), (int64_t)g.second.size()), "grid " + g.first);
      pd_[g.first] = g.second;
    }
    for (auto& n : PD_GRIDS) check(pd_.count(n) > 0, "All internal grid quantities for PlasmaDomain must be initialized");
    gpu(spruce_gpu_setup(dom_), "setup");
    time_ = st_.t;
    if (!continue_mode && overwrite_init) {
      std::cout << "Writing out init.state...\n";
      write_state("init");
    }
    if (write_interval_ > 0) lines_.reserve((size_t)write_interval_ * (1 + 2 * cfg_.output_flags.size()));
    if (!continue_mode && write_interval_ > 0) {
      preamble();
      store();
      flush_out();
    }
  }
  ~Runner() {
    if (dom_) spruce_gpu_destroy(dom_);
  }

  // PlasmaDomain::run (evolution.cpp:8-57); returns the process exit status
  int run(double time_duration, double cluster_time) {
    if (time_duration > 0.0) {
      duration_ = time_duration;
      max_time_ = time_ + duration_;
    } else {
      check(duration_ >= 0.0, "Duration must be specified on command line, or in state file");
      max_time_ = duration_;
    }
    std::cout << "Begining simulation...\n";
    const auto t0 = std::chrono::steady_clock::now();
    const size_t nm = cfg_.modules.size();
    std::vector<int32_t> subs(std::max<size_t>(nm, 1));
    while (time_ < max_time_ && (max_iterations_ < 0 || iter_ < max_iterations_)) {
      const int old_t = (int)(time_ / time_out_);
      double dt = 0.0;
      gpu(spruce_gpu_step(dom_, 1, &dt, subs.data()), "step");
      if (std_out_ > 0 && iter_ % std_out_ == 0) print_update(dt, subs);
      time_ += dt;
      ++iter_;
      const int new_t = (int)(time_ / time_out_);
      const bool c1 = iter_out_ > 0 && iter_ % iter_out_ == 0;
      const bool c2 = time_out_ > 0.0 && new_t > old_t;
      if (c1 || c2) {
        store();
        // cumulative multispecies heating restarts after every store (evolution.cpp:35-41)
        if (multispecies_) gpu(spruce_gpu_multispecies_reset(dom_), "multispecies reset");
      }
      if (write_interval_ > 0 && stores_ > 0 && stores_ % write_interval_ == 0) {
        flush_out();
        write_state("end");
      }
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (el > cluster_time && cluster_time > 0) break;
    }
    flush_out();
    write_state("end");
    if (time_ >= max_time_ || (max_iterations_ > 0 && iter_ >= max_iterations_)) {
      std::cerr << "Simulation successfully reached max simulation time or iterations. Printing this error to "
                   "end recursive scripts."
                << std::endl;
      return 134;
    }
    return 0;
  }

 private:
  std::vector<double> grid(const std::string& n) {
    std::vector<double> g((size_t)st_.nx * st_.ny);
    auto it = pd_.find(n);
    if (it != pd_.end() && std::find(PD_GRIDS.begin(), PD_GRIDS.end(), n) != PD_GRIDS.end()) return it->second;
    gpu(spruce_gpu_get_grid(dom_, n.c_str(), g.data(), (int64_t)g.size()), "grid " + n);
    return g;
  }
  std::string fmt(const std::string& n, int precision) { return format_grid(grid(n), st_.nx, st_.ny, precision); }

  // PlasmaDomain::outputPreamble (fileio.cpp:124-137)
  void preamble() {
    std::ofstream f(out_ / "mhd.out");
    for (auto& c : st_.comments) f << c << std::endl;
    f << "xdim,ydim" << std::endl;
    f << st_.nx << "," << st_.ny << std::endl;
    for (const char* n : {"pos_x", "pos_y", "be_x", "be_y", "be_z"}) {
      f << n << std::endl;
      f << fmt(n, 4);
    }
  }
  // PlasmaDomain::storeGrids (fileio.cpp:142-200): output-flagged variables in index order, the
  // multispecies heating grids, then every module's fileOutput grids in module order
  void store() {
    lines_.push_back("t=" + num2str(time_) + '\n');
    for (auto& v : names_.vars) {
      if (std::find(cfg_.output_flags.begin(), cfg_.output_flags.end(), v) == cfg_.output_flags.end()) continue;
      lines_.push_back(v + '\n');
      lines_.push_back(fmt(v, write_precision_));
    }
    const size_t n = (size_t)st_.nx * st_.ny;
    std::vector<double> g(n);
    if (multispecies_) {
      const char* ms[3] = {"cumulative_electron_heating", "cumulative_ion_heating", "cumulative_joule_heating"};
      for (int w = 0; w < 3; ++w) {
        gpu(spruce_gpu_multispecies(dom_, w, g.data(), (int64_t)n), ms[w]);
        lines_.push_back(std::string(ms[w]) + '\n');
        lines_.push_back(format_grid(g, st_.nx, st_.ny, write_precision_));
      }
    }
    for (int m = 0; m < (int)cfg_.modules.size(); ++m) {
      int32_t count = 0;
      gpu(spruce_gpu_module_output(dom_, m, 0, nullptr, 0, nullptr, 0, &count), "module output");
      for (int k = 0; k < count; ++k) {
        char name[128] = {0};
        gpu(spruce_gpu_module_output(dom_, m, k, name, (int32_t)sizeof name, g.data(), (int64_t)n, nullptr),
            "module output");
        lines_.push_back(std::string(name) + '\n');
        lines_.push_back(format_grid(g, st_.nx, st_.ny, write_precision_));
      }
    }
    ++stores_;
  }
  // PlasmaDomain::writeToOutFile (fileio.cpp:205-216)
  void flush_out() {
    std::ofstream f(out_ / "mhd.out", std::ofstream::app);
    for (auto& l : lines_) f << l;
    lines_.clear();
    stores_ = 0;
  }
  // PlasmaDomain::writeStateFile (fileio.cpp:222-256), precision -1
  void write_state(const std::string& stem) {
    std::ofstream f(out_ / (stem + ".state"));
    for (auto& c : st_.comments) f << c << std::endl;
    f << "xdim,ydim\n";
    f << st_.nx << "," << st_.ny << std::endl;
    f << "ion_mass\n";
    f << st_.ion_mass << std::endl;
    f << "adiabatic_index\n";
    f << st_.gamma << std::endl;
    f << "t=" << time_ << std::endl;
    for (auto& n : PD_GRIDS) {
      f << n << std::endl;
      f << fmt(n, -1);
    }
    for (auto& n : names_.state) {
      f << n << std::endl;
      f << fmt(n, -1);
    }
  }
  // PlasmaDomain::printUpdate (fileio.cpp:283-295): the time printed is that at the start of the
  // step, and each module adds its commandLineMessage unless it is empty (modulehandler.cpp:114-122)
  void print_update(double dt, const std::vector<int32_t>& subs) {
    std::cout << "Iter: " << iter_;
    if (max_iterations_ > 0) std::cout << "/" << max_iterations_;
    std::cout << "|t: " << time_;
    if (max_time_ > 0.0) std::cout << "/" << max_time_;
    std::cout << "|dt: " << dt;
    for (size_t k = 0; k < cfg_.modules.size(); ++k) {
      const std::string msg = module_message(cfg_.modules[k], subs[k], (int)k);
      if (!msg.empty()) std::cout << "|" << msg;
    }
    std::cout << std::endl;
  }
  static bool opt_flag(const Module& m, const std::string& k, bool def) {
    return m.has(k) ? m.opt(k) == "true" : def;
  }
  static double opt_num(const Module& m, const std::string& k) {
    const std::string v = m.opt(k);
    return v.empty() ? 0.0 : std::stod(v);
  }
  // commandLineMessage of the modules this build runs
  std::string module_message(const Module& m, int32_t sub, int k) const {
    const std::string na = opt_flag(m, "inactive_mode", false) ? " (Not Applied)" : "";
    if (m.name == "thermal_conduction") return "Thermal Subcycles: " + std::to_string(sub) + na;  // thermalconduction.cpp:242-245
    if (m.name == "radiative_losses") return "Radiative Subcycles: " + std::to_string(sub) + na;  // radiativelosses.cpp:169-172
    if (m.name == "ambient_heating" || m.name == "ambient_heating_sink")  // ambientheating.cpp:51-54, ambientheatingsink.cpp:44-47
      return m.opt("exp_mode") == "true" ? "Ambient Heating On (Exp. Mode)" : "Ambient Heating On";
    if (m.name == "localized_heating" || m.name == "mass_injection" || m.name == "momentum_injection") {
      // localizedheating.cpp:69-78, massinjection.cpp:55-64, momentuminjection.cpp:78-87
      std::ostringstream oss;
      oss.precision(4);
      oss << opt_num(m, "center_x") << "," << opt_num(m, "center_y");
      const char* head = m.name == "localized_heating" ? "Heating at " :
                         m.name == "mass_injection" ? "Mass injection at " : "Momentum injection at ";
      const double start = opt_num(m, "start_time"), dur = opt_num(m, "duration");
      return head + oss.str() + ((time_ < start || time_ > start + dur) ? " Off" : " On");
    }
    if (m.name == "field_heating")  // fieldheating.cpp:64-71
      return std::string("Field Heating") + (opt_num(m, "coeff") == 0.0 ? " Zero" : " On") + na;
    if (m.name == "div_cleaning") return "Div. Cleaning On";  // divcleaning.cpp:50-53
    if (m.name == "boundary_outflow") {  // boundaryoutflow.cpp:69-74
      double v[2] = {0.0, 0.0};
      int32_t cnt = 0;
      gpu(spruce_gpu_module_report(dom_, k, v, 2, &cnt), "module report");
      check(cnt == 2, "boundary_outflow reports two values");
      const std::string side = m.has("boundary") ? m.opt("boundary") : "y_bound_2";
      return side + " boundary outflow enforced (max " + std::to_string(v[0]) + " cm/s outflow) (accel. " +
             std::to_string(v[1]) + " cm/s^2)";
    }
    if (m.name == "anomalous_resistivity")  // anomalousresistivity.cpp:315-318
      return "Anomalous Resistivity Subcycles: " + std::to_string(sub);
    if (m.name == "physical_viscosity") {  // physicalviscosity.cpp:273-290
      const bool heat = opt_flag(m, "heating_on", true), force = opt_flag(m, "force_on", true);
      const bool zero = opt_num(m, "coeff") == 0.0;
      std::string msg;
      if (heat) {
        msg += std::string("Viscous Heating") + (zero ? " Zero" : " On");
        if (force) msg += ", ";
      }
      if (force) msg += std::string("Viscous Force") + (zero ? " Zero" : " On");
      if (force || heat) msg += ", " + std::to_string(sub) + " Subcycle(s)";
      return msg + na;
    }
    return "";
  }

  fs::path out_;
  bool continue_;
  Config cfg_;
  State st_;
  EqNames names_;
  spruce_gpu_domain* dom_ = nullptr;
  std::map<std::string, std::vector<double>> pd_;
  std::vector<std::string> lines_;
  int stores_ = 0, iter_ = 0;
  bool multispecies_ = false;
  double time_ = 0.0, max_time_ = -1.0, duration_ = -1.0, time_out_ = -1.0;
  int max_iterations_ = 100, iter_out_ = 1, std_out_ = 1, write_precision_ = 4, write_interval_ = 1;
};

int run_main(int argc, char** argv) {
  const std::string mode = cli_arg(argc, argv, "-m", "--mode");
  const std::string dur_s = cli_arg(argc, argv, "-d", "--duration");
  const double duration = dur_s.empty() ? -1.0 : std::stod(dur_s);
  const std::string rt_s = cli_arg(argc, argv, "-r", "--runtime");
  const double runtime = rt_s.empty() ? -1.0 : std::stod(rt_s);
  const fs::path out(cli_arg(argc, argv, "-o", "--output"));
  check(!out.empty(), "output directory must be specified");
  if (mode == "continue") {                                         // main.cpp:29-39, mhd.cpp:7-25
    check(!dur_s.empty(), "In Continue Mode, duration of simulation must be specified on command line");
    check(fs::is_directory(out), "Given output directory of previous run must be existing directory");
    fs::path cfg, st;
    for (auto const& e : fs::directory_iterator{out}) {
      if (e.path().extension() == ".config") {
        check(cfg.empty(), "There must be only one .config file in the specified directory");
        cfg = e.path();
      } else if (e.path().filename() == "end.state") {
        st = e.path();
      }
    }
    check(!cfg.empty() && !st.empty(), "There must be a .state and a .config file in the specified directory");
    std::cout << "Running in Continue Mode for " << duration << " s...\n";
    Runner r(out, cfg, st, true, false);
    return r.run(duration, runtime);
  }
  check(mode == "input", "run mode must be input or continue");
  fs::path cfg(cli_arg(argc, argv, "-c", "--config"));              // main.cpp:40-90
  fs::path st(cli_arg(argc, argv, "-s", "--state"));
  const bool seek_cfg = cfg.empty(), seek_st = st.empty();
  if (seek_cfg || seek_st) {
    for (auto const& e : fs::directory_iterator{out}) {
      if (seek_cfg && e.path().extension() == ".config") {
        check(cfg.empty(), "There must be only one .config file in the specified directory");
        cfg = e.path();
        std::cout << "Found configuration file " << cfg << std::endl;
      } else if (seek_st && e.path().filename() == "init.state") {
        st = e.path();
        std::cout << "Found initializing state file " << st << std::endl;
      }
    }
  }
  check(!cfg.empty(), "Config file not specified and not found in output directory");
  check(!st.empty(), "Initializing state file not specified and not found in output directory");
  fs::create_directories(out);
  check(st.extension() == ".state" && fs::is_regular_file(st), "Given state file must exist and be a .state file");
  if (dur_s.empty())
    std::cout << "Running in Input Mode from the state file " << st.string()
              << " for duration specified in .config file." << std::endl;
  else
    std::cout << "Running in Input Mode from the state file " << st.string() << " for " << duration << " s...\n";
  Runner r(out, cfg, st, false, !seek_st);
  return r.run(duration, runtime);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run_main(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << "spruce_run: " << e.what() << std::endl;
    return 1;
  }
}
