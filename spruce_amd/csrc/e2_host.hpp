// Host drivers of ideal_mhd_2E and ideal_mhd_cons.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "e2.hpp"
#include "host_common.hpp"

namespace spruce {

class E2Domain : public DomainBase {
 public:
  static const char* const STATE[9];
  static const char* const DERIVED[14];
  E2Domain(const spruce_domain_desc& d, const std::vector<std::string>& keys, const std::vector<std::string>& vals);
  ~E2Domain() override;
  void add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>& opts) override;
  void set_grid(const std::string& name, const double* p, int64_t count) override;
  void setup() override;
  void run(int nsteps, double* dt_out, int32_t* subcycles_out) override;
  void get_grid(const std::string& name, double* out, int64_t count) override;
  void rhs(double* out, int64_t count) override;
  double dt_min() override;
  double* device_grid(const std::string& name) override;
  void set_option(const std::string& k, const std::string& v) override;
  size_t num_cells() const override { return N; }
  int num_evolved() const override { return 7; }

 private:
  struct Set {
    DBuf rho, n, mx, my, tei, tee, bx, by, rho_pre;
    E2Set view();
    void alloc(size_t n_);
    void swap(Set& o) {
      rho.swap(o.rho); n.swap(o.n); mx.swap(o.mx); my.swap(o.my); tei.swap(o.tei); tee.swap(o.tee);
      bx.swap(o.bx); by.swap(o.by); rho_pre.swap(o.rho_pre);
    }
  };
  void ghost(Set& G, const int* kinds);
  void stage(Set& S, Set& B, Set& T, double mult, int mode, int kslot);
  void finish_step(int idx);
  void sync_and_check();
  void advance(int s);

  spruce_domain_desc desc{};
  bool ready = false, any_open = false;
  int nx = 0, ny = 0, dt_blocks = 0;
  size_t N = 0;
  std::map<std::string, std::vector<double>> host_grids;
  AxisBufs ax, ay;
  Mesh mesh{};
  MhdConst kc{};
  double s1[4] = {1, 1, 1, 1}, s2[4] = {1, 1, 1, 1};
  int kinds_all[4] = {0, 0, 0, 0}, kinds_ucnp[4] = {0, 0, 0, 0};
  Set L, W, W2;
  DBuf kbuf[3];
  DBuf bex, bey, gx, gy, partial, step, history, scratch;
  int* err_flag = nullptr;
  int eic = 0;            // eic_thermalization modules, applied in order to every RHS
  double eic_nu = 0.0;    // 2 m_e / m_ion (eic_thermalization.cpp:40)
};

class ConsDomain : public DomainBase {
 public:
  static const char* const NAMES[23];
  explicit ConsDomain(const spruce_domain_desc& d);
  ~ConsDomain() override;
  void add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>& opts) override;
  void set_grid(const std::string& name, const double* p, int64_t count) override;
  void setup() override;
  void run(int nsteps, double* dt_out, int32_t* subcycles_out) override;
  void get_grid(const std::string& name, double* out, int64_t count) override;
  void rhs(double* out, int64_t count) override;
  double dt_min() override { return step_h; }
  double* device_grid(const std::string& name) override;
  void set_option(const std::string& k, const std::string& v) override;
  size_t num_cells() const override { return N; }
  int num_evolved() const override { return 6; }

 private:
  ConsGrids view();
  spruce_domain_desc desc{};
  bool ready = false;
  int nx = 0, ny = 0, dt_blocks = 0;
  size_t N = 0;
  double step_h = 0.0;
  std::map<std::string, std::vector<double>> host_grids;
  AxisBufs ax, ay;
  Mesh mesh{};
  MhdConst kc{};
  double s1[4] = {1, 1, 1, 1}, s2[4] = {1, 1, 1, 1};
  DBuf g[24];   // NAMES order, then rho_pre
  DBuf bex, bey, step;
  int* err_flag = nullptr;
};

}  // namespace spruce
