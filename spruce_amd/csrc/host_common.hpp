// Host-side helpers shared by the domain drivers: errors, device buffers, mesh geometry, timing.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spruce_gpu.h"
#include "mhd.hpp"

namespace spruce {

extern thread_local std::string g_err;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void require(bool c, int code, const std::string& msg) { if (!c) throw Error(code, msg); }
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(SPRUCE_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer owned by the domain.
struct DBuf {
  double* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  void swap(DBuf& o) { std::swap(p, o.p); std::swap(n, o.n); }
  void alloc(size_t count) {
    release();
    hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double)), "hipMalloc");
    n = count;
    hip_check(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(double)), "hipMemset");
    // the domain's stream does not synchronise with the null stream: finish the clear now
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
  ~DBuf() { release(); }
};

struct AxisBufs {
  DBuf d, rd, h, hsq, rhsq, fden, rfden, gup, gdn;
  std::vector<double> hd;  // host copy of the cell sizes
  Axis view(int n, int lo, int hi, int per) const {
    return Axis{d.p, rd.p, h.p, hsq.p, rhsq.p, fden.p, rfden.p, gup.p, gdn.p, n, lo, hi, per, lo, hi};
  }
  // Precomputes the geometry the stencils divide by; every reciprocal is a correctly rounded
  // IEEE division on the host.
  void build(const std::vector<double>& dcell, bool per) {
    const int n = (int)dcell.size();
    hd = dcell;
    std::vector<double> vd(n), vrd(n), vh(n), vhsq(n), vrhsq(n), vf(n, 0.0), vrf(n, 0.0), vgu(n, 0.0), vgd(n, 0.0);
    for (int i = 0; i < n; ++i) {
      vd[i] = dcell[i];
      vrd[i] = 1.0 / dcell[i];
      vh[i] = 0.5 * dcell[i];
      vhsq[i] = vh[i] * vh[i];
      vrhsq[i] = 1.0 / vhsq[i];
    }
    auto w = [&](int k) { return per ? ((k % n) + n) % n : k; };
    for (int c = 0; c < n; ++c) {
      const int cm1 = w(c - 1), cm2 = w(c - 2), cp1 = w(c + 1);
      if (cm1 >= 0 && cm1 < n) { vf[c] = vh[c] + vh[cm1]; vrf[c] = 1.0 / vf[c]; }
      if (cm2 >= 0 && cm2 < n && cm1 >= 0) vgu[c] = vh[cm2] + 2.0 * vh[cm1];
      if (cp1 >= 0 && cp1 < n) vgd[c] = vh[cp1] + 2.0 * vh[c];
    }
    auto up = [&](DBuf& b, const std::vector<double>& v) {
      b.alloc(n);
      hip_check(hipMemcpy(b.p, v.data(), n * sizeof(double), hipMemcpyHostToDevice), "upload geometry");
    };
    up(d, vd); up(rd, vrd); up(h, vh); up(hsq, vhsq); up(rhsq, vrhsq); up(fden, vf); up(rfden, vrf);
    up(gup, vgu); up(gdn, vgd);
  }
};

struct KernelClock {
  bool enabled = false;
  struct Pending { std::string name; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<double, int>> acc;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  void collect() {
    for (auto& p : pending) {
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
      auto& s = acc[p.name];
      s.first += ms;
      s.second += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  ~KernelClock() {
    for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
  // Brackets the launches of f on stream s with events when profiling is on.
  template <class F> void timed(hipStream_t s, const char* name, F&& f) {
    if (!enabled) { f(); return; }
    hipEvent_t a = get(), b = get();
    hip_check(hipEventRecord(a, s), "hipEventRecord");
    f();
    hip_check(hipEventRecord(b, s), "hipEventRecord");
    pending.push_back({name, a, b});
  }
  // Writes the per-kernel averages in the layout of spruce_gpu_kernel_times; returns the count.
  int report(char* names, int names_len, double* ms_avg, int32_t* launches, int max_kernels) const {
    std::string all;
    int k = 0;
    for (auto& kv : acc) {
      if (k < max_kernels) {
        if (ms_avg) ms_avg[k] = kv.second.first / std::max(1, kv.second.second);
        if (launches) launches[k] = kv.second.second;
        if (!all.empty()) all += ",";
        all += kv.first;
      }
      ++k;
    }
    if (names && names_len > 0) {
      all.copy(names, names_len - 1);
      names[std::min<size_t>(all.size(), names_len - 1)] = 0;
    }
    return k;
  }
};

// The PlasmaDomain role behind the C ABI: one implementation per equation-set family.
class DomainBase {
 public:
  virtual ~DomainBase() = default;
  virtual void add_module(const std::string& name, const std::vector<std::pair<std::string, std::string>>& opts) = 0;
  virtual void set_grid(const std::string& name, const double* p, int64_t count) = 0;
  virtual void setup() = 0;
  virtual void run(int nsteps, double* dt_out, int32_t* subcycles_out) = 0;
  virtual void get_grid(const std::string& name, double* out, int64_t count) = 0;
  virtual void rhs(double* out, int64_t count) = 0;
  virtual double dt_min() = 0;
  virtual double* device_grid(const std::string& name) = 0;
  virtual void set_option(const std::string& k, const std::string& v) = 0;
  virtual int num_evolved() const = 0;
  // unit entry points on the live grids: overwrite an evolved grid (no propagate), the ghost fill
  // alone, and a whole propagate returning the next step size
  virtual void put_grid(const std::string& name, const double* p, int64_t count) {
    (void)name; (void)p; (void)count;
    throw Error(SPRUCE_E_UNSUPPORTED, "put_grid is available for ideal_mhd only");
  }
  virtual void ghost_fill() { throw Error(SPRUCE_E_UNSUPPORTED, "ghost_fill is available for ideal_mhd only"); }
  virtual double propagate() { throw Error(SPRUCE_E_UNSUPPORTED, "propagate is available for ideal_mhd only"); }
  // rhs() staged through the library's static device array instead of memory allocated per call
  virtual void rhs_noalloc(double* out, int64_t count) {
    (void)out; (void)count;
    throw Error(SPRUCE_E_UNSUPPORTED, "rhs_noalloc is available for ideal_mhd only");
  }
  // discrete adjoint of nsteps steps (only some equation sets provide it)
  virtual void adjoint(int nsteps, int checkpoint, const double* seed, double* grad, double* dt_out) {
    (void)nsteps; (void)checkpoint; (void)seed; (void)grad; (void)dt_out;
    throw Error(SPRUCE_E_UNSUPPORTED, "the adjoint is available for ideal_mhd only");
  }
  virtual void adjoint_state(const double* grad_evolved, double* grad_state) {
    (void)grad_evolved; (void)grad_state;
    throw Error(SPRUCE_E_UNSUPPORTED, "the adjoint is available for ideal_mhd only");
  }
  virtual std::vector<double> param_grad() {
    throw Error(SPRUCE_E_UNSUPPORTED, "the adjoint is available for ideal_mhd only");
  }
  virtual int64_t adjoint_bytes() const { return 0; }
  // the device values module k's command-line message reports (empty when it reports none)
  virtual std::vector<double> module_report(int k) { (void)k; return {}; }
  // xdim * ydim (0 before the grids are known)
  virtual size_t num_cells() const = 0;
  // Module::fileOutput of module k (fileio.cpp:186-194): the number of grids it writes to every
  // stored frame; with idx >= 0 also that grid's name and, when host is not null, its values
  virtual int module_output(int k, int idx, std::string* name, double* host, int64_t count) {
    (void)k; (void)name; (void)host; (void)count;
    require(idx < 0, SPRUCE_E_ARG, "module output index out of range");
    return 0;
  }
  // multispecies mode (plasmadomain.cpp:55-58, evolution.cpp:35-41): the cumulative electron (0),
  // ion (1) and Joule (2) heating since the last store.  Only the ideal_mhd modules add to them;
  // the other equation sets keep them at zero.
  bool multispecies_on = false;
  virtual void multispecies(int which, double* host, int64_t count) {
    require(multispecies_on, SPRUCE_E_STATE, "multispecies_mode is off");
    require(which >= 0 && which < 3, SPRUCE_E_ARG, "multispecies grid index must be 0, 1 or 2");
    require(count == (int64_t)num_cells() && count > 0, SPRUCE_E_ARG, "count must equal xdim*ydim");
    for (int64_t c = 0; c < count; ++c) host[c] = 0.0;
  }
  virtual void multispecies_reset() { require(multispecies_on, SPRUCE_E_STATE, "multispecies_mode is off"); }

  double time = 0.0;
  int iter = 0;
  int device = 0;                 // the GPU the handle's memory and stream live on
  hipStream_t stream = nullptr;
  KernelClock clock;

  template <class F> void timed(const char* name, F&& f) { clock.timed(stream, name, f); }
};

class Domain;

}  // namespace spruce
