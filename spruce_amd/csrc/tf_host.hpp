// Host driver of the two-fluid equation set: one domain (the PlasmaDomain role), and a batch of
// domains advanced together by the same launches, optionally replayed from a captured HIP graph.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "twofluid.hpp"

namespace spruce {

class TfBatchRun;

// Launch shape of a batch: the stage kernel in strips of R rows (R > 0) or one thread per cell in
// bx x by blocks; the other kernels always use the bx x by cell tiles.
struct TfShape {
  int bx = 64, by = 4, R = 0;
  std::string str() const { return R > 0 ? "strip" + std::to_string(R) : std::to_string(bx) + "x" + std::to_string(by); }
};
// "<columns>x<rows>" (64, 128, 192 or 256 threads) or "strip<rows>" (rows a multiple of 4, 4..64)
TfShape parse_tf_shape(const std::string& v);

class TfDomain : public DomainBase {
 public:
  TfDomain(const spruce_domain_desc& d, const std::vector<std::string>& keys, const std::vector<std::string>& vals);
  ~TfDomain() override;
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
  int num_evolved() const override { return TF_NE; }

  // state shared with batches
  spruce_domain_desc desc{};
  bool ready = false;
  TfBatchRun* batch = nullptr;  // batch this domain belongs to, if any
  bool use_sub_cycling = true, remove_curl = false;
  int nx = 0, ny = 0;
  size_t N = 0;
  TfShape shape;
  std::map<std::string, std::vector<double>> host_grids;
  AxisBufs ax, ay;
  DBuf L[TF_NE], W[TF_NE], W2[TF_NE], K[3];
  DBuf bex, bey, gx, gy, kx3, kxv, ky3, kyv, step, history, partial, itemp, etemp, scratch;
  int* counter = nullptr;
  TfDom dom{};                        // device-pointer descriptor of this domain
  std::unique_ptr<TfBatchRun> self;   // batch of one, used by run()
  void fill_descriptor();
};

// A set of two-fluid domains stepped together: every launch covers the tiles of all domains.
class TfBatchRun {
 public:
  TfBatchRun(const std::vector<TfDomain*>& members, TfShape shape);
  ~TfBatchRun();
  void rebuild(TfShape shape);    // new launch shape (autotuning)
  void check_history();           // error if a member stepped past its step-size history
  void enqueue_step(hipStream_t s);
  void run(int nsteps, bool graph, std::vector<std::vector<double>>* dts);
  double autotune(int trial_steps, std::string* chosen);
  std::vector<TfDomain*> members;
  TfBatch b{};
  int maxn = 0;
  int integrator = 1;
  bool any_ucnp = false, any_other = false;
  bool use_graph = true;              // replay a captured step (the batch API default)
  bool persistent = false;            // all steps of a call in one cooperative launch (strip shapes)
  int persistent_blocks = 0;          // co-resident one-wavefront blocks of that launch
  bool broken = false;                // a member was destroyed while in the batch
  std::vector<TfDom> uploaded;        // host copy of d_doms, to notice reallocated member buffers
  KernelClock clock;                  // per-kernel event timing (profiling runs skip the graph)
  TfDom* d_doms = nullptr;
  int* d_tile_dom = nullptr;
  int* d_strip_dom = nullptr;
  TfShape shape;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  void destroy_graph();
};

}  // namespace spruce
