// Two-fluid equation set (source/equationsets/ideal2F.cpp) on the GPU, for one domain or a batch of
// domains of different sizes advanced by the same launches.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

constexpr int TF_NE = 14;  // evolved variables, ideal2F.hpp:40-42 order
enum TfVar { TF_IRHO = 0, TF_ERHO, TF_IMX, TF_IMY, TF_EMX, TF_EMY, TF_ITE, TF_ETE, TF_EX, TF_EY, TF_EZ,
             TF_BX, TF_BY, TF_BZ };

struct TfSet {
  double* v[TF_NE];
};

struct TfConst {
  double mi, rmi;          // ion mass and RN(1/m_i)
  double me, rme;          // electron mass and RN(1/m_e)
  double gamma, gm1, rgm1;
  double nmin, temin, eps, obs;
  double rC;               // RN(1/c)
  double r6, r3;
  int curl;                // Maxwell curl terms present (remove_curl_terms == false)
  int em_dt;               // dt includes the light-crossing limit (!use_sub_cycling && !remove_curl_terms)
};

// One domain as the batched kernels see it (all pointers are device memory).
struct TfDom {
  TfSet L, W, W2;          // live and stage sets
  double* k[3];            // RK4 k1..k3 (14 fields each)
  const double *bex, *bey, *gx, *gy;
  const double* kx3;       // per row: (k_x)^2 * 3 with k_x = 2 pi / (2 d_x)   (ideal2F.cpp:175-182)
  const double* kxv;       // per row: k_x
  const double* ky3;
  const double* kyv;
  TfConst c;
  Mesh mesh;
  int bc[4];
  double s1[4], s2[4];     // open-boundary decay factors per side
  double* step;            // device step size of this domain
  double* history;         // step sizes, indexed by *counter
  int* counter;
  int hist_cap;            // doubles history holds; later step sizes are not stored
  double* partial;         // one dt minimum per tile
  int tile_base, tiles_x, tiles_y;
  int strip_base, strips_x;  // strip kernel: first strip tile of this domain, strips per column tile
};

// tile -> domain map and block shape shared by every launch of a batch
struct TfBatch {
  TfDom* doms;             // device array
  const int* tile_dom;     // device: domain of each tile
  int ndom, ntiles;
  int bx, by;              // block shape: bx columns (y) x by rows (x) per tile
  const int* strip_dom;    // device: domain of each strip tile (strip stage kernel)
  int nstrip_tiles;
  int strip_rows;          // rows per strip of the stage kernel; 0: one thread per cell
};

enum TfStage { TF_STAGE_PLAIN = 0, TF_STAGE_STORE_K = 1, TF_STAGE_RK4_FINAL = 2 };
enum TfSetSel { TF_SEL_L = 0, TF_SEL_W = 1, TF_SEL_W2 = 2 };

// T = minimums(B + mult*step * k(S)) for every domain of the batch.
void launch_tf_stage(const TfBatch& b, int S, int B, int T, double mult, int mode, int kslot, hipStream_t s);
// ghost fill of set G of every domain; kinds_mask selects which boundary kinds act (bit per kind).
void launch_tf_ghost(const TfBatch& b, int G, int kinds_mask, int maxn, hipStream_t s);
// per-tile dt minima of the live sets, then one step size per domain.
void launch_tf_dt(const TfBatch& b, hipStream_t s);
void launch_tf_populate(const TfBatch& b, const double* const* itemp, const double* const* etemp, hipStream_t s);
// variable `which` (TF_DERIVED order of tf_host.hip) of the live set of domain d, and its RHS
// (14 grids, evolved order); nx, ny are that domain's sizes.
void launch_tf_derived(const TfBatch& b, int d, int which, double* out, int nx, int ny, hipStream_t s);
void launch_tf_rhs(const TfBatch& b, int d, double* out, int nx, int ny, hipStream_t s);
// nsteps steps of the whole batch in one cooperative launch of `blocks` one-wavefront blocks (the
// strip stage, ghost fills, dt minima and step sizes, separated by grid barriers); needs a strip
// launch shape.  ucnp_mask / final_mask: the boundary-kind masks of the intermediate and final ghost
// fills (0: none).  tf_persistent_blocks(): how many such blocks the GPU holds at once.
int tf_persistent_blocks();
void launch_tf_persistent(const TfBatch& b, int nsteps, int integrator, int ucnp_mask, int final_mask, int maxn,
                          int blocks, hipStream_t s);
// history[0] = current step size and counter = 1 for every domain (start of a run of steps).
void launch_tf_reset_counters(const TfBatch& b, hipStream_t s);
// set `to` = set `from` for every domain (the Euler update, which cannot write the set it reads).
void launch_tf_copy(const TfBatch& b, int from, int to, hipStream_t s);

}  // namespace spruce
