/* spruce_gpu.h -- C ABI of the MI355X (gfx950) SPRUCE time-step library (libspruce_gpu.so).
 *
 * The reference exposes its time step as a C++ plugin API: PlasmaDomain owns the grids and calls
 * EquationSet and Module hooks in a fixed order (source/mhd/plasmadomain.hpp:18-257,
 * source/equationsets/equationset.hpp:18-171, source/modules/module.hpp:15-58).  This header is the
 * drop-in boundary a host binding (ctypes, cgo, JNI, ...) would call instead: plain pointers,
 * sizes, strings and status codes; no C++ or torch types.
 *
 * Conventions
 *   - Every function returns SPRUCE_OK (0) or a negative status; spruce_gpu_last_error() returns a
 *     message for the calling thread's most recent failure.  Conditions the reference asserts on
 *     (unknown names, bad shapes, missing grids) become status codes; nothing aborts.
 *   - Host arrays are borrowed for the duration of the call.  Device memory is owned by the handle.
 *   - One handle is used by one host thread at a time.  Grids are xdim*ydim doubles, row-major with
 *     x as the row index and y contiguous (Grid, source/mhd/grid.cpp:516-526).
 *   - All arithmetic is IEEE FP64 in the reference's operation order.
 */
#ifndef SPRUCE_GPU_H
#define SPRUCE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum spruce_status {
  SPRUCE_OK = 0,
  SPRUCE_E_ARG = -1,      /* invalid argument, unknown name, inconsistent configuration */
  SPRUCE_E_STATE = -2,    /* call made in the wrong phase (e.g. step before setup) */
  SPRUCE_E_HIP = -3,      /* HIP runtime error */
  SPRUCE_E_NUMERIC = -4,  /* a reduction found no finite value (reference: grid.cpp:80 assert) */
  SPRUCE_E_UNSUPPORTED = -5
};

/* Boundary kinds in the reference's order (plasmadomain.hpp:22). */
enum spruce_bc { SPRUCE_BC_PERIODIC = 0, SPRUCE_BC_OPEN = 1, SPRUCE_BC_FIXED = 2, SPRUCE_BC_REFLECT = 3,
                 SPRUCE_BC_OPEN_MOC = 4, SPRUCE_BC_OPEN_UCNP = 5 };
/* Time integrators (plasmadomain.hpp:29); the reference default is Euler (plasmadomain.hpp:95). */
enum spruce_integrator { SPRUCE_EULER = 0, SPRUCE_RK2 = 1, SPRUCE_RK4 = 2 };

/* PlasmaDomain settings (plasmadomain.hpp:93-153) plus the two .state header scalars. */
typedef struct spruce_domain_desc {
  int32_t xdim, ydim;
  int32_t bc[4];                  /* x_bound_1, x_bound_2, y_bound_1, y_bound_2 */
  int32_t time_integrator;
  double epsilon;                 /* CFL safety factor (evolution.cpp:62) */
  double density_min;             /* number-density floor, cm^-3 */
  double temp_min;
  double thermal_energy_min;
  double open_boundary_strength;
  double open_boundary_decay_base;
  double ion_mass;                /* g */
  double adiabatic_index;
  double time;                    /* initial simulation time */
} spruce_domain_desc;

typedef struct spruce_gpu_domain spruce_gpu_domain;

/* Replaces PlasmaDomain construction + EquationSet::instantiateWithConfig (fileio.cpp:84-122,
 * equationset.cpp:24-85): eqset is one of "ideal_mhd", "ideal_mhd_cons", "ideal_mhd_2E",
 * "ideal_2F"; keys/vals are the lines of its .config block. Returns NULL on failure. */
spruce_gpu_domain* spruce_gpu_create(const spruce_domain_desc* desc, const char* eqset, int32_t nopts,
                                     const char* const* keys, const char* const* vals);

/* Replaces ModuleHandler::instantiateModule + Module::configureModule (modulehandler.cpp:77-112,
 * module.cpp:11-35).  Modules run in the order they are added (modulehandler.cpp:27-65). */
int spruce_gpu_add_module(spruce_gpu_domain* d, const char* name, int32_t nopts, const char* const* keys,
                          const char* const* vals);

/* Replaces the grid assignment in readStateFile (fileio.cpp:53-78): a PlasmaDomain grid (d_x, d_y,
 * pos_x, pos_y, be_x, be_y, be_z) or a state variable of the equation set.  count = xdim*ydim. */
int spruce_gpu_set_grid(spruce_gpu_domain* d, const char* name, const double* host, int64_t count);

/* Replaces the tail of the PlasmaDomain constructor (plasmadomain.cpp:48-53): validation,
 * computeIterationBounds, EquationSet::setupEquationSet (populateVariablesFromState) and
 * ModuleHandler::setupModules. */
int spruce_gpu_setup(spruce_gpu_domain* d);

/* Replaces nsteps calls of PlasmaDomain::advanceTime (evolution.cpp:59-82).  dt_out (may be NULL)
 * receives the nsteps step sizes; subcycles_out (may be NULL) receives nsteps*nmodules subcycle
 * counts, -1 for modules that do not subcycle. */
int spruce_gpu_step(spruce_gpu_domain* d, int32_t nsteps, double* dt_out, int32_t* subcycles_out);

/* Copies any equation-set variable (state, evolved or derived, including "dt") or PlasmaDomain grid
 * of the live state to host memory.  count = xdim*ydim. */
int spruce_gpu_get_grid(spruce_gpu_domain* d, const char* name, double* host, int64_t count);

/* Simulation time and iteration counter (plasmadomain.hpp:96-98). */
int spruce_gpu_time(spruce_gpu_domain* d, double* t, int32_t* iter);

/* The device-side values that Module::commandLineMessage (module.hpp:45; boundaryoutflow.cpp:69-74)
 * prints for module `module` (index in the order modules were added) after the latest step:
 * boundary_outflow reports {largest outflow speed, applied acceleration}; other modules report
 * nothing.  count_out (may be NULL) receives the number of values; out (may be NULL) must hold
 * count >= that many doubles. */
int spruce_gpu_module_report(spruce_gpu_domain* d, int32_t module, double* out, int32_t count, int32_t* count_out);

/* The grids module `module` adds to every stored frame (Module::fileOutput, called by
 * PlasmaDomain::storeGrids, fileio.cpp:186-194), as of the last step: thermal_conduction,
 * radiative_losses, physical_viscosity, field_heating and anomalous_resistivity with
 * output_to_file = true, artificial_viscosity with visc_output_*.  count_out (may be NULL)
 * receives how many; grid k's name goes to name (name_len bytes, NUL-terminated; may be NULL) and
 * its xdim*ydim values to host (count = xdim*ydim; may be NULL). */
int spruce_gpu_module_output(spruce_gpu_domain* d, int32_t module, int32_t k, char* name, int32_t name_len,
                             double* host, int64_t count, int32_t* count_out);

/* Multispecies mode (set_option "multispecies_mode" = "true" before setup; PlasmaDomain,
 * plasmadomain.cpp:55-58): the cumulative electron (which = 0), ion (1) and Joule (2) heating
 * grids that storeGrids writes after the variables (fileio.cpp:164-184), and their reset to zero
 * after each store (evolution.cpp:35-41). */
int spruce_gpu_multispecies(spruce_gpu_domain* d, int32_t which, double* host, int64_t count);
int spruce_gpu_multispecies_reset(spruce_gpu_domain* d);

/* Unit entry points for parity tests.
 * rhs: EquationSet::computeTimeDerivatives on the live grids (equationset.cpp:204-210), written as
 *      nevolved consecutive xdim*ydim blocks in evolved-variable order.
 * dt_min: epsilon * min of the dt grid over the dt bounds, the step size advanceTime would take
 *      next (evolution.cpp:62; grid.cpp:71-82).
 * num_evolved: number of evolved variables of the equation set. */
int spruce_gpu_rhs(spruce_gpu_domain* d, double* host_out, int64_t count);
int spruce_gpu_dt_min(spruce_gpu_domain* d, double* step_out);
int spruce_gpu_num_evolved(spruce_gpu_domain* d);

/* put_grid: overwrites one evolved variable of the live state (rho, mom_x/y/z, thermal_energy,
 *      bi_x/y/z) from host memory without propagating, as a plugin assigning m_eqs->grid(v)
 *      does (e.g. thermalconduction.cpp:153); derived variables keep their values until the next
 *      propagate.  count = xdim*ydim.  ideal_mhd only.
 * ghost_fill: PlasmaDomain::updateGhostZones on the live grids (evolution.cpp:126-152): the ghost
 *      cells' evolved variables for every non-periodic side.  ideal_mhd only.
 * propagate: EquationSet::propagateChanges on the live grids (equationset.cpp:212-220): minimums,
 *      ghost fill, derived variables and the dt grid; the next step's size epsilon * min dt
 *      (evolution.cpp:62) becomes the domain's and is written to step_out (may be NULL).
 *      ideal_mhd only. */
int spruce_gpu_put_grid(spruce_gpu_domain* d, const char* name, const double* host, int64_t count);

/* Allocation-free variant of spruce_gpu_rhs; it has no counterpart in the reference.  It is for
 * callers that cannot allocate device memory while they run (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the eight derivative grids are staged in a
 * device array the library holds for the life of the process, one per GPU, instead of in memory
 * allocated per call.  Limit: grids of at most 65,536 cells (the array is 4 MiB); larger grids
 * return SPRUCE_E_ARG.  Every call on one GPU uses the same array, so calls that overlap on one GPU
 * (from two host threads, or two handles on two streams) race.  ideal_mhd only. */
int spruce_gpu_rhs_noalloc(spruce_gpu_domain* d, double* host_out, int64_t count);
int spruce_gpu_ghost_fill(spruce_gpu_domain* d);
int spruce_gpu_propagate(spruce_gpu_domain* d, double* step_out);

/* Device-resident access for zero-copy callers: the HIP stream the handle launches on and the
 * device address of a live grid (valid until the next step call). */
int spruce_gpu_stream(spruce_gpu_domain* d, void** hip_stream);
int spruce_gpu_device_grid(spruce_gpu_domain* d, const char* name, double** dev_ptr);

/* Execution options (no reference counterpart; step results are identical for every setting):
 *   "stage_kernel" = "auto" (default: strips when the grid gives at least 1024 4-row strip
 *                    wavefronts, else one thread per cell) | "strip" (wavefront strips marching
 *                    along x) | "cell" (one thread per cell, faces recomputed per cell);
 *   "strip_rows"   = rows per wavefront strip, a multiple of 4 (default: chosen from the grid size);
 *   "tc_kernel"    = thermal-conduction update: "auto" (default: strips when 8-row strips give at
 *                    least 4096 wavefronts, else one thread per cell) | "strip" (wavefront strips,
 *                    every face computed once, b_hat derivatives formed in the kernel) | "cell";
 *   "tc_rows"      = rows per conduction strip, 1..256 (default: chosen from the grid size);
 *   "adjoint_tc_rows" = spruce_gpu_adjoint's reverse of Euler conduction subcycles without flux
 *                    saturation: rows per wavefront strip, 1..256, 0 (default: 8 where the grid
 *                    has at least 64 columns and 16 rows), or -1 (tiles of one thread per cell);
 *                    gradients agree to rounding across settings;
 *   "adjoint_tc_records" = "auto" (default: the segment recompute keeps the conduction subcycle
 *                    results of every step of the segment where they fit in 80% of the free device
 *                    memory, and the replay of a step takes them instead of running its subcycles)
 *                    | "step" (each replay runs its subcycles again); same gradients;
 *   "multispecies_mode" = "true" | "false" (the config setting of that name, fileio.cpp:322):
 *                    accumulate the cumulative heating grids (before setup);
 *   "block"        = ideal_2F launch shape: "<columns>x<rows>" thread block of every kernel, stage
 *                    one thread per cell (default 64x4), or "strip<rows>" (rows a multiple of 4
 *                    up to 64): the stage kernel in wavefront strips, the other kernels in 64x4. */
int spruce_gpu_set_option(spruce_gpu_domain* d, const char* key, const char* value);

/* Average duration (ms) of each kernel of the most recent spruce_gpu_step call, measured with HIP
 * events on the handle's stream when profiling is enabled (spruce_gpu_set_profiling).  names is a
 * comma-separated list written to a caller buffer. */
int spruce_gpu_set_profiling(spruce_gpu_domain* d, int32_t enable);
int spruce_gpu_kernel_times(spruce_gpu_domain* d, char* names, int32_t names_len, double* ms_avg,
                            int32_t* launches, int32_t max_kernels);

/* Device arithmetic self-check (no reference counterpart): evaluates a/b, sqrt(a) and the
 * reciprocal-corrected quotient used by the stencils on n inputs, for comparison with the host. */
int spruce_gpu_arith_probe(const double* a, const double* b, int64_t n, double* div_out, double* sqrt_out,
                           double* qdiv_out);

/* Device power self-check (no reference counterpart): out[i] = x[i]^y for the exponents the solar
 * modules use (which: 0 -> 1.5, 1 -> 2.5, 2 -> -1.5, 3 -> 1/3), correctly rounded. */
int spruce_gpu_pow_probe(const double* x, int64_t n, int32_t which, double* out);

const char* spruce_gpu_last_error(void);
void spruce_gpu_destroy(spruce_gpu_domain* d);

/* Discrete adjoint (no reference counterpart: the reference has no gradient).  Runs nsteps steps
 * exactly as spruce_gpu_step does (the handle advances; dt_out, may be NULL, receives the step
 * sizes), keeping a checkpoint of the evolved grids every checkpoint_interval steps, then sweeps
 * backward through the hand-written reverse of every stage, minimum/derived update and step-size
 * reduction, recomputing each segment from its checkpoint.  seed = dJ/d(evolved grids after the
 * steps); grad_out = dJ/d(evolved grids before them); both 8*xdim*ydim doubles in evolved order
 * (rho, mom_x, mom_y, mom_z, thermal_energy, bi_x, bi_y, bi_z).  Covers ideal_mhd, euler, rk2 or
 * rk4, with periodic, fixed, reflect, open (only with the other axis periodic) and open_ucnp boundaries
 * and the modules thermal_conduction (any subcycle integrator, with or without flux saturation),
 * radiative_losses, artificial_viscosity (right-hand-side and hyperviscous terms), physical_viscosity,
 * field_heating, div_cleaning, boundary_outflow,
 * ambient_heating, ambient_heating_sink, localized_heating, mass_injection, momentum_injection and
 * parametric_heating (any of them in inactive_mode too);
 * SPRUCE_E_UNSUPPORTED otherwise.  With modules or rk4 the reverse needs each step's intermediate
 * states (stages, conduction subcycle results, module outputs): the recompute of a segment keeps them
 * for all its steps where they fit in 80% of the free device memory (see "adjoint_tc_records"), and
 * each step of the backward sweep is replayed once more otherwise. */
int spruce_gpu_adjoint(spruce_gpu_domain* d, int32_t nsteps, int32_t checkpoint_interval, const double* seed,
                       double* grad_out, double* dt_out);
/* Gradient of the last spruce_gpu_adjoint call with respect to module parameters (no reference
 * counterpart): the coefficients coeff_0 .. coeff_{K-1} of each parametric_heating module, in module
 * order.  count_out (may be NULL) receives the number of parameters; grad_out (may be NULL) must
 * hold count >= that many doubles. */
int spruce_gpu_adjoint_params(spruce_gpu_domain* d, double* grad_out, int32_t count, int32_t* count_out);
/* Device memory (bytes) the last spruce_gpu_adjoint call allocated beyond the domain's own grids
 * (no reference counterpart). */
int spruce_gpu_adjoint_device_bytes(spruce_gpu_domain* d, int64_t* bytes);
/* Reverse of populateVariablesFromState (equationset.cpp:96-104): maps dJ/d(evolved grids right
 * after setup) to dJ/d(.state grids rho, temp, mom_x, mom_y, mom_z, bi_x, bi_y, bi_z), 8*xdim*ydim. */
int spruce_gpu_adjoint_state(spruce_gpu_domain* d, const double* grad_evolved, double* grad_state);

/* Batched mode (no reference counterpart: the reference runs one PlasmaDomain per process,
 * main.cpp:17-94).  A batch advances several set-up ideal_2F domains of any sizes with one launch
 * per kernel of the step (evolution.cpp:59-124 for each domain), each domain keeping its own step
 * size; the members must share the time integrator.  While in a batch a domain is stepped only
 * through it (spruce_gpu_step returns SPRUCE_E_STATE); get_grid/rhs/dt_min stay available.
 * Destroy the batch before its domains. */
typedef struct spruce_gpu_batch spruce_gpu_batch;
spruce_gpu_batch* spruce_gpu_batch_create(spruce_gpu_domain* const* domains, int32_t n);
/* "graph" = "true" (default: one step is captured once into a HIP graph and replayed) | "false";
 * "block" = launch shape, as the domain option "block" ("<columns>x<rows>" or "strip<rows>");
 * "persistent" = "true": every spruce_gpu_batch_step call is one cooperative launch that runs all
 * its steps for every member, phase by phase with grid-wide barriers (strip stage; a cell shape is
 * replaced by strips of 4 rows) | "false" (default). */
int spruce_gpu_batch_set_option(spruce_gpu_batch* b, const char* key, const char* value);
/* Times trial_steps steps for each candidate block shape, keeps the fastest and restores the state
 * of every member, so tuning does not advance the simulation.  chosen receives "<cols>x<rows>". */
int spruce_gpu_batch_autotune(spruce_gpu_batch* b, int32_t trial_steps, char* chosen, int32_t chosen_len,
                              double* ms_per_step);
/* nsteps steps of every member; dt_out (may be NULL) receives n*nsteps step sizes, member-major. */
int spruce_gpu_batch_step(spruce_gpu_batch* b, int32_t nsteps, double* dt_out);
/* Per-kernel event timing of batch steps, as spruce_gpu_set_profiling / spruce_gpu_kernel_times;
 * while profiling is on, steps are launched one by one instead of replayed from the graph. */
int spruce_gpu_batch_set_profiling(spruce_gpu_batch* b, int32_t enable);
int spruce_gpu_batch_kernel_times(spruce_gpu_batch* b, char* names, int32_t names_len, double* ms_avg,
                                  int32_t* launches, int32_t max_kernels);
void spruce_gpu_batch_destroy(spruce_gpu_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* SPRUCE_GPU_H */
