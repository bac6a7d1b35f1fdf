/* CPU FP64 restatement of the SPRUCE time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the GPU path: only tests/, __graft_entry__.smoke() and the CPU
 * leg of bench.py may load it.  The product (spruce_amd/ and include/spruce_gpu.h) never links,
 * calls or falls back to it.
 *
 * It restates the reference algorithm with the reference's evaluation order, rounding points and
 * min/max NaN rules, so that its outputs are bit-identical to the compiled reference on the
 * golden vectors in tests/golden/.  Each function cites the reference file:line it follows.
 *
 * Handles are opaque; every call returns 0 on success and a negative code on failure, with
 * oracle_last_error() describing the failure.
 */
#ifndef SPRUCE_ORACLE_H
#define SPRUCE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* scal[] layout: epsilon, density_min, temp_min, thermal_energy_min, open_boundary_strength,
 *                open_boundary_decay_base, ion_mass, adiabatic_index */
void* oracle_create(int xdim, int ydim, const int bc[4], int integrator, const double scal[8],
                    const char* eqset, int nopt, const char* const* keys, const char* const* vals);
int oracle_add_module(void* h, const char* name, int nopt, const char* const* keys, const char* const* vals);
int oracle_set_grid(void* h, const char* name, const double* data);
int oracle_setup(void* h);
int oracle_step(void* h, int nsteps, double* dt_out);
int oracle_get_grid(void* h, const char* name, double* out);
/* Time derivatives of the evolved variables of the live grids (EquationSet::computeTimeDerivatives),
 * written as nevolved consecutive xdim*ydim blocks in evolved-variable order. */
int oracle_rhs(void* h, double* out);
int oracle_num_evolved(void* h);
double oracle_time(void* h);
/* Subcycle counts of the most recent step, one per module (-1 for modules without subcycles). */
int oracle_subcycles(void* h, int* out, int nmax);
/* pow used by the module restatements: mode 0 = the C library's pow (as the reference), mode 1 =
 * correctly rounded (quad-precision powq rounded once), as the GPU computes it. */
int oracle_set_pow_mode(int mode);
double oracle_pow(double x, double y, int mode);
const char* oracle_last_error(void);
void oracle_destroy(void* h);

#ifdef __cplusplus
}
#endif
#endif
