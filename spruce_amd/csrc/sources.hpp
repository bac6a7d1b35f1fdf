// Source-term modules on the GPU: localized heating, mass and momentum injection, the ambient
// heating sink, field heating and divergence cleaning (source/modules/solar/localizedheating.cpp,
// massinjection.cpp, momentuminjection.cpp, ambientheatingsink.cpp, fieldheating.cpp,
// divcleaning.cpp).  Templates that depend on positions only are built on the host with the C
// library's exp, sin and cos; the kernels apply them with the reference's operation order.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

enum SourceMode : int {
  SRC_ADD = 0,         // dst = dst + mask * (a * t)                  (localizedheating.cpp:61)
  SRC_ADD_SCALED = 1,  // dst = dst + mask * ((a * t) * b)            (massinjection.cpp:52)
  SRC_ADD_WEIGHTED = 2,// dst = dst + mask * ((a * t) * w)            (momentuminjection.cpp:73-74)
  SRC_SUB = 3          // dst = dst - a * t                           (ambientheatingsink.cpp:36)
};

// One pointwise source update on every cell.  mask is 1 on the interior and 0 on ghost cells
// (the ghost-zone mask of plasmadomain.cpp:136-161).  With a_dev non-null the coefficient a is read
// from the device (the step size) instead of the argument.
void launch_source(double* dst, const double* t, const double* w, double a, double b, const double* a_dev, int mode,
                   Mesh mesh, hipStream_t s);

struct FheatParams {
  double coeff;   // heating coefficient
  double cc;      // C / (4 pi)
  double cpow, bpow, npow, rpow;
};

// FieldHeating::computeHeating (fieldheating.cpp:34-50) from the live grids and the live b_hat.
void launch_fheat_pre(MhdSet L, MhdStatic st, const double* bhx, const double* bhy, FheatParams p, Mesh mesh,
                      double* heat, hipStream_t s);
// FieldHeating::iterateModule (fieldheating.cpp:56-61): heat = mask * (dt * heat), and with
// apply != 0, te = te + heat.  dt is read from the device.
void launch_fheat_apply(double* te, double* heat, const double* step, Mesh mesh, int apply, hipStream_t s);

// b = bi + be on every cell (the derived b_x, b_y that divergence cleaning starts from).
void launch_field_sum(const double* bix, const double* biy, const double* bex, const double* bey, size_t n,
                      double* bx, double* by, hipStream_t s);
// One divergence-cleaning subcycle (divcleaning.cpp:34-39): the inner first derivatives
// dx_bx = derivative1D(b_x, 0) and dy_by = derivative1D(b_y, 1) (0 off the interior), then
//   bi_x += mask*h*coeff*(derivative1D(dy_by, 0) + secondDerivative1D(b_x, 0))
//   bi_y += mask*h*coeff*(derivative1D(dx_bx, 1) + secondDerivative1D(b_y, 1))
// and the next b_x = bi_x + be_x, b_y = bi_y + be_y.
void launch_dc_inner(const double* bx, const double* by, Mesh mesh, double* dx_bx, double* dy_by, hipStream_t s);
void launch_dc_update(const double* bx, const double* by, const double* dx_bx, const double* dy_by, double* bix,
                      double* biy, const double* bex, const double* bey, const double* coeff, double h, Mesh mesh,
                      double* nbx, double* nby, hipStream_t s);

// ---- anomalous resistivity (anomalousresistivity.cpp) ----
enum ArModel : int { AR_TIME_SCALE = 0, AR_SYNTELIS = 1, AR_YS94 = 2 };
struct ArParams {
  int model;
  double ts;          // time_scale (AR_TIME_SCALE)
  double p0, p1, p2;  // eta_0, eta_1, J_crit (Syntelis-19) or v_c, alpha, eta_max (YS-94)
};
struct ArRhsArgs {
  const double *bix, *biy, *biz, *coeff, *J, *lbx, *lby, *lbz;
  double *d0, *d1, *d2, *d3;
  double eta_factor;  // 4 pi / C / C
  int grad_corr;
  Mesh mesh;
};
struct ArAxpyArgs {
  const double* base[4];
  const double* k1[4];
  const double* k2[4];
  const double* k3[4];
  const double* k4[4];
  double* out[4];
  double s;
  int nf, rk4;
  size_t n;
};
void launch_ar_current(const double* bix, const double* biy, const double* biz, double cc, Mesh mesh, double* J,
                       hipStream_t s);
void launch_ar_field(const double* bex, const double* bey, const double* bix, const double* biy, size_t n, double* bx,
                     double* by, double* bm, hipStream_t s);
void launch_ar_frob(const double* bx, const double* by, const double* bm, double coeff, Mesh mesh, double* ratio,
                    double* tmpl, hipStream_t s);
// out[0] = smallest value within the radius in the box, out[1] = its cell index (-1: none),
// out[2] = the value at the centre
void launch_ar_argmin_box(const double* q, const double* px, const double* py, int ci, int cj, double radius, int il,
                          int iu, int jl, int ju, int ny, double* out, hipStream_t s);
void launch_ar_flood_seed(const double* q, size_t n, size_t start, double thr, double* res, hipStream_t s);
void launch_ar_flood_sweep(const double* q, double thr, Mesh mesh, double* res, int* changed, hipStream_t s);
void launch_ar_masks(double* tmpl, const double* J, const double* px, const double* py, size_t cnull, double max_radius,
                     double min_current, double ramp, Mesh mesh, hipStream_t s);
void launch_ar_smooth(const double* t, const double* kern, int kr, Mesh mesh, double* out, hipStream_t s);
void launch_ar_frob_final(double* t, size_t n, hipStream_t s);
void launch_ar_diffusion(const double* J, const double* n, ArParams p, Mesh mesh, double* diff, hipStream_t s);
void launch_ar_coeff(const double* tmpl, const double* diff, Mesh mesh, double* coeff, double* td, hipStream_t s);
void launch_ar_rhs(const ArRhsArgs& a, hipStream_t s);
void launch_ar_laplacian(const double* q, Mesh mesh, double* out, hipStream_t s);
void launch_ar_axpy(const ArAxpyArgs& a, hipStream_t s);
int ar_scale_blocks(const Mesh& m);
// out[0] = max over every cell of td, out[1] = min over [xl..xu] x [yl..yu] of (1/(1/dx^2+1/dy^2))/2/td
void launch_ar_scales(const double* td, Mesh mesh, int xl, int xu, int yl, int yu, double* partial, double* out,
                      hipStream_t s);

// ---- boundary outflow (boundaryoutflow.cpp) ----
enum OutflowSide : int { OF_X1 = 0, OF_X2 = 1, OF_Y1 = 2, OF_Y2 = 3 };
struct OutflowParams {
  int side;           // which boundary the outflow leaves through
  int aligned;        // field_aligned_mode: accelerate along b_hat instead of the boundary normal
  int dynamic;        // dynamic_mode: throttle towards dynamic_target_speed
  double amax;        // max_accel
  double dtime;       // dynamic_time
  double target;      // dynamic_target_speed
  int wxl, wxu, wyl, wyu;  // the cells computeMeanOutflow scans (inclusive)
};
int outflow_blocks(const OutflowParams& p);
// computeMeanOutflow (boundaryoutflow.cpp:140-236) and the acceleration of postIterateModule
// (:39-46): sc[0] = the largest outflow speed along the field over the window, starting the fold
// from -dynamic_target_speed (ties keep the first value in row order, as std::max does), and with
// set_acc != 0, sc[1] = max_accel or the throttled acceleration.  partial holds 2 * outflow_blocks.
void launch_outflow_max(MhdSet L, MhdStatic st, OutflowParams p, Mesh mesh, double* partial, double* sc, int set_acc,
                        hipStream_t s);
// postIterateModule's momentum update (:47-66): a = dt * (sc[1] * template), then
// mom_x += (a * b_hat_x) * rho and mom_y += (a * b_hat_y) * rho along the field, or the momentum
// normal to the boundary += a * rho.  dt is read from the device.
void launch_outflow_apply(MhdSet L, MhdStatic st, const double* tmpl, const double* sc, const double* step,
                          OutflowParams p, Mesh mesh, hipStream_t s);

}  // namespace spruce
