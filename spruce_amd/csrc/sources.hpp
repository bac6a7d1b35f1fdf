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

}  // namespace spruce
