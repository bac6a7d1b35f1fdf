// Module file output and multispecies accumulation kernels (outputs.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mhd.hpp"

namespace spruce {

// out = (now - before) / dt: a module's average rate of change of a grid over the step
// (thermalconduction.cpp:102, radiativelosses.cpp:93, anomalousresistivity.cpp:167)
void launch_out_rate(const double* now, const double* before, double dt, size_t n, double* out, hipStream_t s);
// acc += f * (now - before) when scaled (thermalconduction.cpp:106-107), acc += (now - before) otherwise
// (anomalousresistivity.cpp:168-170)
void launch_ms_change(double* acc, const double* now, const double* before, double f, bool scaled, size_t n,
                      hipStream_t s);

enum MsSourceMode : int {
  MS_SRC_TEMPLATE = 0,   // acc += (f * mask) * (dt * src)      (localizedheating.cpp:63-65)
  MS_SRC_RATE = 1,       // acc += ((f * mask) * src) * dt      (ambientheating.cpp:45-47)
  MS_SRC_SINK = 2        // acc -= ((f * mask) * src) * dt      (ambientheatingsink.cpp:38-40)
};
// dt from dt_dev when it is not null
void launch_ms_source(double* acc, const double* src, double dt, const double* dt_dev, double f, int mode, Mesh mesh,
                      hipStream_t s);

}  // namespace spruce
