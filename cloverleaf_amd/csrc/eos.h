// The ideal-gas equation of state, gamma = 1.4, with the reference's expressions
// and their order (src/serial/ideal_gas.cpp:40-45): every kernel that forms a cell's
// pressure or sound speed on the fly uses these, so the doubles match what
// ideal_gas writes.
#pragma once

#include <hip/hip_runtime.h>

namespace clv {

__device__ __forceinline__ void ideal_gas_cell(double rho, double e, double &p, double &c) {
  double v = 1.0 / rho;
  p = (1.4 - 1.0) * rho * e;
  double pressurebyenergy = (1.4 - 1.0) * rho;
  double pressurebyvolume = -rho * p;
  double sound_speed_squared = v * v * (p * pressurebyenergy - pressurebyvolume);
  c = sqrt(sound_speed_squared);
}

__device__ __forceinline__ double pressure_of(double rho, double e) { return (1.4 - 1.0) * rho * e; }

}  // namespace clv
