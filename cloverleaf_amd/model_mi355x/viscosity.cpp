// void viscosity(global_variables&)   driver/viscosity.h:24
#include "viscosity.h"

#include "mi355x_model.h"

void viscosity(global_variables &globals) {
  if (mi355x::defer_viscosity(globals)) return;  // runs in calc_dt's launch
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    mi355x::check(clv_viscosity(mi355x::chunk(globals, tile)), "viscosity");
    mi355x::sync_if_profiling(globals, tile);
  }
}
