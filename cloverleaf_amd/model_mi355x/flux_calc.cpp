// void flux_calc(global_variables&)   driver/flux_calc.h:24
#include "flux_calc.h"

#include "mi355x_model.h"
#include "timer.h"

void flux_calc(global_variables &globals) {
  if (mi355x::deferred_flux_calc(globals)) return;  // the recorded Lagrangian chain with it
  double kernel_time = 0;
  if (globals.profiler_on) kernel_time = timer();
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    mi355x::check(clv_flux_calc(mi355x::chunk(globals, tile), globals.dt), "flux_calc");
    mi355x::sync_if_profiling(globals, tile);
  }
  if (globals.profiler_on) globals.profiler.flux += timer() - kernel_time;
}
