// void reset_field(global_variables&)   driver/reset_field.h:24
#include "reset_field.h"

#include "mi355x_model.h"
#include "timer.h"

void reset_field(global_variables &globals) {
  double kernel_time = 0;
  if (globals.profiler_on) kernel_time = timer();
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    if (mi355x::deferred_reset_field(globals, tile)) continue;  // done with the recorded advection sweep
    mi355x::check(clv_reset_field(mi355x::chunk(globals, tile)), "reset_field");
    mi355x::sync_if_profiling(globals, tile);
  }
  if (globals.profiler_on) globals.profiler.reset += timer() - kernel_time;
}
