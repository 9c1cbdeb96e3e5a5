// void accelerate(global_variables&)   driver/accelerate.h:24
#include "accelerate.h"

#include "mi355x_model.h"
#include "timer.h"

void accelerate(global_variables &globals) {
  if (mi355x::defer_accelerate(globals)) return;  // runs in flux_calc's launch
  double kernel_time = 0;
  if (globals.profiler_on) kernel_time = timer();
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    mi355x::check(clv_accelerate(mi355x::chunk(globals, tile), globals.dt), "accelerate");
    mi355x::sync_if_profiling(globals, tile);
  }
  if (globals.profiler_on) globals.profiler.acceleration += timer() - kernel_time;
}
