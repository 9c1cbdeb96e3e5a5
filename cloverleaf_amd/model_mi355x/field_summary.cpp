// void field_summary(global_variables&, parallel_&)   driver/field_summary.h:25
// The sequence of src/serial/field_summary.cpp:44-104: header, ideal_gas, the five
// sums on the GPU (CLOVER_MI355X_SUMMARY=serial selects the serial model's own
// summation order), clover_sum, report.  The pressure column is the pressure sum,
// as the serial model reports it (src/serial/field_summary.cpp:103).  Without the
// profiler, each tile's ideal_gas and sums run as one pass (clv_hydro_summary);
// with it, as two, so that the profiler's ideal_gas line keeps its meaning.
#include "field_summary.h"

#include "comms.h"
#include "ideal_gas.h"
#include "mi355x_model.h"
#include "report.h"
#include "timer.h"

void field_summary(global_variables &globals, parallel_ &parallel) {
  clover_report_step_header(globals, parallel);
  double kernel_time = 0;
  if (globals.profiler_on) {
    kernel_time = timer();
    for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) ideal_gas(globals, tile, false);
    for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) mi355x::sync_if_profiling(globals, tile);
    globals.profiler.ideal_gas += timer() - kernel_time;
    kernel_time = timer();
  }
  double vol = 0.0, mass = 0.0, ie = 0.0, ke = 0.0, press = 0.0;
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    double s[5];
    if (globals.profiler_on) mi355x::check(clv_field_summary(mi355x::chunk(globals, tile), s), "field_summary");
    else mi355x::check(clv_hydro_summary(mi355x::chunk(globals, tile), s), "field_summary");
    vol += s[0];
    mass += s[1];
    ie += s[2];
    ke += s[3];
    press += s[4];
  }
  clover_sum(vol);
  clover_sum(mass);
  clover_sum(ie);
  clover_sum(ke);
  clover_sum(press);
  if (globals.profiler_on) globals.profiler.summary += timer() - kernel_time;
  clover_report_step(globals, parallel, vol, mass, ie, ke, press);
}
