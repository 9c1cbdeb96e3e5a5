// void PdV(global_variables&, bool predict)   driver/PdV.h:24
// The sequence of src/serial/PdV.cpp:92-133: the kernel, the error check, and for
// the predictor ideal_gas(predict), the pressure halo and revert -- as one launch
// (clv_PdV_driver under CLV_ENTRY_FUSED) for a single tile with every face external.
#include "PdV.h"

#include "comms.h"
#include "ideal_gas.h"
#include "mi355x_model.h"
#include "report.h"
#include "revert.h"
#include "timer.h"
#include "update_halo.h"

void PdV(global_variables &globals, bool predict) {
  if (mi355x::defer_pdv(globals, predict)) return;  // runs in flux_calc's launch
  double kernel_time = 0;
  if (globals.profiler_on) kernel_time = timer();
  globals.error_condition = 0;
  // one tile with every face external: the whole predictor chain in one launch (its
  // pressure halo is the reflective one of every face; the profiler counts it as PdV)
  if (predict && mi355x::fused_entries(globals) && globals.config.tiles_per_chunk == 1) {
    int external[4];
    mi355x::external_faces(globals, 0, external);
    if (external[0] && external[1] && external[2] && external[3]) {
      mi355x::check(clv_PdV_driver(mi355x::chunk(globals, 0), 1, globals.dt), "PdV");
      mi355x::sync_if_profiling(globals, 0);
      clover_check_error(globals.error_condition);
      if (globals.profiler_on) globals.profiler.PdV += timer() - kernel_time;
      if (globals.error_condition == 1) report_error((char *)"PdV", (char *)"error in PdV");
      return;
    }
  }
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    mi355x::check(clv_PdV(mi355x::chunk(globals, tile), predict ? 1 : 0, globals.dt), "PdV");
    mi355x::sync_if_profiling(globals, tile);
  }
  clover_check_error(globals.error_condition);
  if (globals.profiler_on) globals.profiler.PdV += timer() - kernel_time;
  if (globals.error_condition == 1) report_error((char *)"PdV", (char *)"error in PdV");
  if (predict) {
    if (globals.profiler_on) kernel_time = timer();
    for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
      ideal_gas(globals, tile, true);
      mi355x::sync_if_profiling(globals, tile);
    }
    if (globals.profiler_on) globals.profiler.ideal_gas += timer() - kernel_time;
    int fields[NUM_FIELDS];
    for (int &field : fields) field = 0;
    fields[field_pressure] = 1;
    update_halo(globals, fields, 1);
    if (globals.profiler_on) kernel_time = timer();
    revert(globals);
    if (globals.profiler_on) globals.profiler.revert += timer() - kernel_time;
  }
}
