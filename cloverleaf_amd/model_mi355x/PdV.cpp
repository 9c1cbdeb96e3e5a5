// void PdV(global_variables&, bool predict)   driver/PdV.h:24
// The sequence of src/serial/PdV.cpp:92-133: the kernel, the error check, and for
// the predictor ideal_gas(predict), the pressure halo and revert.
#include "PdV.h"

#include "comms.h"
#include "ideal_gas.h"
#include "mi355x_model.h"
#include "report.h"
#include "revert.h"
#include "timer.h"
#include "update_halo.h"

void PdV(global_variables &globals, bool predict) {
  double kernel_time = 0;
  if (globals.profiler_on) kernel_time = timer();
  globals.error_condition = 0;
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
