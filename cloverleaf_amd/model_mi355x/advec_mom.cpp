// void advec_mom_driver(global_variables&, int tile, int which_vel, int direction, int sweep_number)   driver/advec_mom.h:24
#include "advec_mom.h"

#include "mi355x_model.h"

void advec_mom_driver(global_variables &globals, int tile, int which_vel, int direction, int sweep_number) {
  if (mi355x::defer_advec_mom(globals, tile, which_vel, direction, sweep_number)) return;  // the recorded sweep
  mi355x::check(clv_advec_mom(mi355x::chunk(globals, tile), which_vel, direction, sweep_number), "advec_mom");
  mi355x::refresh_buffers(globals, tile);  // the velocity may now be another array
  mi355x::sync_if_profiling(globals, tile);
}
