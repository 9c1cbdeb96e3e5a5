// void advec_cell_driver(global_variables&, int tile, int sweep_number, int direction)   driver/advec_cell.h:24
#include "advec_cell.h"

#include "mi355x_model.h"

void advec_cell_driver(global_variables &globals, int tile, int sweep_number, int direction) {
  if (mi355x::defer_advec_cell(globals, tile, sweep_number, direction)) return;  // runs with the sweep's advec_mom
  mi355x::check(clv_advec_cell(mi355x::chunk(globals, tile), sweep_number, direction), "advec_cell");
  mi355x::refresh_buffers(globals, tile);  // density1 / energy1 may now be other arrays
  mi355x::sync_if_profiling(globals, tile);
}
