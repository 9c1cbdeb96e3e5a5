// void revert(global_variables&)   driver/revert.h:24
#include "revert.h"

#include "mi355x_model.h"

void revert(global_variables &globals) {
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    mi355x::check(clv_revert(mi355x::chunk(globals, tile)), "revert");
    mi355x::sync_if_profiling(globals, tile);
  }
}
