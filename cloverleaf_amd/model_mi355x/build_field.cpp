// void build_field(global_variables&)   driver/build_field.h:24
#include "build_field.h"

#include "mi355x_model.h"

void build_field(global_variables &globals) {
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) mi355x::check(clv_build_field(mi355x::chunk(globals, tile)), "build_field");
}
