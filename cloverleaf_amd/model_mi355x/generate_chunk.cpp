// void generate_chunk(const int tile, global_variables&)   driver/generate_chunk.h:24
#include "generate_chunk.h"

#include "mi355x_model.h"

namespace mi355x {
clv_config config_of(const global_variables &globals);
}

void generate_chunk(const int tile, global_variables &globals) {
  clv_config c = mi355x::config_of(globals);
  mi355x::check(clv_generate_chunk(mi355x::chunk(globals, tile), &c), "generate_chunk");
}
