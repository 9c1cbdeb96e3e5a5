// void ideal_gas(global_variables&, const int tile, bool predict)   driver/ideal_gas.h:24
#include "ideal_gas.h"

#include "mi355x_model.h"

void ideal_gas(global_variables &globals, const int tile, bool predict) {
  if (mi355x::defer_ideal_gas(globals, tile, predict)) return;  // runs in calc_dt's launch
  mi355x::check(clv_ideal_gas(mi355x::chunk(globals, tile), predict ? 1 : 0), "ideal_gas");
}
