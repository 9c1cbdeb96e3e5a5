// void initialise_chunk(const int tile, global_variables&)   driver/initialise_chunk.h:24
#include "initialise_chunk.h"

#include <cstring>

#include "mi355x_model.h"

namespace mi355x {
// the grid and the states of the deck in the C ABI's form
clv_config config_of(const global_variables &globals) {
  clv_config c;
  std::memset(&c, 0, sizeof(c));
  const global_config &g = globals.config;
  c.xmin = g.grid.xmin;
  c.ymin = g.grid.ymin;
  c.xmax = g.grid.xmax;
  c.ymax = g.grid.ymax;
  c.x_cells = g.grid.x_cells;
  c.y_cells = g.grid.y_cells;
  c.dtinit = g.dtinit;
  c.dtmin = g.dtmin;
  c.dtmax = g.dtmax;
  c.dtrise = g.dtrise;
  c.dtc_safe = g.dtc_safe;
  c.dtu_safe = g.dtu_safe;
  c.dtv_safe = g.dtv_safe;
  c.dtdiv_safe = g.dtdiv_safe;
  c.end_time = g.end_time;
  c.end_step = g.end_step;
  c.summary_frequency = g.summary_frequency;
  c.number_of_states = g.number_of_states;
  if (g.number_of_states > CLV_MAX_STATES) unsupported("more than CLV_MAX_STATES states");
  for (int s = 0; s < g.number_of_states; ++s) {
    const state_type &st = g.states[s];
    c.states[s] = clv_state{st.density, st.energy, st.xvel, st.yvel, (int)st.geometry, st.xmin, st.ymin, st.xmax, st.ymax, st.radius};
  }
  return c;
}
}  // namespace mi355x

// src/serial/initialise_chunk.cpp:32-37: the tile's mesh starts at its first cell
// (t_left, t_bottom) of the grid
void initialise_chunk(const int tile, global_variables &globals) {
  clv_config c = mi355x::config_of(globals);
  const tile_info &t = globals.chunk.tiles[tile].info;
  mi355x::check(clv_initialise_tile(mi355x::chunk(globals, tile), &c, t.t_left, t.t_bottom), "initialise_chunk");
}
