// void update_halo(global_variables&, int fields[NUM_FIELDS], int depth)   driver/update_halo.h:24
// The order of src/serial/update_halo.cpp:690-718: the tile-to-tile copies
// (driver/update_tile_halo.cpp), the chunk exchange (clover_exchange), then the
// reflective halo on each tile's external faces (one launch when all four are).
#include "update_halo.h"

#include "comms_kernel.h"
#include "mi355x_model.h"
#include "timer.h"
#include "update_tile_halo.h"

void update_halo(global_variables &globals, int fields_in[NUM_FIELDS], const int depth) {
  // pressure / viscosity of a recorded timestep chain get their halo in its launch
  int fields[NUM_FIELDS];
  for (int k = 0; k < NUM_FIELDS; ++k) fields[k] = fields_in[k];
  mi355x::defer_halo(globals, fields, depth);
  double kernel_time = 0;
  if (globals.profiler_on) kernel_time = timer();
  update_tile_halo(globals, fields, depth);
  if (globals.profiler_on) {
    globals.profiler.tile_halo_exchange += timer() - kernel_time;
    kernel_time = timer();
  }
  clover_exchange(globals, fields, depth);
  if (globals.profiler_on) {
    globals.profiler.mpi_halo_exchange += timer() - kernel_time;
    kernel_time = timer();
  }
  const auto &nb = globals.chunk.chunk_neighbours;
  if (nb[chunk_left] == external_face || nb[chunk_right] == external_face || nb[chunk_bottom] == external_face ||
      nb[chunk_top] == external_face) {
    for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
      int external[4];
      mi355x::external_faces(globals, tile, external);
      // defer_halo has taken the recorded chain's arrays out of `fields` (or run the chain)
      mi355x::check(clv_update_halo_faces(mi355x::chunk_keep(globals, tile), fields, depth, external), "update_halo");
      mi355x::sync_if_profiling(globals, tile);
    }
  }
  if (globals.profiler_on) globals.profiler.self_halo_exchange += timer() - kernel_time;
}
