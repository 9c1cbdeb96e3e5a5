// Tile-to-tile halo copies (driver/update_tile_halo_kernel.h:25-86), reached with
// tiles_per_chunk > 1 from driver/update_tile_halo.cpp:26-91.  Each call fills the
// ghosts of one tile on one face from its neighbouring tile's interior with one
// launch for every requested field (clv_update_tile_halo).  The tiles are found
// from their density0 buffers.
#include "update_tile_halo_kernel.h"

#include "mi355x_model.h"

using B2 = clover::Buffer2D<double>;

#define TILE_PARAMS(p)                                                                                                                 \
  int p##xmin, int p##xmax, int p##ymin, int p##ymax, B2 &p##density0, B2 &p##energy0, B2 &p##pressure, B2 &p##viscosity,          \
      B2 &p##soundspeed, B2 &p##density1, B2 &p##energy1, B2 &p##xvel0, B2 &p##yvel0, B2 &p##xvel1, B2 &p##yvel1, B2 &p##vol_flux_x, \
      B2 &p##vol_flux_y, B2 &p##mass_flux_x, B2 &p##mass_flux_y

namespace {
void tile_copy(global_variables &globals, B2 &density0, B2 &neighbour_density0, int face, const int fields[NUM_FIELDS], int depth,
               const char *what) {
  clv_chunk *t = mi355x::chunk_of(globals, density0.data);
  clv_chunk *n = mi355x::chunk_of(globals, neighbour_density0.data);
  mi355x::check(clv_update_tile_halo(t, n, face, fields, depth), what);
}
}  // namespace

void update_tile_halo_l_kernel(global_variables &globals, TILE_PARAMS(), TILE_PARAMS(left_), const int fields[NUM_FIELDS], int depth) {
  tile_copy(globals, density0, left_density0, CLV_FACE_LEFT, fields, depth, "update_tile_halo_l_kernel");
}
void update_tile_halo_r_kernel(global_variables &globals, TILE_PARAMS(), TILE_PARAMS(right_), const int fields[NUM_FIELDS], int depth) {
  tile_copy(globals, density0, right_density0, CLV_FACE_RIGHT, fields, depth, "update_tile_halo_r_kernel");
}
void update_tile_halo_t_kernel(global_variables &globals, TILE_PARAMS(), TILE_PARAMS(top_), const int fields[NUM_FIELDS], int depth) {
  tile_copy(globals, density0, top_density0, CLV_FACE_TOP, fields, depth, "update_tile_halo_t_kernel");
}
void update_tile_halo_b_kernel(global_variables &globals, TILE_PARAMS(), TILE_PARAMS(bottom_), const int fields[NUM_FIELDS], int depth) {
  tile_copy(globals, density0, bottom_density0, CLV_FACE_BOTTOM, fields, depth, "update_tile_halo_b_kernel");
}
