// The reference serial kernels behind one C entry point -- TEST INFRASTRUCTURE ONLY.
//
// Each call builds a reference field_type (driver/definitions.cpp:4-41, the
// reference's own extents), copies the caller's x-contiguous arrays (pitch
// nx + 5, the oracle's layout) into the serial model's y-contiguous buffers
// (src/serial/context.h:69), runs one reference kernel, and copies every array
// back.  tests/ use it to pin the oracle bitwise, kernel by kernel.
#include <array>
#include <cstring>
#include <string>

#include "definitions.h"

// Kernel signatures as defined (with external linkage) in src/serial/*.cpp.
using B1 = clover::Buffer1D<double>;
using B2 = clover::Buffer2D<double>;
void ideal_gas_kernel(int x_min, int x_max, int y_min, int y_max, B2 &density, B2 &energy, B2 &pressure, B2 &soundspeed);
void viscosity_kernel(int x_min, int x_max, int y_min, int y_max, B1 &celldx, B1 &celldy, B2 &density0, B2 &pressure, B2 &viscosity,
                      B2 &xvel0, B2 &yvel0);
void calc_dt_kernel(int x_min, int x_max, int y_min, int y_max, double dtmin, double dtc_safe, double dtu_safe, double dtv_safe,
                    double dtdiv_safe, B2 &xarea, B2 &yarea, B1 &cellx, B1 &celly, B1 &celldx, B1 &celldy, B2 &volume, B2 &density0,
                    B2 &energy0, B2 &pressure, B2 &viscosity_a, B2 &soundspeed, B2 &xvel0, B2 &yvel0, double &dt_min_val, int &dtl_control,
                    double &xl_pos, double &yl_pos, int &jldt, int &kldt, int &small);
void PdV_kernel(bool predict, int x_min, int x_max, int y_min, int y_max, double dt, B2 &xarea, B2 &yarea, B2 &volume, B2 &density0,
                B2 &density1, B2 &energy0, B2 &energy1, B2 &pressure, B2 &viscosity, B2 &xvel0, B2 &xvel1, B2 &yvel0, B2 &yvel1,
                B2 &volume_change);
void revert_kernel(int x_min, int x_max, int y_min, int y_max, B2 &density0, B2 &density1, B2 &energy0, B2 &energy1);
void accelerate_kernel(int x_min, int x_max, int y_min, int y_max, double dt, B2 &xarea, B2 &yarea, B2 &volume, B2 &density0, B2 &pressure,
                       B2 &viscosity, B2 &xvel0, B2 &yvel0, B2 &xvel1, B2 &yvel1);
void flux_calc_kernel(int x_min, int x_max, int y_min, int y_max, double dt, B2 &xarea, B2 &yarea, B2 &xvel0, B2 &yvel0, B2 &xvel1,
                      B2 &yvel1, B2 &vol_flux_x, B2 &vol_flux_y);
void advec_cell_kernel(int x_min, int x_max, int y_min, int y_max, int dir, int sweep_number, B1 &vertexdx, B1 &vertexdy, B2 &volume,
                       B2 &density1, B2 &energy1, B2 &mass_flux_x, B2 &vol_flux_x, B2 &mass_flux_y, B2 &vol_flux_y, B2 &pre_vol,
                       B2 &post_vol, B2 &pre_mass, B2 &post_mass, B2 &advec_vol, B2 &post_ener, B2 &ener_flux);
void advec_mom_kernel(int x_min, int x_max, int y_min, int y_max, B2 &vel1, B2 &mass_flux_x, B2 &vol_flux_x, B2 &mass_flux_y,
                      B2 &vol_flux_y, B2 &volume, B2 &density1, B2 &node_flux, B2 &node_mass_post, B2 &node_mass_pre, B2 &mom_flux,
                      B2 &pre_vol, B2 &post_vol, B1 &celldx, B1 &celldy, int which_vel, int sweep_number, int direction);
void reset_field_kernel(int x_min, int x_max, int y_min, int y_max, B2 &density0, B2 &density1, B2 &energy0, B2 &energy1, B2 &xvel0,
                        B2 &xvel1, B2 &yvel0, B2 &yvel1);
void update_halo_kernel(int x_min, int x_max, int y_min, int y_max, const std::array<int, 4> &chunk_neighbours,
                        const std::array<int, 4> &tile_neighbours, field_type &field, const int fields[NUM_FIELDS], int depth);
// tile-to-tile copies, exchange messages and the decompositions: declared by the
// reference's own headers
#include "comms.h"
#include "pack_kernel.h"
#include "update_tile_halo_kernel.h"

// The driver objects linked in reference the report functions of driver/report.cpp
// and g_out of driver/clover_leaf.cpp; none is reached by the kernels run here.
#include <cstdio>
#include <cstdlib>
#include <ostream>
#include "report.h"
std::ostream g_out(nullptr);
void report_error(char *location, char *error) {
  std::fprintf(stderr, "reference error in %s: %s\n", location, error);
  std::abort();
}
void clover_report_step_header(global_variables &, parallel_ &) {}
void clover_report_step(global_variables &, parallel_ &, double, double, double, double, double) {}

namespace {

// the oracle's field ids (oracle/clover_oracle.h) -> field_type members
B2 *field2d(field_type &f, int id) {
  switch (id) {
    case 0: return &f.density0;
    case 1: return &f.density1;
    case 2: return &f.energy0;
    case 3: return &f.energy1;
    case 4: return &f.pressure;
    case 5: return &f.viscosity;
    case 6: return &f.soundspeed;
    case 7: return &f.xvel0;
    case 8: return &f.xvel1;
    case 9: return &f.yvel0;
    case 10: return &f.yvel1;
    case 11: return &f.vol_flux_x;
    case 12: return &f.vol_flux_y;
    case 13: return &f.mass_flux_x;
    case 14: return &f.mass_flux_y;
    case 15: return &f.work_array1;
    case 16: return &f.work_array2;
    case 17: return &f.work_array3;
    case 18: return &f.work_array4;
    case 19: return &f.work_array5;
    case 20: return &f.work_array6;
    case 21: return &f.work_array7;
    case 22: return &f.volume;
    case 23: return &f.xarea;
    case 24: return &f.yarea;
  }
  return nullptr;
}
B1 *field1d(field_type &f, int id) {
  switch (id) {
    case 0: return &f.cellx;
    case 1: return &f.celly;
    case 2: return &f.vertexx;
    case 3: return &f.vertexy;
    case 4: return &f.celldx;
    case 5: return &f.celldy;
    case 6: return &f.vertexdx;
    case 7: return &f.vertexdy;
  }
  return nullptr;
}

// The tile-halo and pack kernels take a global_variables& that they never read.
global_variables &unread_globals() {
  alignas(global_variables) static unsigned char storage[sizeof(global_variables)];
  return *reinterpret_cast<global_variables *>(storage);
}

void load(field_type &f, int nx, double **f2d, double **m1d) {
  const size_t px = nx + 5;
  for (int a = 0; a < 25; ++a) {
    B2 &b = *field2d(f, a);
    for (size_t i = 0; i < b.sizeX; ++i)
      for (size_t j = 0; j < b.sizeY; ++j) b(i, j) = f2d[a][j * px + i];
  }
  for (int a = 0; a < 8; ++a) {
    B1 &b = *field1d(f, a);
    for (size_t i = 0; i < b.size; ++i) b[i] = m1d[a][i];
  }
}
void store(field_type &f, int nx, double **f2d) {
  const size_t px = nx + 5;
  for (int a = 0; a < 25; ++a) {
    B2 &b = *field2d(f, a);
    for (size_t i = 0; i < b.sizeX; ++i)
      for (size_t j = 0; j < b.sizeY; ++j) f2d[a][j * px + i] = b(i, j);
  }
}

} // namespace

// update_tile_halo_{l,r,b,t}_kernel (face 0 left, 1 right, 2 bottom, 3 top) on tile
// (nx, ny) from its neighbour (nnx, nny); the tile's arrays are updated in place.
extern "C" int refk_tile_halo(int face, int nx, int ny, double **f2d, double **m1d, int nnx, int nny, double **nf2d, double **nm1d,
                              const int *fields, int depth) {
  clover::context ctx{};
  field_type f(nx + 4, ny + 4, ctx), n(nnx + 4, nny + 4, ctx);
  load(f, nx, f2d, m1d);
  load(n, nnx, nf2d, nm1d);
  global_variables &g = unread_globals();
#define TILE_ARGS(t) t.density0, t.energy0, t.pressure, t.viscosity, t.soundspeed, t.density1, t.energy1, t.xvel0, t.yvel0, t.xvel1, \
                     t.yvel1, t.vol_flux_x, t.vol_flux_y, t.mass_flux_x, t.mass_flux_y
  switch (face) {
    case 0: update_tile_halo_l_kernel(g, 1, nx, 1, ny, TILE_ARGS(f), 1, nnx, 1, nny, TILE_ARGS(n), fields, depth); break;
    case 1: update_tile_halo_r_kernel(g, 1, nx, 1, ny, TILE_ARGS(f), 1, nnx, 1, nny, TILE_ARGS(n), fields, depth); break;
    case 2: update_tile_halo_b_kernel(g, 1, nx, 1, ny, TILE_ARGS(f), 1, nnx, 1, nny, TILE_ARGS(n), fields, depth); break;
    case 3: update_tile_halo_t_kernel(g, 1, nx, 1, ny, TILE_ARGS(f), 1, nnx, 1, nny, TILE_ARGS(n), fields, depth); break;
    default: return -1;
  }
#undef TILE_ARGS
  store(f, nx, f2d);
  return 0;
}

// clover_decompose (driver/comms.cpp:64-156) as seen by rank `task`: out = {left,
// right, bottom, top, chunk_neighbours[0..3]} (neighbours are 1-based ranks, -1 external)
extern "C" int refk_chunk_decompose(int x_cells, int y_cells, int chunks, int task, int *out) {
  global_config cfg{};
  cfg.number_of_chunks = chunks;
  parallel_ par;
  par.task = task;
  par.boss = false;
  int left = 0, right = 0, bottom = 0, top = 0;
  const std::array<int, 4> nb = clover_decompose(cfg, par, x_cells, y_cells, left, right, bottom, top);
  const int v[8] = {left, right, bottom, top, nb[0], nb[1], nb[2], nb[3]};
  for (int i = 0; i < 8; ++i) out[i] = v[i];
  return 0;
}

// clover_tile_decompose (driver/comms.cpp:158-253) of a chunk whose first cell is
// (left, bottom): out[12 t + ...] = {t_left, t_right, t_bottom, t_top,
// tile_neighbours[0..3], external_tile_mask[0..3]} for each of the `tiles` tiles
extern "C" int refk_tile_decompose(int x_cells, int y_cells, int tiles, int left, int bottom, int *out) {
  global_config cfg{};
  cfg.tiles_per_chunk = tiles;
  const std::array<int, 4> nb{external_face, external_face, external_face, external_face};
  chunk_type chunk(nb, 0, 1, 1, x_cells, y_cells, left, left + x_cells - 1, bottom, bottom + y_cells - 1, 0, 0, 0, 0, tiles);
  global_variables g(cfg, clover::context{}, chunk);
  const std::vector<tile_info> t = clover_tile_decompose(g, x_cells, y_cells);
  for (int i = 0; i < tiles; ++i) {
    const int v[12] = {t[i].t_left, t[i].t_right, t[i].t_bottom, t[i].t_top, t[i].tile_neighbours[0], t[i].tile_neighbours[1],
                       t[i].tile_neighbours[2], t[i].tile_neighbours[3], t[i].external_tile_mask[0], t[i].external_tile_mask[1],
                       t[i].external_tile_mask[2], t[i].external_tile_mask[3]};
    for (int k = 0; k < 12; ++k) out[12 * i + k] = v[k];
  }
  return 0;
}

// clover_pack_message_* (unpack = 0) / clover_unpack_message_* (unpack = 1) of one
// (nx + 5) x (ny + 5) array for face 0 left, 1 right, 2 bottom, 3 top; buf has buflen
// elements; both are updated in place.
extern "C" int refk_pack(int face, int unpack, int nx, int ny, double *field, int field_type_, int depth, double *buf, int buflen,
                         int offset) {
  clover::context ctx{};
  B2 a(ctx, nx + 5, ny + 5);
  const size_t px = nx + 5;
  for (size_t i = 0; i < a.sizeX; ++i)
    for (size_t j = 0; j < a.sizeY; ++j) a(i, j) = field[j * px + i];
  B1 b(ctx, buflen);
  for (int i = 0; i < buflen; ++i) b[i] = buf[i];
  global_variables &g = unread_globals();
  const int c = cell_data, v = vertex_data, xf = x_face_data, yf = y_face_data;
  if (!unpack) {
    if (face == 0) clover_pack_message_left(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else if (face == 1) clover_pack_message_right(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else if (face == 2) clover_pack_message_bottom(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else if (face == 3) clover_pack_message_top(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else return -1;
  } else {
    if (face == 0) clover_unpack_message_left(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else if (face == 1) clover_unpack_message_right(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else if (face == 2) clover_unpack_message_bottom(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else if (face == 3) clover_unpack_message_top(g, 1, nx, 1, ny, a, b, c, v, xf, yf, depth, field_type_, offset);
    else return -1;
  }
  for (size_t i = 0; i < a.sizeX; ++i)
    for (size_t j = 0; j < a.sizeY; ++j) field[j * px + i] = a(i, j);
  for (int i = 0; i < buflen; ++i) buf[i] = b[i];
  return 0;
}

extern "C" int refk_run(const char *kernel, int nx, int ny, double **f2d, double **m1d, const int *ia, const double *da, double *dout,
                        int *iout) {
  clover::context ctx{};
  field_type f(nx + 4, ny + 4, ctx);
  const size_t px = nx + 5;
  for (int a = 0; a < 25; ++a) {
    B2 &b = *field2d(f, a);
    for (size_t i = 0; i < b.sizeX; ++i)
      for (size_t j = 0; j < b.sizeY; ++j) b(i, j) = f2d[a][j * px + i];
  }
  for (int a = 0; a < 8; ++a) {
    B1 &b = *field1d(f, a);
    for (size_t i = 0; i < b.size; ++i) b[i] = m1d[a][i];
  }
  const int x_min = 1, y_min = 1, x_max = nx, y_max = ny;
  std::string k(kernel);
  if (k == "ideal_gas") {
    if (ia[0]) ideal_gas_kernel(x_min, x_max, y_min, y_max, f.density1, f.energy1, f.pressure, f.soundspeed);
    else ideal_gas_kernel(x_min, x_max, y_min, y_max, f.density0, f.energy0, f.pressure, f.soundspeed);
  } else if (k == "viscosity") {
    viscosity_kernel(x_min, x_max, y_min, y_max, f.celldx, f.celldy, f.density0, f.pressure, f.viscosity, f.xvel0, f.yvel0);
  } else if (k == "calc_dt") {
    double xl = 0, yl = 0;
    calc_dt_kernel(x_min, x_max, y_min, y_max, da[0], da[1], da[2], da[3], da[4], f.xarea, f.yarea, f.cellx, f.celly, f.celldx, f.celldy,
                   f.volume, f.density0, f.energy0, f.pressure, f.viscosity, f.soundspeed, f.xvel0, f.yvel0, dout[0], iout[0], xl, yl,
                   iout[1], iout[2], iout[3]);
  } else if (k == "PdV") {
    PdV_kernel(ia[0] != 0, x_min, x_max, y_min, y_max, da[0], f.xarea, f.yarea, f.volume, f.density0, f.density1, f.energy0, f.energy1,
               f.pressure, f.viscosity, f.xvel0, f.xvel1, f.yvel0, f.yvel1, f.work_array1);
  } else if (k == "revert") {
    revert_kernel(x_min, x_max, y_min, y_max, f.density0, f.density1, f.energy0, f.energy1);
  } else if (k == "accelerate") {
    accelerate_kernel(x_min, x_max, y_min, y_max, da[0], f.xarea, f.yarea, f.volume, f.density0, f.pressure, f.viscosity, f.xvel0,
                      f.yvel0, f.xvel1, f.yvel1);
  } else if (k == "flux_calc") {
    flux_calc_kernel(x_min, x_max, y_min, y_max, da[0], f.xarea, f.yarea, f.xvel0, f.yvel0, f.xvel1, f.yvel1, f.vol_flux_x, f.vol_flux_y);
  } else if (k == "advec_cell") {
    // ia = {sweep_number, dir}; binding as in src/serial/advec_cell.cpp:228-231
    advec_cell_kernel(x_min, x_max, y_min, y_max, ia[1], ia[0], f.vertexdx, f.vertexdy, f.volume, f.density1, f.energy1, f.mass_flux_x,
                      f.vol_flux_x, f.mass_flux_y, f.vol_flux_y, f.work_array1, f.work_array2, f.work_array3, f.work_array4,
                      f.work_array5, f.work_array6, f.work_array7);
  } else if (k == "advec_mom") {
    // ia = {which_vel, sweep_number, direction}; binding as in src/serial/advec_mom.cpp:233-243
    B2 &vel1 = ia[0] == 1 ? f.xvel1 : f.yvel1;
    advec_mom_kernel(x_min, x_max, y_min, y_max, vel1, f.mass_flux_x, f.vol_flux_x, f.mass_flux_y, f.vol_flux_y, f.volume, f.density1,
                     f.work_array1, f.work_array2, f.work_array3, f.work_array4, f.work_array5, f.work_array6, f.celldx, f.celldy, ia[0],
                     ia[1], ia[2]);
  } else if (k == "reset_field") {
    reset_field_kernel(x_min, x_max, y_min, y_max, f.density0, f.density1, f.energy0, f.energy1, f.xvel0, f.xvel1, f.yvel0, f.yvel1);
  } else if (k == "update_halo_ext") {
    // ia = {depth, mask[0..14], ext[0..3]}: a face is reflected where the chunk's
    // neighbour is external (src/serial/update_halo.cpp:41-71); every tile face external
    std::array<int, 4> chunk_nb{};
    for (int face = 0; face < 4; ++face) chunk_nb[face] = ia[16 + face] ? external_face : 1;
    std::array<int, 4> tile_nb{external_tile, external_tile, external_tile, external_tile};
    update_halo_kernel(x_min, x_max, y_min, y_max, chunk_nb, tile_nb, f, ia + 1, ia[0]);
  } else if (k == "update_halo") {
    // ia = {depth, mask[0..14]}; one chunk, one tile: every neighbour external (driver/comms.cpp:221-237)
    std::array<int, 4> chunk_nb{external_face, external_face, external_face, external_face};
    std::array<int, 4> tile_nb{external_tile, external_tile, external_tile, external_tile};
    update_halo_kernel(x_min, x_max, y_min, y_max, chunk_nb, tile_nb, f, ia + 1, ia[0]);
  } else {
    return -1;
  }
  for (int a = 0; a < 25; ++a) {
    B2 &b = *field2d(f, a);
    for (size_t i = 0; i < b.sizeX; ++i)
      for (size_t j = 0; j < b.sizeY; ++j) f2d[a][j * px + i] = b(i, j);
  }
  return 0;
}
