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

} // namespace

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
