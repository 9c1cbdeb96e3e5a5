// void finalise(global_variables&) (driver/finalise.h:24): releases the chunk
// bindings and every buffer of every tile (the handles have no destructor).
#include "finalise.h"

#include "mi355x_model.h"

void finalise(global_variables &globals) {
  mi355x::release_chunks();
  for (auto &t : globals.chunk.tiles) {
    field_type &f = t.field;
    for (auto *b : {&f.density0, &f.density1, &f.energy0, &f.energy1, &f.pressure, &f.viscosity, &f.soundspeed, &f.xvel0, &f.xvel1,
                    &f.yvel0, &f.yvel1, &f.vol_flux_x, &f.vol_flux_y, &f.mass_flux_x, &f.mass_flux_y, &f.work_array1, &f.work_array2,
                    &f.work_array3, &f.work_array4, &f.work_array5, &f.work_array6, &f.work_array7, &f.volume, &f.xarea, &f.yarea})
      b->release();
    for (auto *b : {&f.cellx, &f.celly, &f.vertexx, &f.vertexy, &f.celldx, &f.celldy, &f.vertexdx, &f.vertexdy}) b->release();
  }
}
