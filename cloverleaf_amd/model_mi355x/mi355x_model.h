// Shared helpers of the MI355X plugin model: the tile -> clover_mi355x chunk
// binding and the profiler synchronisation the reference's GPU models do.
#pragma once

#include "clover_mi355x.h"
#include "definitions.h"

namespace mi355x {

// The chunk over tile `tile`'s buffers, wrapped on first use (clv_chunk_wrap), after
// running a recorded timestep chain (below).
clv_chunk *chunk(global_variables &globals, int tile);
// The same without running the recorded chain: for work that touches none of the
// chain's arrays (pressure, soundspeed, viscosity).
clv_chunk *chunk_keep(global_variables &globals, int tile);
// The chunk of the tile one of whose 2D buffers is `field` (aborts if none is).
clv_chunk *chunk_of(global_variables &globals, const double *field);
// external[face] (face: left, right, bottom, top) = the tile's face lies on the
// chunk's face and the chunk has no neighbour there -- the faces the reflective
// halo updates (src/serial/update_halo.cpp:41-71).  Taken from the tile's extent:
// the driver's tile_neighbours array keeps bottom and top in one slot
// (driver/definitions.h:58).
void external_faces(const global_variables &globals, int tile, int external[4]);
// When the reference profiler is on, wait for the tile's queued kernels so the
// driver's host timers measure them (as src/hip/update_halo.cpp:635-661 does).
void sync_if_profiling(global_variables &globals, int tile);
// Destroys the chunk bindings (finalise).
void release_chunks();
// Whether the tiles run advec_cell, advec_mom and the PdV predictor chain as one launch
// each (CLV_ENTRY_FUSED): yes unless the driver dumps fields (--dump writes the work
// arrays after every phase, which the fused kernels leave as scratch) or
// CLOVER_MI355X_ENTRY=reference.
bool fused_entries(const global_variables &globals);
// After a call that may exchange a tile's buffers (CLV_ENTRY_FUSED advection): point the
// driver's Buffer2D handles at the arrays the chunk now uses.
void refresh_buffers(global_variables &globals, int tile);

// Three chains of entry points run as one launch each on a single tile of a single chunk
// (without the profiler, with the one-launch entry points; CLOVER_MI355X_CHAINS=off
// turns them off):
// - the timestep, driver/timestep.cpp:47-69: ideal_gas(predict = false), the halo
//   update with pressure, viscosity, the viscosity halo update, then calc_dt runs
//   them (clv_timestep_kernels); the halo updates run at once for the other fields;
// - the Lagrangian phase, driver/hydro.cpp:63-70: PdV(predict), accelerate,
//   PdV(correct), then flux_calc runs them (clv_lagrange_kernels);
// - each advection sweep, driver/advection.cpp:47-73 and :75-103: advec_cell, the
//   halo update of density1, energy1, the velocities and the mass fluxes, advec_mom
//   for xvel1, then advec_mom for yvel1 runs them (clv_advection_pass); the
//   velocities' halo runs at once (advec_cell does not change them).  The second
//   sweep stays recorded until the next call: reset_field runs it with reset_field's
//   copy done by the pass's own stores (clv_advection_pass_reset), anything else
//   runs it as clv_advection_pass first.
// The calls before the last are recorded.  Anything else that touches a tile first
// runs what is recorded one entry point at a time (flush_deferred), so the arrays are
// the reference's whenever they can be seen.  The timestep chain leaves pressure and
// soundspeed unwritten when the Lagrangian chain, which rewrites them, is next; the
// first other call writes them (ideal_gas and the depth-1 pressure halo) before it runs.
bool defer_ideal_gas(global_variables &globals, int tile, bool predict);
bool defer_viscosity(global_variables &globals);
// Takes pressure / viscosity out of a halo request when their halo can be recorded.
void defer_halo(global_variables &globals, int fields[CLV_NUM_HALO_FIELDS], int depth);
// calc_dt of the recorded chain: true when it ran (outputs as clv_calc_dt).
bool deferred_calc_dt(global_variables &globals, int tile, double &dt, int &control, int &jldt, int &kldt, int &small);
bool defer_pdv(global_variables &globals, bool predict);
bool defer_accelerate(global_variables &globals);
// flux_calc of the recorded Lagrangian chain: true when it ran.
bool deferred_flux_calc(global_variables &globals);
bool defer_advec_cell(global_variables &globals, int tile, int sweep, int dir);
// advec_mom: recorded (xvel, and yvel of the second sweep) or the recorded sweep run
// (yvel of the first): true when handled.
bool defer_advec_mom(global_variables &globals, int tile, int which_vel, int dir, int sweep);
// reset_field of a recorded second sweep: true when it ran (clv_advection_pass_reset).
bool deferred_reset_field(global_variables &globals, int tile);
void flush_deferred();

inline void check(clv_status st, const char *what) { clover::check_clv(st, what); }
[[noreturn]] void unsupported(const char *what);

}  // namespace mi355x
