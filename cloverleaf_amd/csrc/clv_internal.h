// Internal declarations shared by the gfx950 kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "clover_mi355x.h"

namespace clv {

constexpr double g_small = 1.0e-16;  // driver/definitions.h:32
constexpr double g_big = 1.0e+21;    // driver/definitions.h:33

// Kernel-side view of a chunk: one base pointer per array, one row pitch.
// Passed by value (kernarg), 288 bytes.
struct Dev {
  double *f[CLV_NUM_FIELDS];
  double *m[CLV_NUM_MESH];
  int nx, ny;
  int64_t pitch;
  // geom_mesh: the chunk is in CLV_GEOMETRY_MESH mode and volume, xarea and yarea
  // still hold what initialise_chunk wrote (volume = dx * dy, xarea(i, j) = celldy(j),
  // yarea(i, j) = celldx(i)), so the kernels compute them from the 1D mesh.  Set by
  // initialise_chunk in that mode, cleared by every entry point through which they
  // could change.
  int geom_mesh;
};

// Row pitch for an x_cells-wide chunk: holds nx + 5 doubles, a multiple of 16
// doubles (128 B) so that with the arena offset below every row's first
// interior element (i = 2) starts a 128-byte line.
inline int64_t pitch_for(int nx) { return ((int64_t)nx + 5 + 15) / 16 * 16; }
// doubles placed before element (0, 0) of each array so that element (2, j) is 128 B aligned
constexpr int64_t kLeadPad = 14;

// --- launch helpers ----------------------------------------------------------------------
struct Profiler;

}  // namespace clv

struct clv_chunk {
  clv::Dev d{};
  void *arena = nullptr;
  size_t arena_bytes = 0;
  bool owns = false;
  hipStream_t stream = nullptr;
  int device = 0;
  // reduction scratch: block partials (device) and results (pinned host)
  double *partials = nullptr;
  size_t partials_len = 0;
  double *result_dev = nullptr;
  double *result_host = nullptr;
  clv::Profiler *prof = nullptr;
  int summary_serial = 0;  // CLV_SUMMARY_SERIAL: reproduce the serial model's summation order
  int seg[4] = {0};        // rows per segment of the fused plan's kernel families (chosen on first use)
  int geom_mode = CLV_GEOMETRY_ARRAYS;  // CLV_GEOMETRY_MESH: initialise_chunk sets d.geom_mesh
  int entry_fused = 0;  // CLV_ENTRY_FUSED: the per-kernel entry points run one launch each (kernels_entry.hip)
};

namespace clv {

clv_status set_error(clv_status code, const std::string &msg);
clv_status hip_check(hipError_t e, const char *what);
// after a launch: reports launch errors; records profiling events when enabled
clv_status after_launch(clv_chunk *c, const char *name);
void before_launch(clv_chunk *c, const char *name);

// Kernel wrappers (defined in the .hip files), each returning CLV_OK or an error.
clv_status k_fill_zero(clv_chunk *c);
clv_status k_initialise_chunk(clv_chunk *c, const clv_config *cfg, int t_left, int t_bottom);
clv_status k_generate_chunk(clv_chunk *c, const clv_config *cfg);
clv_status k_ideal_gas(clv_chunk *c, int predict);
clv_status k_viscosity(clv_chunk *c);
clv_status k_PdV(clv_chunk *c, int predict, double dt);
clv_status k_revert(clv_chunk *c);
clv_status k_accelerate(clv_chunk *c, double dt);
clv_status k_flux_calc(clv_chunk *c, double dt);
// flux_calc's extra ghost row of vol_flux_x and ghost column of vol_flux_y only
clv_status k_flux_calc_edges(clv_chunk *c, double dt);
clv_status k_advec_cell(clv_chunk *c, int sweep_number, int dir);
clv_status k_advec_mom(clv_chunk *c, int which_vel, int direction, int sweep_number);
clv_status k_reset_field(clv_chunk *c);
clv_status k_update_halo(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth);
clv_status k_update_halo_faces(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth, const int external[4]);
clv_status k_tile_halo(clv_chunk *c, const clv_chunk *n, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth);
clv_status k_pack_array(clv_chunk *c, double *a, int field_type, int face, int depth, double *buf, int64_t offset, int unpack);
clv_status k_pack_fields(clv_chunk *c, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth, int64_t stride, int64_t tile_offset,
                         double *buf, int unpack);
clv_status k_calc_dt(clv_chunk *c, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe, double *dt_min_val);
clv_status k_field_summary(clv_chunk *c, double out[5]);
// ideal_gas(predict = false) then field_summary, one pass over the cells
clv_status k_field_summary_gas(clv_chunk *c, double out[5]);

// one launch per entry point (kernels_entry.hip); the advection kernels write the
// updated fields to the given arrays
clv_status e_PdV_predict_chain(clv_chunk *c, double dt);
clv_status e_advec_cell(clv_chunk *c, int sweep_number, int dir, double *rho_out, double *e_out);
clv_status e_advec_mom(clv_chunk *c, int which_vel, int direction, int sweep_number, double *vel_out);

// fused plan (kernels_fused.hip)
size_t fused_partials_needed(int nx, int ny);
clv_status fused_timestep_kernels(clv_chunk *c, const clv_config *cfg, double *dt_min_val);
clv_status fused_timestep_entry(clv_chunk *c, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe, int p_depth,
                                double *dt_min_val);
clv_status fused_lagrange_entry(clv_chunk *c, double dt, double *rho_pred = nullptr, double *e_pred = nullptr);
clv_status fused_advection_entry(clv_chunk *c, int sweep, int dir, bool reset);
// dst's depth-2 ghost frame (node extent) <- src's
clv_status k_copy_frame(clv_chunk *c, int src, int dst);
clv_status fused_lagrange(clv_chunk *c, double dt);
clv_status fused_advection(clv_chunk *c, int advect_x);

// Location/label bookkeeping of calc_dt: constants (src/serial/calc_dt.cpp:40,80-83).
inline void dt_location(int nx, int *control, int *jldt, int *kldt) {
  double jk_control = 1.1;
  *control = (int)(10.01 * (jk_control - (int)(jk_control)));
  jk_control = jk_control - (jk_control - (int)(jk_control));
  *jldt = ((int)jk_control) % nx;
  *kldt = (int)(1.f + (jk_control / nx));
}

// Columns each wave of a fused kernel owns (computes and stores).  A wave's 64
// lanes hold its owned columns plus the neighbours their stencils need.
constexpr int kOwnTimestep = 62, kOwnLagrange = 60, kOwnCmX = 56, kOwnCmY = 62;

// ---- batched form (kernels_fused.hip, kernels_misc.hip, batch.cpp) ------------------------
// Tile kernels of a batched step, one per fused kernel and sweep variant (the
// advection passes: cell + mom of sweep 1 / 2 in x / y).
enum BatchKind { BK_TIMESTEP, BK_LAGRANGE, BK_CM_X1, BK_CM_X2, BK_CM_Y1, BK_CM_Y2, BK_COUNT };
// One tile of a persistent launch: strip block bx of problem p, rows
// [2 + k * seg, 2 + (k + 1) * seg) clipped to the problem.
struct BatchItem {
  uint16_t p, bx, k, seg;
};
// Per problem, constant while a schedule holds.
struct BatchConst {
  double dtc_safe, dtu_safe, dtv_safe, dtdiv_safe;
  double dtrise, dtmax, dtmin, end_time;
  double *partials;  // timestep minima: strips x row segments of the timestep schedule
  double *dt_hist;   // dt of step s at [s - 1], s <= hist_len
  int nparts, end_step, hist_len, pad;
};
// Per problem, advanced on the device by batch_dt_k.
struct BatchState {
  double dt, dtold, time;
  int step;         // steps begun (driver/hydro.cpp:57)
  int complete;     // driver/hydro.cpp:108 held after the last step
  int run;          // this step's kernels run for the problem
  int ran;          // steps completed through the batch
  int failed_step;  // step whose dt fell below dtmin (0: none)
  int pad;
};
// mesh: every member's geometry comes from its mesh (Dev::geom_mesh)
hipError_t batch_launch_tiles(int kind, hipStream_t s, int grid, const Dev *tab, const BatchConst *pc, const BatchState *ps,
                              const BatchItem *items, int nitems, int mesh);
hipError_t batch_launch_dt(hipStream_t s, int nprob, const BatchConst *pc, BatchState *ps);
// halo of the listed fields for every running problem (grid: problems x fields)
hipError_t batch_launch_halo(hipStream_t s, int nprob, const Dev *tab, const BatchState *ps, const int *ids, int nids, int depth);

}  // namespace clv
