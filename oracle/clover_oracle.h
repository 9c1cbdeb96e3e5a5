/*
 * CPU oracle for the CloverLeaf hydro hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A C/C++ restatement (C interface) of the reference's serial model (UoB-HPC/CloverLeaf,
 * src/serial/<kernel>.cpp) and of its step sequencing (driver/hydro.cpp,
 * driver/timestep.cpp, driver/advection.cpp, driver/start.cpp).  Each function
 * cites the reference file:line it follows and keeps the reference's
 * floating-point expression order, so that results are bit-identical.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only to check the GPU library's results against or to time a
 * CPU run beside it.  The product library never calls it.
 *
 * Pinning (how each function is checked; the tests live in tests/):
 *   - every kernel: bitwise against the reference serial kernel compiled from
 *     its own sources (oracle/_ref/librefkern.so, built by oracle/Makefile
 *     where the reference tree is present) on random grids, every direction,
 *     sweep and branch;
 *   - full runs: dt sequence and field summaries against the full-precision
 *     output of the reference serial model built here (tests/golden/<deck>.json)
 *     and the reference's own known answers (driver/report.cpp:77-92);
 *   - closed forms and invariants: ideal gas c^2 = gamma p / rho; a uniform
 *     single-state deck stays at rest and its dt equals
 *     min(0.7 min(dx,dy) / sqrt(1.4 * 0.4 * e), dtold * dtrise, dtmax);
 *     volume and mass are conserved; halos reflect with the reference's signs.
 *
 * Layout: every 2D array is x-contiguous with one row pitch px = nx + 5 and
 * ny + 5 rows; element (i, j) (reference storage indices, 0 = x_min - 2) lives
 * at [j * px + i].  Cell arrays use (nx+4) x (ny+4) of that, node arrays
 * (nx+5) x (ny+5), x-face (nx+5) x (ny+4), y-face (nx+4) x (ny+5).
 */
#ifndef CLOVER_ORACLE_H
#define CLOVER_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* 2D field ids; 0..14 follow the reference enum field_parameter
 * (driver/definitions.h:62-78) so that update_halo field masks match. */
enum {
  ORC_DENSITY0 = 0, ORC_DENSITY1 = 1, ORC_ENERGY0 = 2, ORC_ENERGY1 = 3,
  ORC_PRESSURE = 4, ORC_VISCOSITY = 5, ORC_SOUNDSPEED = 6,
  ORC_XVEL0 = 7, ORC_XVEL1 = 8, ORC_YVEL0 = 9, ORC_YVEL1 = 10,
  ORC_VOL_FLUX_X = 11, ORC_VOL_FLUX_Y = 12, ORC_MASS_FLUX_X = 13, ORC_MASS_FLUX_Y = 14,
  ORC_WORK1 = 15, ORC_WORK2, ORC_WORK3, ORC_WORK4, ORC_WORK5, ORC_WORK6, ORC_WORK7,
  ORC_VOLUME = 22, ORC_XAREA = 23, ORC_YAREA = 24,
  ORC_NUM_2D = 25
};
/* 1D mesh arrays (driver/definitions.cpp:28-35). */
enum {
  ORC_CELLX = 0, ORC_CELLY, ORC_VERTEXX, ORC_VERTEXY,
  ORC_CELLDX, ORC_CELLDY, ORC_VERTEXDX, ORC_VERTEXDY,
  ORC_NUM_1D = 8
};

typedef struct orc_chunk {
  int nx, ny;             /* x_max, y_max; x_min = y_min = 1 */
  int px;                 /* row pitch of every 2D array = nx + 5 */
  double *f[ORC_NUM_2D];  /* (ny + 5) * px doubles each */
  double *m[ORC_NUM_1D];  /* x arrays nx + 5, y arrays ny + 5 doubles */
} orc_chunk;

typedef struct orc_state {
  double density, energy, xvel, yvel;
  int geometry;           /* 1 rectangle, 2 circle, 3 point (driver/definitions.h:50) */
  double xmin, ymin, xmax, ymax, radius;
} orc_state;

typedef struct orc_config {
  double xmin, ymin, xmax, ymax;
  int x_cells, y_cells;
  double dtinit, dtmax, dtmin, dtrise;
  double dtc_safe, dtu_safe, dtv_safe, dtdiv_safe;
  double end_time;
  int end_step;
  int summary_frequency;
  int number_of_states;
  const orc_state *states;  /* states[0] is the background; extents already shrunk (read_input.cpp:268-276) */
} orc_config;

/* Per-step driver state (driver/definitions.h:265-274). */
typedef struct orc_globals {
  int step;
  int advect_x;
  double time, dt, dtold;
} orc_globals;

/* allocation (zeroed) */
int orc_chunk_alloc(orc_chunk *c, int nx, int ny);
void orc_chunk_free(orc_chunk *c);

/* kernels */
void orc_ideal_gas(orc_chunk *c, int predict);
void orc_viscosity(orc_chunk *c);
void orc_calc_dt(orc_chunk *c, double dtmin, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe,
                 double *dt_min_val, int *dtl_control, int *jldt, int *kldt, int *small);
void orc_PdV(orc_chunk *c, int predict, double dt);  /* kernel only */
void orc_revert(orc_chunk *c);
void orc_accelerate(orc_chunk *c, double dt);
void orc_flux_calc(orc_chunk *c, double dt);
void orc_advec_cell(orc_chunk *c, int sweep_number, int dir);
void orc_advec_mom(orc_chunk *c, int which_vel, int sweep_number, int direction);
void orc_reset_field(orc_chunk *c);
void orc_update_halo(orc_chunk *c, const int fields[15], int depth);
/* the reflective halo on the faces with ext[face] != 0 only (face 0 left, 1 right,
 * 2 bottom, 3 top): a face shared with another tile or chunk is left to its exchange */
void orc_update_halo_ext(orc_chunk *c, const int fields[15], int depth, const int ext[4]);
/* ghost layers of c on `face` from the neighbouring tile (update_tile_halo_{l,r,b,t}_kernel) */
void orc_tile_halo(orc_chunk *c, orc_chunk *neighbour, int face, const int fields[15], int depth);
/* one field's exchange message (clover_pack_message_* / clover_unpack_message_*):
 * field is an array of c's layout, field_type 1 cell / 2 vertex / 3 x face / 4 y face */
void orc_pack_message(orc_chunk *c, double *field, int field_type, int face, int depth, double *buf, long offset);
void orc_unpack_message(orc_chunk *c, double *field, int field_type, int face, int depth, const double *buf, long offset);
void orc_field_summary_kernel(orc_chunk *c, double out[5]);
void orc_build_field(orc_chunk *c);
void orc_initialise_chunk(orc_chunk *c, const orc_config *cfg);
void orc_generate_chunk(orc_chunk *c, const orc_config *cfg);

/* sequencing */
/* driver/start.cpp:69-127: build, init, generate, ideal_gas, depth-2 halo, step-0 summary. */
void orc_start(orc_chunk *c, const orc_config *cfg, orc_globals *g, double summary0[5]);
/* driver/timestep.cpp:37-109 (dt selection; returns 1 if dt < dtmin). */
int orc_timestep(orc_chunk *c, const orc_config *cfg, orc_globals *g);
/* driver/PdV driver src/serial/PdV.cpp:92-133 (kernel + predictor chain). */
void orc_PdV_driver(orc_chunk *c, int predict, double dt);
/* driver/advection.cpp:29-105 */
void orc_advection(orc_chunk *c, int advect_x);
/* One pass of driver/hydro.cpp:57-89. Returns 1 if dt fell below dtmin. */
int orc_hydro_step(orc_chunk *c, const orc_config *cfg, orc_globals *g);
/* src/serial/field_summary.cpp:44-104: ideal_gas then the five sums. */
void orc_field_summary(orc_chunk *c, double out[5]);
/* One step with a given dt: driver/timestep.cpp:47-69 without calc_dt, then
 * driver/hydro.cpp:63-89.  Equal to orc_hydro_step whenever calc_dt would not
 * limit dt (e.g. dtc/u/v/div_safe large, dtinit = dtmax = dt). */
void orc_step_fixed(orc_chunk *c, double dt, int advect_x);

/* ---- tangent-linear model ----------------------------------------------------------------
 * The same kernels on forward-mode dual numbers: every 2D array element holds a
 * value v and a directional derivative d.  Branches follow the value; at a tie
 * fmin/fmax take their first argument and fabs(0) has derivative +1.
 * Pinned: values bit-identical to the oracle's; derivatives against central
 * finite differences of the oracle (tests/test_oracle_tl.py). */
typedef struct orc_dual {
  double v, d;
} orc_dual;
typedef struct orc_tl_chunk {
  int nx, ny, px;
  orc_dual *f[ORC_NUM_2D];
  double *m[ORC_NUM_1D];
} orc_tl_chunk;
int orc_tl_alloc(orc_tl_chunk *t, int nx, int ny);
void orc_tl_free(orc_tl_chunk *t);
/* values and mesh from c, derivatives zero */
void orc_tl_load(orc_tl_chunk *t, const orc_chunk *c);
void orc_tl_ideal_gas(orc_tl_chunk *t, int predict);
void orc_tl_viscosity(orc_tl_chunk *t);
void orc_tl_PdV(orc_tl_chunk *t, int predict, double dt);
void orc_tl_revert(orc_tl_chunk *t);
void orc_tl_accelerate(orc_tl_chunk *t, double dt);
void orc_tl_flux_calc(orc_tl_chunk *t, double dt);
void orc_tl_advec_cell(orc_tl_chunk *t, int sweep_number, int dir);
void orc_tl_advec_mom(orc_tl_chunk *t, int which_vel, int sweep_number, int direction);
void orc_tl_reset_field(orc_tl_chunk *t);
void orc_tl_update_halo(orc_tl_chunk *t, const int fields[15], int depth);
void orc_tl_field_summary_kernel(orc_tl_chunk *t, orc_dual out[5]);
void orc_tl_field_summary(orc_tl_chunk *t, orc_dual out[5]);
void orc_tl_PdV_driver(orc_tl_chunk *t, int predict, double dt);
void orc_tl_advection(orc_tl_chunk *t, int advect_x);
void orc_tl_step_fixed(orc_tl_chunk *t, double dt, int advect_x);
/* Kink census of tangent-linear runs: with eps > 0, every piecewise choice (comparison,
 * fabs, fmin, fmax) whose deciding difference x satisfies |x| < eps |dx| is counted --
 * a choice that differs between the +eps and -eps runs of a central difference along
 * the run's direction, to first order.  Setting eps resets the count. */
void orc_tl_set_kink_eps(double eps);
long long orc_tl_kinks(void);

#ifdef __cplusplus
}
#endif
#endif
