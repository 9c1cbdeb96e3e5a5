/*
 * clover_mi355x -- C ABI of the MI355X (gfx950) CloverLeaf hydro kernels.
 *
 * One chunk = one 2D staggered grid of x_cells * y_cells cells with two ghost
 * layers, stored on one GPU as 25 FP64 fields (8 cell, 4 node velocity, 2 x-face,
 * 2 y-face, 7 work, volume, xarea, yarea) and 8 1D mesh arrays.  Indices follow
 * the reference's Fortran-heritage convention: x_min = y_min = 1, storage
 * index 0 is x_min - 2, interior cells are [2, x_cells + 1], interior nodes
 * [2, x_cells + 2] (driver/definitions.cpp:4-41, src/serial/ideal_gas.cpp:38-39).
 *
 * Every entry point returns CLV_OK (0) or a negative clv_status; the message of
 * the last failure on the calling thread is clv_last_error().  Calls on one
 * chunk are ordered on that chunk's stream and do not block the host, except
 * those that return host scalars (clv_calc_dt, clv_timestep_kernels,
 * clv_field_summary, clv_timestep, clv_hydro_step) and the explicit copies.  A chunk is not
 * thread-safe.  No HIP or C++ type crosses this boundary.
 *
 * Each per-kernel entry point replaces one reference model entry point for a
 * single chunk and a single tile; the reference declaration it replaces is cited.
 */
#ifndef CLOVER_MI355X_H
#define CLOVER_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int clv_status;
enum {
  CLV_OK = 0,
  CLV_ERR_ARG = -1,      /* bad argument (null pointer, field id, size, depth ...) */
  CLV_ERR_HIP = -2,      /* a HIP runtime call failed */
  CLV_ERR_ALLOC = -3,    /* device or host allocation failed */
  CLV_ERR_SMALL_DT = -4, /* the timestep fell below dtmin (driver/timestep.cpp:96,104-106) */
  CLV_ERR_DECK = -5      /* the input deck is malformed (driver/read_input.cpp:110,197) */
};

/* 2D field ids.  0..14 equal the reference enum field_parameter
 * (driver/definitions.h:62-78), so a halo field mask is the reference's
 * int fields[NUM_FIELDS] array unchanged. */
enum {
  CLV_DENSITY0 = 0, CLV_DENSITY1 = 1, CLV_ENERGY0 = 2, CLV_ENERGY1 = 3,
  CLV_PRESSURE = 4, CLV_VISCOSITY = 5, CLV_SOUNDSPEED = 6,
  CLV_XVEL0 = 7, CLV_XVEL1 = 8, CLV_YVEL0 = 9, CLV_YVEL1 = 10,
  CLV_VOL_FLUX_X = 11, CLV_VOL_FLUX_Y = 12, CLV_MASS_FLUX_X = 13, CLV_MASS_FLUX_Y = 14,
  CLV_WORK1 = 15, CLV_WORK2 = 16, CLV_WORK3 = 17, CLV_WORK4 = 18, CLV_WORK5 = 19, CLV_WORK6 = 20, CLV_WORK7 = 21,
  CLV_VOLUME = 22, CLV_XAREA = 23, CLV_YAREA = 24,
  CLV_NUM_FIELDS = 25,
  CLV_NUM_HALO_FIELDS = 15
};
/* 1D mesh arrays (driver/definitions.cpp:28-35): cell arrays hold n + 4
 * doubles, vertex arrays n + 5 (n = x_cells or y_cells). */
enum {
  CLV_CELLX = 0, CLV_CELLY = 1, CLV_VERTEXX = 2, CLV_VERTEXY = 3,
  CLV_CELLDX = 4, CLV_CELLDY = 5, CLV_VERTEXDX = 6, CLV_VERTEXDY = 7,
  CLV_NUM_MESH = 8
};
static inline size_t clv_mesh_length(int x_cells, int y_cells, int mesh) {
  const int n = (mesh == CLV_CELLX || mesh == CLV_VERTEXX || mesh == CLV_CELLDX || mesh == CLV_VERTEXDX) ? x_cells : y_cells;
  return (size_t)n + ((mesh == CLV_VERTEXX || mesh == CLV_VERTEXY || mesh == CLV_VERTEXDX || mesh == CLV_VERTEXDY) ? 5 : 4);
}

/* Sweep directions (driver/definitions.h:84). */
enum { CLV_XDIR = 1, CLV_YDIR = 2 };

/* Kernel plan for the whole-step entry points.  Both give bit-identical fields. */
enum {
  CLV_PLAN_REFERENCE = 0, /* one launch per reference loop, as the per-kernel entry points */
  CLV_PLAN_FUSED = 1      /* fused, register/LDS-tiled sweeps */
};

typedef struct clv_chunk clv_chunk;

/* A device view of one chunk: where each array lives.  All 2D arrays share
 * one row pitch (in doubles); element (i, j) of field f is f[j * pitch + i]. */
typedef struct clv_chunk_view {
  int x_cells, y_cells;
  int64_t pitch;
  double *field[CLV_NUM_FIELDS];
  double *mesh[CLV_NUM_MESH];
} clv_chunk_view;

/* Geometry and dt controls of a problem (driver/definitions.h:97-155). */
typedef struct clv_state {
  double density, energy, xvel, yvel;
  int geometry; /* 1 rectangle, 2 circle, 3 point (driver/definitions.h:50) */
  double xmin, ymin, xmax, ymax, radius;
} clv_state;

#define CLV_MAX_STATES 16

typedef struct clv_config {
  double xmin, ymin, xmax, ymax;
  int x_cells, y_cells;
  double dtinit, dtmin, dtmax, dtrise;
  double dtc_safe, dtu_safe, dtv_safe, dtdiv_safe;
  double end_time;
  int end_step;
  int summary_frequency, visit_frequency, tiles_per_chunk, test_problem, profiler_on;
  int number_of_states;
  clv_state states[CLV_MAX_STATES]; /* states[0] = background; extents already shrunk by dx/100 */
} clv_config;

/* Driver state carried between steps (driver/definitions.h:265-274). */
typedef struct clv_hydro_state {
  int step;
  int advect_x;
  double time, dt, dtold;
  int jdt, kdt;          /* location of the controlling cell, as printed */
  int dt_control;        /* 1 sound, 2 xvel, 3 yvel, 4 div (src/serial/calc_dt.cpp:131-134) */
  int complete;
} clv_hydro_state;

/* ---- library / errors ------------------------------------------------------------------- */
const char *clv_last_error(void);
const char *clv_version(void);
/* Current HIP device count (0 if none); does not create a context. */
int clv_device_count(void);

/* ---- chunks ----------------------------------------------------------------------------- */
/* Allocates a chunk of x_cells * y_cells on the current device, in one arena.
 * Replaces the per-field allocations of field_type (driver/definitions.cpp:4-41). */
clv_status clv_chunk_create(int x_cells, int y_cells, clv_chunk **out);
/* Wraps device arrays the caller owns (e.g. a reference model's Buffer2D
 * storage); the caller keeps them alive while the chunk exists.  A wrapped
 * chunk always reads volume, xarea and yarea from their arrays (the caller
 * may write them at any time), so CLV_GEOMETRY_MESH is refused for it. */
clv_status clv_chunk_wrap(const clv_chunk_view *view, clv_chunk **out);
clv_status clv_chunk_destroy(clv_chunk *c);
/* The view's pointers allow writes to any array, so taking a view also switches the
 * kernels to reading volume, xarea and yarea (see clv_chunk_geometry_from_mesh). */
clv_status clv_chunk_get_view(const clv_chunk *c, clv_chunk_view *out);
/* Where the kernels take the cell volume and face areas from (no counterpart in
 * the reference, whose kernels always read the arrays).
 * CLV_GEOMETRY_ARRAYS (the default): from the volume, xarea and yarea arrays, as
 *   the reference asks of every optimisation (driver/clover_leaf.cpp:31-36: "the
 *   volume, though constant for all cells, should remain array").
 * CLV_GEOMETRY_MESH: while the three arrays hold what clv_initialise_chunk wrote,
 *   from the 1D spacings: volume(i, j) = celldx(i) * celldy(j), xarea(i, j) =
 *   celldy(j), yarea(i, j) = celldx(i) -- the same doubles (dx * dy on the
 *   reference's uniform mesh), three fewer 2D arrays streamed per kernel.  Takes
 *   effect at the next clv_initialise_chunk; refused for wrapped chunks.
 * Setting CLV_GEOMETRY_ARRAYS takes effect at once. */
enum { CLV_GEOMETRY_ARRAYS = 0, CLV_GEOMETRY_MESH = 1 };
clv_status clv_chunk_set_geometry_mode(clv_chunk *c, int mode);
/* 1 while the kernels (both plans, the per-kernel entry points, calc_dt and
 * field_summary) compute the geometry from the mesh: the chunk is in
 * CLV_GEOMETRY_MESH mode and clv_initialise_chunk has run since.  clv_build_field,
 * clv_field_upload of volume / xarea / yarea, clv_unpack_message into them,
 * clv_mesh_upload of celldx or celldy and clv_chunk_get_view clear it.  0 otherwise,
 * -1 for a null chunk. */
int clv_chunk_geometry_from_mesh(const clv_chunk *c);
/* Orders the chunk's work on a caller stream (a hipStream_t passed as void*;
 * NULL = the null stream). */
clv_status clv_chunk_set_stream(clv_chunk *c, void *hip_stream);
clv_status clv_chunk_synchronize(clv_chunk *c);
/* Device bytes held by the chunk's arena (0 for wrapped chunks). */
size_t clv_chunk_bytes(const clv_chunk *c);

/* Device memory for a plugin model's buffer types: the allocation/free of the
 * reference's Buffer1D/Buffer2D (src/hip/context.h:67-79 alloc/dealloc) and the
 * host copies behind mirrored() (src/hip/context.h:115-123). */
clv_status clv_device_alloc(size_t bytes, void **out);
clv_status clv_device_free(void *p);
clv_status clv_copy_to_host(void *dst, const void *device_src, size_t bytes);
clv_status clv_copy_to_device(void *device_dst, const void *src, size_t bytes);

/* Blocking host copies of one array.  2D arrays use host pitch host_pitch
 * doubles (>= x_cells + 5) and y_cells + 5 rows; 1D arrays their full length. */
clv_status clv_field_upload(clv_chunk *c, int field, const double *host, int64_t host_pitch);
clv_status clv_field_download(clv_chunk *c, int field, double *host, int64_t host_pitch);
clv_status clv_mesh_upload(clv_chunk *c, int mesh, const double *host);
clv_status clv_mesh_download(clv_chunk *c, int mesh, double *host);

/* ---- set-up kernels --------------------------------------------------------------------- */
/* void build_field(global_variables&)              driver/build_field.h:24 */
clv_status clv_build_field(clv_chunk *c);
/* void initialise_chunk(int tile, global_variables&) driver/initialise_chunk.h:24 */
clv_status clv_initialise_chunk(clv_chunk *c, const clv_config *cfg);
/* void generate_chunk(int tile, global_variables&)   driver/generate_chunk.h:24 */
clv_status clv_generate_chunk(clv_chunk *c, const clv_config *cfg);

/* ---- per-step kernels (one reference entry point each) ---------------------------------- */
/* How the per-kernel entry points below run (no counterpart in the reference, whose
 * models all run one loop nest per loop).
 * CLV_ENTRY_REFERENCE (the default): one launch per reference loop, every work array
 *   written as the reference writes it.
 * CLV_ENTRY_FUSED: clv_advec_cell, clv_advec_mom and clv_PdV_driver(predict = 1) run
 *   one launch each, the values the reference passes between its loops through work
 *   arrays kept in registers (the same doubles).  clv_advec_cell writes the new
 *   density1 / energy1 into work arrays 6 / 7 and exchanges the two pairs of pointers
 *   in the chunk; clv_advec_mom does the same with the velocity and work array 5.  A
 *   caller that holds the arrays' addresses (a wrapped chunk's owner) reads them again
 *   with clv_chunk_get_view after those calls.  Every array the reference reads later
 *   holds the reference's values; the work arrays and the ghost layers of an
 *   exchanged array hold other values until the next halo update of that field.
 *   clv_PdV_driver's fused chain writes the pressure halo of every face (a tile whose
 *   four faces are external). */
enum { CLV_ENTRY_REFERENCE = 0, CLV_ENTRY_FUSED = 1 };
clv_status clv_chunk_set_entry_plan(clv_chunk *c, int plan);
/* void ideal_gas(global_variables&, const int tile, bool predict)   driver/ideal_gas.h:24 */
clv_status clv_ideal_gas(clv_chunk *c, int predict);
/* void viscosity(global_variables&)                                 driver/viscosity.h:24 */
clv_status clv_viscosity(clv_chunk *c);
/* void calc_dt(global_variables&, int tile, double& local_dt, std::string& local_control,
 *              double& xl_pos, double& yl_pos, int& jldt, int& kldt) driver/calc_dt.h:26-27
 * Blocks for the minimum.  control/jldt/kldt are the reference's constant
 * bookkeeping (src/serial/calc_dt.cpp:80-83); small = dt_min_val < dtmin. */
clv_status clv_calc_dt(clv_chunk *c, double dtmin, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe,
                       double *dt_min_val, int *dt_control, int *jldt, int *kldt, int *small);
/* driver/timestep.cpp:47-69 in one pass over the cells, for a driver that has already
 * refreshed the density0, energy0, xvel0 and yvel0 halos: ideal_gas(predict = 0),
 * viscosity and calc_dt (src/serial/ideal_gas.cpp:28-48, viscosity.cpp:30-68,
 * calc_dt.cpp:30-111) with pressure, soundspeed and viscosity written as those
 * entry points write them, the pressure halo of depth `pressure_halo_depth` (1 or 2)
 * and the viscosity halo of depth 1 (the reflective halo of every face, as
 * clv_update_halo); outputs as clv_calc_dt.  It replaces the sequence clv_ideal_gas,
 * clv_update_halo(pressure...), clv_viscosity, clv_update_halo(viscosity, 1),
 * clv_calc_dt with the same arrays and dt.  pressure_halo_depth 0: pressure and
 * soundspeed are left as they were and no pressure halo runs (viscosity and dt as
 * above), for a caller whose next call rewrites both before anything reads them --
 * clv_lagrange_kernels stores the predictor's pressure with its depth-1 halo and
 * soundspeed over the same cells.  x_cells, y_cells >= 4. */
clv_status clv_timestep_kernels(clv_chunk *c, double dtmin, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe,
                                int pressure_halo_depth, double *dt_min_val, int *dt_control, int *jldt, int *kldt, int *small);
/* The kernel of void PdV(global_variables&, bool predict)           driver/PdV.h:24
 * (src/serial/PdV.cpp:36-87); clv_PdV_driver adds the predictor chain
 * ideal_gas(predict) + update_halo(pressure, 1) + revert (src/serial/PdV.cpp:113-132). */
clv_status clv_PdV(clv_chunk *c, int predict, double dt);
clv_status clv_PdV_driver(clv_chunk *c, int predict, double dt);
/* The Lagrangian phase of driver/hydro.cpp:63-70 in one pass over the cells, for a
 * driver whose timestep left pressure = ideal_gas(density0, energy0), the
 * density0/energy0/xvel0/yvel0/pressure halos of depth 1 and the viscosity halo of
 * depth 1 (driver/timestep.cpp:47-69): PdV(predict) with its ideal_gas(predict),
 * pressure halo and revert (src/serial/PdV.cpp:92-133), accelerate, PdV(correct) and
 * flux_calc, every array as those entry points leave it (the predictor's pressure,
 * its depth-1 reflective halo on every face, and its soundspeed included).  It
 * replaces clv_PdV_driver(c, 1, dt), clv_accelerate, clv_PdV_driver(c, 0, dt),
 * clv_flux_calc.  x_cells, y_cells >= 4. */
clv_status clv_lagrange_kernels(clv_chunk *c, double dt);
/* One sweep of driver/advection.cpp -- advec_cell_driver(sweep_number, direction),
 * update_halo(density1, energy1, xvel1, yvel1, mass_flux_x, mass_flux_y, 2),
 * advec_mom_driver(xvel), advec_mom_driver(yvel) (:47-73 for sweep 1, :75-103 for
 * sweep 2) -- in one pass, for a driver that has already refreshed the xvel1/yvel1
 * halo of depth 2 (and, before sweep 1, the energy1, density1, vol_flux_x/y halo of
 * depth 2, :39-45).  density1, energy1, xvel1 and yvel1 come back in other buffers
 * (re-read them with clv_chunk_get_view), the mass fluxes in place, every value as
 * those calls leave it, ghosts included; the work arrays hold scratch.  The halo is
 * the reflective one of every face.  x_cells, y_cells >= 4. */
clv_status clv_advection_pass(clv_chunk *c, int sweep_number, int direction);
/* clv_advection_pass followed by void reset_field(global_variables&)   driver/reset_field.h:24
 * (src/serial/reset_field.cpp:29-42): the pass also stores its new density, energy and
 * velocities into density0, energy0, xvel0 and yvel0 (interior cells and nodes, as
 * reset_field copies them), so that no separate copy pass runs.  Every array as
 * clv_advection_pass then clv_reset_field leave it. */
clv_status clv_advection_pass_reset(clv_chunk *c, int sweep_number, int direction);
/* void revert(global_variables&)                                    driver/revert.h:24 */
clv_status clv_revert(clv_chunk *c);
/* void accelerate(global_variables&)                                driver/accelerate.h:24 */
clv_status clv_accelerate(clv_chunk *c, double dt);
/* void flux_calc(global_variables&)                                 driver/flux_calc.h:24 */
clv_status clv_flux_calc(clv_chunk *c, double dt);
/* void advec_cell_driver(global_variables&, int tile, int sweep_number, int direction)
 *                                                                   driver/advec_cell.h:24 */
clv_status clv_advec_cell(clv_chunk *c, int sweep_number, int direction);
/* void advec_mom_driver(global_variables&, int tile, int which_vel, int direction,
 *                       int sweep_number)                           driver/advec_mom.h:24 */
clv_status clv_advec_mom(clv_chunk *c, int which_vel, int direction, int sweep_number);
/* void reset_field(global_variables&)                               driver/reset_field.h:24 */
clv_status clv_reset_field(clv_chunk *c);
/* void update_halo(global_variables&, int fields[NUM_FIELDS], int depth)
 *                                                                   driver/update_halo.h:24
 * The reflective self halo of a single chunk and tile (every face external);
 * fields[f] == 1 selects field f (0..14), depth is 1 or 2. */
clv_status clv_update_halo(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth);

/* ---- decomposed domains: tiles of a chunk, chunks of a mesh ----------------------------- */
/* Faces, in the order of the reference's chunk_neighbour_type (driver/definitions.h:57),
 * and the data types of exchange messages (driver/definitions.h:80). */
enum { CLV_FACE_LEFT = 0, CLV_FACE_RIGHT = 1, CLV_FACE_BOTTOM = 2, CLV_FACE_TOP = 3 };
enum { CLV_CELL_DATA = 1, CLV_VERTEX_DATA = 2, CLV_X_FACE_DATA = 3, CLV_Y_FACE_DATA = 4 };
/* void initialise_chunk(int tile, global_variables&)   driver/initialise_chunk.h:24
 * for a tile whose first cell is (t_left, t_bottom) of the grid
 * (src/serial/initialise_chunk.cpp:32-37); the chunk is the tile's size. */
clv_status clv_initialise_tile(clv_chunk *c, const clv_config *cfg, int t_left, int t_bottom);
/* update_halo_kernel of a tile whose faces are not all external
 *                                       src/serial/update_halo.cpp:33-684, :41-71
 * external[face] != 0 reflects that face; the other faces keep the ghosts their
 * tile or chunk exchange wrote (corners included). */
clv_status clv_update_halo_faces(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth, const int external[4]);
/* void update_tile_halo_{l,r,b,t}_kernel(...)   driver/update_tile_halo_kernel.h:25-86
 * The ghosts of `tile` on `face` from the neighbouring tile's interior
 * (src/serial/update_tile_halo_kernel.cpp:30-804).  Neighbours across a
 * left/right face have equal y_cells, across a bottom/top face equal x_cells;
 * both chunks on one device.  Ordered on the tile's stream (the neighbour's
 * stream is synchronised first when it differs).  The reference's own driver
 * runs every tile's bottom/top copies before any left/right copy
 * (driver/update_tile_halo.cpp:26-91). */
clv_status clv_update_tile_halo(clv_chunk *tile, clv_chunk *neighbour, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth);
/* void clover_pack_message_{left,right,bottom,top}(...) / clover_unpack_message_*(...)
 *                                                        driver/pack_kernel.h:23-47
 * One exchange message of one array (src/serial/pack_kernel.cpp:27-303).  field:
 * a device array of c's layout (pitch, y_cells + 5 rows); buffer: device memory of
 * buffer_len doubles; the message starts at element offset.  Left/right
 * messages hold offset + j + k*depth (rows k, layers j), bottom/top
 * offset + k + j*depth. */
clv_status clv_pack_message(clv_chunk *c, const double *field, int field_type, int face, int depth, double *buffer, size_t buffer_len,
                            int64_t offset);
clv_status clv_unpack_message(clv_chunk *c, double *field, int field_type, int face, int depth, const double *buffer, size_t buffer_len,
                              int64_t offset);
/* void clover_pack_{left,right,bottom,top}(...) / clover_unpack_*(...)   driver/comms.cpp:282-930
 * Every requested field of tile c into one face message: the k-th requested
 * field (in field order) at k * depth * (chunk_cells + 5) + tile_offset, where
 * chunk_cells is the chunk's cell count along the face (y_cells for left/right)
 * and tile_offset = (the tile's first cell along the face - the chunk's) * depth
 * (src/serial/comms_kernel.cpp:64-73, driver/comms.cpp:286).  buffer_len must be
 * at least clv_message_length. */
size_t clv_message_length(const int fields[CLV_NUM_HALO_FIELDS], int depth, int chunk_cells);
clv_status clv_pack_fields(clv_chunk *c, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth, int chunk_cells, int64_t tile_offset,
                           double *buffer, size_t buffer_len);
clv_status clv_unpack_fields(clv_chunk *c, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth, int chunk_cells, int64_t tile_offset,
                             const double *buffer, size_t buffer_len);
/* The sums of void field_summary(global_variables&, parallel_&)     driver/field_summary.h:25
 * (src/serial/field_summary.cpp:60-92) over the current fields, without the
 * ideal_gas call that precedes them in the reference; out = {vol, mass, ie, ke,
 * press}.  Blocks.  Fixed-order reduction: run-to-run reproducible. */
clv_status clv_field_summary(clv_chunk *c, double out[5]);
/* Summation order of clv_field_summary / clv_hydro_summary on this chunk.
 * CLV_SUMMARY_TREE (default): fixed-order tree, reproducible, within ~1e-15
 * relative of the exact sums.  CLV_SUMMARY_SERIAL: the serial model's own
 * left-to-right order over the linear cell index, so the five sums are
 * bit-identical to the reference serial model's; the fold is sequential
 * (a few hundred M cells/s), meant for validation runs. */
enum { CLV_SUMMARY_TREE = 0, CLV_SUMMARY_SERIAL = 1 };
clv_status clv_chunk_set_summary_order(clv_chunk *c, int order);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the sums of clv_field_summary (tree order, the same kernels, so the
 * same doubles) over the arrays of a caller's view -- density0, energy0, pressure,
 * xvel0, yvel0 and volume are read, the geometry always from the volume array --
 * with no chunk and no allocation.  Its block partials and results live in one fixed
 * device array per GPU (81,984 bytes: 5 x 2048 partials + 5 results), kept until the
 * process ends; the reduction never runs more than 2048 blocks, so no grid is too
 * large (views with x_cells or y_cells < 2 or pitch < x_cells + 5 are refused).
 * Ordered on hip_stream (a hipStream_t as void*, NULL = the null stream); blocks
 * for out.  Calls that overlap on one GPU (other threads, other streams) share that
 * array and race: serialise them. */
clv_status clv_field_summary_noalloc(const clv_chunk_view *view, void *hip_stream, double out[5]);

/* ---- whole-step forms (the host loop of driver/hydro.cpp) -------------------------------- */
/* Parses deck text with driver/read_input.cpp semantics (defaults :39-69). */
clv_status clv_deck_parse(const char *text, clv_config *out);
/* driver/start.cpp:69-127: build, initialise, generate, ideal_gas, depth-2 halo.
 * Resets *st; summary0 (may be NULL) receives the step-0 field summary. */
clv_status clv_hydro_start(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, double summary0[5]);
/* driver/timestep.cpp:37-108: ideal_gas, halos, viscosity, calc_dt, dt selection.
 * With CLV_PLAN_FUSED the pressure, soundspeed and viscosity arrays are not
 * written (the fused step recomputes them where it needs them); follow it with
 * the fused step, not with the per-kernel entry points. */
clv_status clv_timestep(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, int plan);
/* One step of driver/hydro.cpp:57-89 (timestep, PdV predict, accelerate, PdV
 * correct, flux_calc, advection, reset_field, advect_x flip, time += dt); sets
 * st->complete when driver/hydro.cpp:108 says so.  Returns CLV_ERR_SMALL_DT
 * when dt < dtmin.  With CLV_PLAN_FUSED the state the next step reads equals the
 * reference plan's (density0 and energy0 on the interior cells, xvel0 and yvel0 on
 * the interior and boundary nodes, the volume fluxes); pressure, soundspeed,
 * viscosity, the mass fluxes, the 1-fields and the work arrays hold other values
 * (clv_hydro_summary recomputes what the summary reads). */
clv_status clv_hydro_step(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, int plan);
/* field_summary with its ideal_gas (src/serial/field_summary.cpp:51-99).  With the
 * tree summation order the two run as one pass over the cells: pressure and
 * soundspeed are written as ideal_gas writes them, and the sums use the pressure
 * just written (the same doubles as clv_ideal_gas then clv_field_summary). */
clv_status clv_hydro_summary(clv_chunk *c, double out[5]);
/* The phases of one step of driver/hydro.cpp:59-85 on the reference plan, for a host
 * loop that acts between them as the reference's does (its --dump writes the fields
 * after each, driver/hydro.cpp:60-85).  Called in order they make one
 * clv_hydro_step(..., CLV_PLAN_REFERENCE): CLV_PHASE_TIMESTEP advances st->step and
 * chooses dt (clv_timestep; CLV_ERR_SMALL_DT as there); CLV_PHASE_END_STEP runs no
 * kernel: it flips advect_x, advances time and sets complete (:86-89, :108). */
enum {
  CLV_PHASE_TIMESTEP = 0,    /* timestep()            driver/timestep.cpp:37-108 */
  CLV_PHASE_PDV_PREDICT = 1, /* PdV(globals, true)    src/serial/PdV.cpp:92-133 */
  CLV_PHASE_ACCELERATE = 2,  /* accelerate()          src/serial/accelerate.cpp */
  CLV_PHASE_PDV_CORRECT = 3, /* PdV(globals, false) */
  CLV_PHASE_FLUX_CALC = 4,   /* flux_calc() */
  CLV_PHASE_ADVECTION = 5,   /* advection()           driver/advection.cpp:29-105 */
  CLV_PHASE_RESET_FIELD = 6, /* reset_field() */
  CLV_PHASE_END_STEP = 7     /* advect_x, time, complete */
};
clv_status clv_hydro_phase(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, int phase);

/* ---- batched form: many independent problems advanced together ------------------------- */
/* No counterpart in the reference, whose driver advances one problem
 * (driver/hydro.cpp:53-223).  Each member is a chunk of its own (clv_chunk_create
 * + clv_hydro_start); a batch step runs the step of clv_hydro_step(...,
 * CLV_PLAN_FUSED) for every member that is not complete, with one persistent
 * launch per phase over the tiles of all members, dt chosen per member on the
 * device, and the whole step replayed as a captured hipGraph.  Each member's
 * fields and dt sequence are bit-identical to stepping it alone. */
typedef struct clv_batch clv_batch;

/* Tile families of a batched step (one per fused kernel; the advection passes
 * advec_cell + advec_mom in x and in y; sweep variants share). */
enum { CLV_BF_TIMESTEP = 0, CLV_BF_LAGRANGE, CLV_BF_CM_X, CLV_BF_CM_Y, CLV_BF_COUNT };
enum { CLV_BATCH_ORDER_GIVEN = 0, CLV_BATCH_ORDER_LARGEST_FIRST = 1 };
/* Work schedule of one family: rows per tile (8 ... 4096), persistent grid in
 * blocks per CU (0: one block per tile) and the order of members in the tile list. */
typedef struct clv_batch_family {
  int seg;
  int blocks_per_cu;
  int order;
} clv_batch_family;
typedef struct clv_batch_schedule {
  clv_batch_family fam[CLV_BF_COUNT];
} clv_batch_schedule;

/* Deck text of problem `index` of the seeded batch workload: a square grid of
 * side 256 + 128 k (k uniform in 0..14) on [0,10]^2, state 1 density 0.2 energy
 * 1.0, a state-2 rectangle with corners from two sorted U(0,10) draws per axis,
 * density U[0.5,2.0], energy U[1.5,3.5], dt controls of InputDecks/clover_bm_short.in,
 * end_step 100.  Draws come from splitmix64 streams keyed by (seed, index).
 * Writes at most len bytes including the terminating NUL. */
clv_status clv_batch_deck(uint64_t seed, int index, char *text, size_t len);
/* Members: n owned chunks on the current device, already started, with their
 * configs and driver states (all with the same advect_x).  hist_len: dt values
 * kept per member (step s at index s - 1).  While a chunk is in a batch, step it
 * only through the batch; reading it (downloads, summaries) is fine between runs.
 * If every member's volume and face areas come from its mesh at creation
 * (clv_chunk_geometry_from_mesh), the batch computes them for all members, and
 * clv_batch_run refuses to run once a member's flag has been cleared. */
clv_status clv_batch_create(clv_chunk *const *chunks, const clv_config *cfgs, const clv_hydro_state *states, int n, int hist_len,
                            clv_batch **out);
clv_status clv_batch_destroy(clv_batch *b);
clv_status clv_batch_get_schedule(const clv_batch *b, clv_batch_schedule *out);
clv_status clv_batch_set_schedule(clv_batch *b, const clv_batch_schedule *s);
/* Times each family's kernel on the current fields over candidate schedules
 * (rows per tile, persistent grid, member order; `reps` launches each, by
 * coordinate descent) and sets the fastest.  Does not change the members'
 * state.  out / family_ms (may be NULL) receive the schedule and its per-launch
 * times in ms. */
clv_status clv_batch_autotune(clv_batch *b, int reps, clv_batch_schedule *out, double family_ms[CLV_BF_COUNT]);
/* Runs up to max_steps batch steps, stopping early once every member is
 * complete.  use_graph: replay the step as a hipGraph (else direct launches).
 * *steps_done (may be NULL) receives the batch steps issued.  Blocks.  Returns
 * CLV_ERR_SMALL_DT if a member's dt fell below its dtmin (that member stops). */
clv_status clv_batch_run(clv_batch *b, int max_steps, int use_graph, int *steps_done);
/* Driver state of every member (n entries), as clv_hydro_step would have left it. */
clv_status clv_batch_states(clv_batch *b, clv_hydro_state *out);
/* dt of steps 1 .. *count of member `member` (at most max values). */
clv_status clv_batch_dt_history(clv_batch *b, int member, double *dt, int max, int *count);

/* ---- adjoint form ----------------------------------------------------------------------- */
/* No counterpart in the reference (it has no derivatives).  Reverse-mode
 * adjoints of the reference kernels, for gradients of a fixed-dt run (dt given,
 * calc_dt not run).  p holds the forward kernel's INPUT state, a (same size and
 * pitch) the adjoints: on entry the adjoints of the kernel's outputs (plus what
 * has accumulated on its inputs), on return the inputs' adjoints accumulated and
 * the overwritten outputs' adjoints cleared.  Piecewise definitions (upwind
 * choice, van Leer limiters, viscosity switch, fmin/fmax, fabs) are
 * differentiated on the branch the forward took.  Geometry is constant;
 * soundspeed (read only by calc_dt) carries no adjoint; the work arrays of both
 * p and a are scratch (p's receive the forward's intermediate values).  Each
 * replaces no reference entry point; the forward kernel each transposes is
 * cited. */
clv_status clv_adj_ideal_gas(clv_chunk *p, clv_chunk *a, int predict);          /* src/serial/ideal_gas.cpp:28-48 */
clv_status clv_adj_viscosity(clv_chunk *p, clv_chunk *a);                       /* src/serial/viscosity.cpp:30-68 */
clv_status clv_adj_PdV(clv_chunk *p, clv_chunk *a, int predict, double dt);     /* src/serial/PdV.cpp:36-87 */
clv_status clv_adj_revert(clv_chunk *p, clv_chunk *a);                          /* src/serial/revert.cpp:29-41 */
clv_status clv_adj_accelerate(clv_chunk *p, clv_chunk *a, double dt);           /* src/serial/accelerate.cpp:28-56 */
clv_status clv_adj_flux_calc(clv_chunk *p, clv_chunk *a, double dt);            /* src/serial/flux_calc.cpp:27-43 */
clv_status clv_adj_advec_cell(clv_chunk *p, clv_chunk *a, int sweep_number, int dir); /* src/serial/advec_cell.cpp:28-220 */
/* both velocities of one sweep (the yvel pass reuses the xvel pass's node
 * arrays, src/serial/advec_mom.cpp:81-103): advec_mom(xvel) then advec_mom(yvel) */
clv_status clv_adj_advec_mom(clv_chunk *p, clv_chunk *a, int sweep_number, int dir); /* src/serial/advec_mom.cpp:30-225 */
clv_status clv_adj_reset_field(clv_chunk *p, clv_chunk *a);                     /* src/serial/reset_field.cpp:28-51 */
clv_status clv_adj_update_halo(clv_chunk *a, const int fields[CLV_NUM_HALO_FIELDS], int depth); /* src/serial/update_halo.cpp:33-684 */
/* ke of field_summary (src/serial/field_summary.cpp:64-90), times ke_bar */
clv_status clv_adj_kinetic_energy(clv_chunk *p, clv_chunk *a, double ke_bar);

/* Checkpointed gradient of the final kinetic energy of a fixed-dt run with
 * respect to the initial density0 and energy0 fields (interior cells).  The
 * forward pass stores the state every checkpoint_every steps; the backward
 * pass re-runs each segment with an undo log and walks it backwards.
 * Memory: two chunks, (steps / checkpoint_every + 1) x 4 arrays, and
 * checkpoint_every x 27 arrays. */
typedef struct clv_adjoint clv_adjoint;
/* initial: a started chunk (clv_hydro_start) holding the initial state; copied. */
clv_status clv_adjoint_create(const clv_chunk *initial, double dt, int steps, int checkpoint_every, clv_adjoint **out);
clv_status clv_adjoint_destroy(clv_adjoint *g);
size_t clv_adjoint_bytes(const clv_adjoint *g);
/* Replace the initial interior density0 / energy0 (y_cells rows of pitch doubles,
 * host or device memory); the start-up halo is re-applied. */
clv_status clv_adjoint_set_initial(clv_adjoint *g, const double *density0, const double *energy0, int64_t pitch);
/* Run the steps; summary (may be NULL) receives the final field_summary
 * {vol, mass, ie, ke, press}. */
clv_status clv_adjoint_forward(clv_adjoint *g, double summary[5]);
/* After clv_adjoint_forward: ke_bar * d ke / d density0 and d energy0 of the
 * initial interior cells into (y_cells rows of pitch doubles, host or device;
 * either may be NULL). */
clv_status clv_adjoint_backward(clv_adjoint *g, double ke_bar, double *grad_density0, double *grad_energy0, int64_t pitch);
/* the primal and adjoint chunks (for inspection; owned by g).  The backward pass
 * restores logged arrays by exchanging buffers, so a field pointer of the primal
 * read before clv_adjoint_backward may be stale after it: read it again. */
clv_status clv_adjoint_chunks(clv_adjoint *g, clv_chunk **primal, clv_chunk **adjoint);

/* ---- measurement ------------------------------------------------------------------------ */
/* Per-kernel HIP-event timing on the chunk's stream: when enabled, every launch
 * is bracketed by events and clv_profile_read() returns, for each kernel name,
 * the launch count and total milliseconds.  names must hold max entries. */
clv_status clv_profile_enable(clv_chunk *c, int on);
int clv_profile_read(clv_chunk *c, const char **names, int64_t *launches, double *total_ms, int max);
clv_status clv_profile_reset(clv_chunk *c);
/* HBM bandwidth probe on the current device (no counterpart in the reference):
 * a plain copy of `bytes` (read + write, 16-byte accesses), timed over `iters`
 * launches with HIP events; *copy_gbps = 2 * bytes * iters / time.  Allocates and
 * frees two buffers of `bytes`. */
clv_status clv_bandwidth_probe(size_t bytes, int iters, double *copy_gbps);
/* Access-pattern probe on the current device (no counterpart in the reference):
 * the fused kernels' march (64-column strips, `seg`-row segments, workgroups
 * x-fastest on one XCD band each) over the interior of an x_cells * y_cells
 * grid, reading `reads` and writing `writes` FP64 arrays per cell with no
 * arithmetic (streaming stores); the arrays are padded like a chunk's, element
 * (2, j) on a 128-byte line.  (reads, writes) is (7, 1), (8, 6), (7, 5), (6, 5), (4, 1),
 * (5, 6), (5, 5), (7, 0), (7, 6), (4, 0), (4, 6), (7, 4), (6, 4) or (5, 4).
 * Allocates and frees the arrays; *ms_per_launch = the mean of `iters` timed
 * launches. */
clv_status clv_march_probe(int x_cells, int y_cells, int reads, int writes, int seg, int iters, double *ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif
