// Adjoint form: the gradient of the final kinetic energy (field_summary's ke,
// src/serial/field_summary.cpp:87) of a fixed-dt run with respect to the initial
// density and energy fields.
//
// Forward.  The fixed-dt step is driver/timestep.cpp:47-69 without calc_dt, then
// driver/hydro.cpp:63-89.  The timestep runs on the reference-plan kernels (one
// launch per reference loop); the Lagrangian phase (PdV predictor chain,
// accelerate, PdV corrector, flux_calc) as one fused pass (fused_lagrange_entry),
// and each advection sweep as one (advec_cell, the halo between, advec_mom for both
// velocities: fused_advection_entry): their arrays are those of the loops bit for
// bit.  The adjoint kernels recompute the loops' intermediates from their inputs,
// so only the inputs are logged.
//
// Checkpointing (two levels).  The forward pass stores the state (density0,
// energy0, xvel0, yvel0, whole arrays) every `every` steps.  The backward pass
// walks the segments last to first: it restores the segment's checkpoint,
// re-runs the segment recording an undo log -- before each kernel that
// overwrites arrays, a copy of those arrays, or the old buffers themselves where a
// pass writes new ones -- and then walks the log backwards:
// restore the kernel's input arrays, run its adjoint.  Halo updates need no
// copy (their adjoint is the transposed reflection and no adjoint reads a value
// a later halo overwrote).  Memory: (steps / every + 1) x 4 state arrays plus
// every x 27 arrays of log (the arrays the step's kernels overwrite).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "clv_internal.h"

namespace clv {
// kernels_adjoint.hip
clv_status adj_ideal_gas(clv_chunk *p, clv_chunk *a, int predict);
clv_status adj_update_halo(clv_chunk *a, const int fields[CLV_NUM_HALO_FIELDS], int depth);
clv_status adj_viscosity(clv_chunk *p, clv_chunk *a);
clv_status adj_PdV(clv_chunk *p, clv_chunk *a, int predict, double dt);
clv_status adj_revert(clv_chunk *p, clv_chunk *a);
clv_status adj_accelerate(clv_chunk *p, clv_chunk *a, double dt);
clv_status adj_flux_calc(clv_chunk *p, clv_chunk *a, double dt);
clv_status adj_advec_cell(clv_chunk *p, clv_chunk *a, int sweep, int dir);
clv_status adj_advec_mom(clv_chunk *p, clv_chunk *a, int sweep, int dir);
clv_status adj_reset_field(clv_chunk *p, clv_chunk *a);
clv_status adj_kinetic_energy(clv_chunk *p, clv_chunk *a, double keb);
}  // namespace clv

using namespace clv;

namespace {

clv_status same_shape(const clv_chunk *p, const clv_chunk *a) {
  if (!p || !a) return set_error(CLV_ERR_ARG, "adjoint: null chunk");
  if (p->d.nx != a->d.nx || p->d.ny != a->d.ny || p->d.pitch != a->d.pitch)
    return set_error(CLV_ERR_ARG, "adjoint: primal and adjoint chunks differ in size or pitch");
  if (p->d.nx < 4 || p->d.ny < 4) return set_error(CLV_ERR_ARG, "adjoint: x_cells, y_cells must be >= 4");
  return CLV_OK;
}

// adjoint and primal on one stream: the adjoint chunk's
clv_status sync_streams(clv_chunk *p, clv_chunk *a) {
  if (p->stream == a->stream) return CLV_OK;
  return hip_check(hipStreamSynchronize(p->stream), "adjoint: primal sync");
}

enum OpKind { OP_IDEAL, OP_HALO, OP_VISC, OP_PDV, OP_REVERT, OP_ACCEL, OP_FLUX, OP_CELL, OP_MOM, OP_RESET };

struct LogEntry {
  OpKind kind;
  int i0, i1;          // predict / sweep, dir / depth
  int fields[CLV_NUM_HALO_FIELDS];
  int nsnap;
  int snap[4];         // field ids copied
  size_t slot;         // first log slot
};

}  // namespace

struct clv_adjoint {
  clv_chunk *p = nullptr;  // primal
  clv_chunk *a = nullptr;  // adjoint
  double dt = 0.0;
  int steps = 0, every = 1;
  size_t arr = 0;          // doubles per 2D array (pitch * (ny + 5))
  double *ckpt = nullptr;  // (nckpt) x 4 arrays
  int nckpt = 0;
  double *log = nullptr;   // the undo log's allocation
  // the slots' arrays (element (0, 0) each; aligned like the chunk arena's arrays).
  // Restoring a snapshot swaps the primal field's pointer with its slot's, so a
  // slot may later hold what was a field's buffer and vice versa.
  std::vector<double *> slot;
  size_t log_slots = 0, log_used = 0;
  std::vector<LogEntry> entries;
  bool forward_done = false;
  double *rho0_bar = nullptr, *e0_bar = nullptr;
};

namespace {

const int kState[4] = {CLV_DENSITY0, CLV_ENERGY0, CLV_XVEL0, CLV_YVEL0};
// undo-log slots per step (forward_step's logged() lists: 15 copies, and 8 buffers
// the advection passes hand over)
constexpr int kSnapsPerStep = 23;

clv_status copy_array(clv_adjoint *g, double *dst, const double *src) {
  return hip_check(hipMemcpyAsync(dst, src, g->arr * sizeof(double), hipMemcpyDeviceToDevice, g->p->stream), "adjoint copy");
}
// element (0, 0) of a 2D array; the array spans pitch * (ny + 5) doubles from there
double *base(clv_chunk *c, int fid) { return c->d.f[fid]; }

clv_status save_state(clv_adjoint *g, int k) {
  for (int q = 0; q < 4; ++q) {
    clv_status st = copy_array(g, g->ckpt + ((size_t)k * 4 + q) * g->arr, base(g->p, kState[q]));
    if (st != CLV_OK) return st;
  }
  return CLV_OK;
}
clv_status load_state(clv_adjoint *g, int k) {
  for (int q = 0; q < 4; ++q) {
    clv_status st = copy_array(g, base(g->p, kState[q]), g->ckpt + ((size_t)k * 4 + q) * g->arr);
    if (st != CLV_OK) return st;
  }
  return CLV_OK;
}

// record (when logging) then run one forward kernel: `snap` copied into the log now;
// the slots of `later` reserved after them, for the caller to fill by buffer exchange
clv_status logged(clv_adjoint *g, bool rec, LogEntry e, std::initializer_list<int> snap, std::initializer_list<int> later = {}) {
  if (!rec) return CLV_OK;
  e.nsnap = 0;
  e.slot = g->log_used;
  for (int f : snap) {
    if (g->log_used >= g->log_slots) return set_error(CLV_ERR_ARG, "adjoint: undo log full");
    clv_status st = copy_array(g, g->slot[g->log_used], base(g->p, f));
    if (st != CLV_OK) return st;
    e.snap[e.nsnap++] = f;
    ++g->log_used;
  }
  for (int f : later) {
    if (g->log_used >= g->log_slots) return set_error(CLV_ERR_ARG, "adjoint: undo log full");
    e.snap[e.nsnap++] = f;
    ++g->log_used;
  }
  g->entries.push_back(e);
  return CLV_OK;
}

LogEntry entry(OpKind k, int i0 = 0, int i1 = 0) {
  LogEntry e{};
  e.kind = k;
  e.i0 = i0;
  e.i1 = i1;
  return e;
}

#define TRY_A(x)                       \
  do {                                 \
    clv_status s_ = (x);               \
    if (s_ != CLV_OK) return s_;       \
  } while (0)

clv_status halo_op(clv_adjoint *g, bool rec, std::initializer_list<int> ids, int depth, bool run = true) {
  LogEntry e = entry(OP_HALO, depth);
  for (int id : ids) e.fields[id] = 1;
  clv_status st = logged(g, rec, e, {});
  if (st != CLV_OK || !run) return st;
  return k_update_halo(g->p, e.fields, depth);
}

// One advection sweep, driver/advection.cpp:47-73 (sweep 1) or :75-103 (sweep 2)
// after its leading halo: advec_cell, update_halo(density1, energy1, xvel1, yvel1,
// mass fluxes, 2), advec_mom(xvel1), advec_mom(yvel1), logged in that order and run
// as one pass.  The velocities' halo runs first (advec_cell does not change them),
// so both snapshots -- advec_cell's inputs, then advec_mom's velocities -- are the
// arrays before the pass.  The pass writes the new density1, energy1, xvel1, yvel1
// into the work arrays and exchanges the buffers, leaving the old ones, unchanged,
// as work1..work4: those become the snapshots (no copy); the mass flux is copied.
clv_status advection_sweep(clv_adjoint *g, bool rec, int sweep, int dir) {
  clv_chunk *p = g->p;
  int vel[CLV_NUM_HALO_FIELDS] = {0};
  vel[CLV_XVEL1] = vel[CLV_YVEL1] = 1;
  TRY_A(k_update_halo(p, vel, 2));
  const size_t cell_slot = g->log_used;
  TRY_A(logged(g, rec, entry(OP_CELL, sweep, dir), {dir == CLV_XDIR ? CLV_MASS_FLUX_X : CLV_MASS_FLUX_Y}, {CLV_DENSITY1, CLV_ENERGY1}));
  TRY_A(halo_op(g, rec, {CLV_DENSITY1, CLV_ENERGY1, CLV_XVEL1, CLV_YVEL1, CLV_MASS_FLUX_X, CLV_MASS_FLUX_Y}, 2, false));
  const size_t mom_slot = g->log_used;
  TRY_A(logged(g, rec, entry(OP_MOM, sweep, dir), {}, {CLV_XVEL1, CLV_YVEL1}));
  TRY_A(fused_advection_entry(p, sweep, dir, false));
  if (rec) {
    std::swap(g->slot[cell_slot + 1], p->d.f[CLV_WORK1]);  // density1 before the pass
    std::swap(g->slot[cell_slot + 2], p->d.f[CLV_WORK2]);  // energy1
    std::swap(g->slot[mom_slot], p->d.f[CLV_WORK3]);       // xvel1 after its halo
    std::swap(g->slot[mom_slot + 1], p->d.f[CLV_WORK4]);   // yvel1
  }
  return CLV_OK;
}

// one fixed-dt step on the reference-plan kernels; advect_x alternates from 1
clv_status forward_step(clv_adjoint *g, int n, bool rec) {
  clv_chunk *p = g->p;
  const double dt = g->dt;
  const int advect_x = (n % 2 == 0) ? 1 : 0;
  // driver/timestep.cpp:47-69 without calc_dt
  TRY_A(logged(g, rec, entry(OP_IDEAL, 0), {CLV_PRESSURE}));
  TRY_A(k_ideal_gas(p, 0));
  TRY_A(halo_op(g, rec, {CLV_PRESSURE, CLV_ENERGY0, CLV_DENSITY0, CLV_XVEL0, CLV_YVEL0}, 1));
  TRY_A(logged(g, rec, entry(OP_VISC), {CLV_VISCOSITY}));
  TRY_A(k_viscosity(p));
  TRY_A(halo_op(g, rec, {CLV_VISCOSITY}, 1));
  // PdV predictor chain (src/serial/PdV.cpp:113-132: PdV, ideal_gas, the pressure
  // halo, revert), accelerate, the PdV corrector and flux_calc, logged in that order
  // and run as one pass (fused_lagrange_entry), whose arrays are those of the
  // kernels bit for bit.  Every snapshot but revert's is an array before the pass;
  // revert's is density1 / energy1 as PdV(predict) leaves them: the arrays before
  // the pass (the ghosts, which PdV leaves alone), their interior written by the
  // pass.  Accelerate and the corrector overwrite u1, v1 and rho1, e1, but no
  // kernel adjoint reads those values before revert (this step) and advec_mom (the
  // previous step) restore their own inputs: nothing to log.
  const size_t revert_slot = g->log_used + 3;
  TRY_A(logged(g, rec, entry(OP_PDV, 1), {CLV_DENSITY1, CLV_ENERGY1}));
  TRY_A(logged(g, rec, entry(OP_IDEAL, 1), {CLV_PRESSURE}));
  TRY_A(halo_op(g, rec, {CLV_PRESSURE}, 1, false));
  TRY_A(logged(g, rec, entry(OP_REVERT), {CLV_DENSITY1, CLV_ENERGY1}));
  TRY_A(logged(g, rec, entry(OP_ACCEL), {}));
  TRY_A(logged(g, rec, entry(OP_PDV, 0), {}));
  TRY_A(logged(g, rec, entry(OP_FLUX), {CLV_VOL_FLUX_X, CLV_VOL_FLUX_Y}));
  if (rec) TRY_A(fused_lagrange_entry(p, dt, g->slot[revert_slot], g->slot[revert_slot + 1]));
  else TRY_A(fused_lagrange_entry(p, dt));
  // driver/advection.cpp:29-105
  const int d1 = advect_x ? CLV_XDIR : CLV_YDIR, d2 = advect_x ? CLV_YDIR : CLV_XDIR;
  TRY_A(halo_op(g, rec, {CLV_ENERGY1, CLV_DENSITY1, CLV_VOL_FLUX_X, CLV_VOL_FLUX_Y}, 2));
  TRY_A(advection_sweep(g, rec, 1, d1));
  TRY_A(advection_sweep(g, rec, 2, d2));
  TRY_A(logged(g, rec, entry(OP_RESET), {CLV_DENSITY0, CLV_ENERGY0, CLV_XVEL0, CLV_YVEL0}));
  TRY_A(k_reset_field(p));
  return CLV_OK;
}

clv_status adjoint_of(clv_adjoint *g, const LogEntry &e) {
  clv_chunk *p = g->p, *a = g->a;
  switch (e.kind) {
    case OP_IDEAL: return adj_ideal_gas(p, a, e.i0);
    case OP_HALO: return adj_update_halo(a, e.fields, e.i0);
    case OP_VISC: return adj_viscosity(p, a);
    case OP_PDV: return adj_PdV(p, a, e.i0, g->dt);
    case OP_REVERT: return adj_revert(p, a);
    case OP_ACCEL: return adj_accelerate(p, a, g->dt);
    case OP_FLUX: return adj_flux_calc(p, a, g->dt);
    case OP_CELL: return adj_advec_cell(p, a, e.i0, e.i1);
    case OP_MOM: return adj_advec_mom(p, a, e.i0, e.i1);
    case OP_RESET: return adj_reset_field(p, a);
  }
  return set_error(CLV_ERR_ARG, "adjoint: unknown log entry");
}

void free_adjoint(clv_adjoint *g) {
  if (g->p) clv_chunk_destroy(g->p);
  if (g->a) clv_chunk_destroy(g->a);
  if (g->ckpt) (void)hipFree(g->ckpt);
  if (g->log) (void)hipFree(g->log);
  delete g;
}

}  // namespace

extern "C" {

// ---- per-kernel adjoints ---------------------------------------------------------------------
#define ADJ_ENTRY(call)                          \
  clv_status st = same_shape(p, a);              \
  if (st != CLV_OK) return st;                   \
  if ((st = sync_streams(p, a)) != CLV_OK) return st; \
  return call;

clv_status clv_adj_ideal_gas(clv_chunk *p, clv_chunk *a, int predict) { ADJ_ENTRY(adj_ideal_gas(p, a, predict ? 1 : 0)) }
clv_status clv_adj_viscosity(clv_chunk *p, clv_chunk *a) { ADJ_ENTRY(adj_viscosity(p, a)) }
clv_status clv_adj_PdV(clv_chunk *p, clv_chunk *a, int predict, double dt) { ADJ_ENTRY(adj_PdV(p, a, predict ? 1 : 0, dt)) }
clv_status clv_adj_revert(clv_chunk *p, clv_chunk *a) { ADJ_ENTRY(adj_revert(p, a)) }
clv_status clv_adj_accelerate(clv_chunk *p, clv_chunk *a, double dt) { ADJ_ENTRY(adj_accelerate(p, a, dt)) }
clv_status clv_adj_flux_calc(clv_chunk *p, clv_chunk *a, double dt) { ADJ_ENTRY(adj_flux_calc(p, a, dt)) }
clv_status clv_adj_reset_field(clv_chunk *p, clv_chunk *a) { ADJ_ENTRY(adj_reset_field(p, a)) }
clv_status clv_adj_kinetic_energy(clv_chunk *p, clv_chunk *a, double ke_bar) { ADJ_ENTRY(adj_kinetic_energy(p, a, ke_bar)) }

clv_status clv_adj_advec_cell(clv_chunk *p, clv_chunk *a, int sweep_number, int dir) {
  if ((sweep_number != 1 && sweep_number != 2) || (dir != CLV_XDIR && dir != CLV_YDIR))
    return set_error(CLV_ERR_ARG, "clv_adj_advec_cell: sweep_number must be 1 or 2, dir CLV_XDIR or CLV_YDIR");
  ADJ_ENTRY(adj_advec_cell(p, a, sweep_number, dir))
}

clv_status clv_adj_advec_mom(clv_chunk *p, clv_chunk *a, int sweep_number, int dir) {
  if ((sweep_number != 1 && sweep_number != 2) || (dir != CLV_XDIR && dir != CLV_YDIR))
    return set_error(CLV_ERR_ARG, "clv_adj_advec_mom: sweep_number must be 1 or 2, dir CLV_XDIR or CLV_YDIR");
  ADJ_ENTRY(adj_advec_mom(p, a, sweep_number, dir))
}

clv_status clv_adj_update_halo(clv_chunk *a, const int fields[CLV_NUM_HALO_FIELDS], int depth) {
  if (!a || !fields) return set_error(CLV_ERR_ARG, "clv_adj_update_halo: null argument");
  if (depth != 1 && depth != 2) return set_error(CLV_ERR_ARG, "clv_adj_update_halo: depth must be 1 or 2");
  if (a->d.nx < 4 || a->d.ny < 4) return set_error(CLV_ERR_ARG, "adjoint: x_cells, y_cells must be >= 4");
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f)
    if (fields[f] == 1 && f == CLV_SOUNDSPEED) return set_error(CLV_ERR_ARG, "clv_adj_update_halo: soundspeed carries no adjoint");
  return adj_update_halo(a, fields, depth);
}

// ---- checkpointed run --------------------------------------------------------------------
clv_status clv_adjoint_create(const clv_chunk *initial, double dt, int steps, int checkpoint_every, clv_adjoint **out) {
  if (!initial || !out) return set_error(CLV_ERR_ARG, "clv_adjoint_create: null argument");
  if (!(dt > 0.0) || steps < 1 || checkpoint_every < 1) return set_error(CLV_ERR_ARG, "clv_adjoint_create: dt > 0, steps >= 1, checkpoint_every >= 1");
  const int nx = initial->d.nx, ny = initial->d.ny;
  if (nx < 4 || ny < 4) return set_error(CLV_ERR_ARG, "clv_adjoint_create: x_cells, y_cells must be >= 4");
  *out = nullptr;
  clv_adjoint *g = new clv_adjoint();
  g->dt = dt;
  g->steps = steps;
  g->every = std::min(checkpoint_every, steps);
  clv_status st = clv_chunk_create(nx, ny, &g->p);
  if (st == CLV_OK) st = clv_chunk_create(nx, ny, &g->a);
  if (st != CLV_OK) {
    free_adjoint(g);
    return st;
  }
  g->arr = (size_t)g->p->d.pitch * (size_t)(ny + 5);
  // copy every field and mesh array of the initial state
  st = hip_check(hipStreamSynchronize(initial->stream), "clv_adjoint_create: sync");
  for (int f = 0; f < CLV_NUM_FIELDS && st == CLV_OK; ++f) {
    if (initial->d.pitch == g->p->d.pitch)
      st = hip_check(hipMemcpy(base(g->p, f), initial->d.f[f], g->arr * sizeof(double), hipMemcpyDeviceToDevice), "copy field");
    else
      st = hip_check(hipMemcpy2D(g->p->d.f[f], g->p->d.pitch * 8, initial->d.f[f], initial->d.pitch * 8, (size_t)(nx + 5) * 8, ny + 5,
                                 hipMemcpyDeviceToDevice),
                     "copy field");
  }
  for (int m = 0; m < CLV_NUM_MESH && st == CLV_OK; ++m)
    st = hip_check(hipMemcpy(g->p->d.m[m], initial->d.m[m], clv_mesh_length(nx, ny, m) * sizeof(double), hipMemcpyDeviceToDevice), "copy mesh");
  g->nckpt = (steps + g->every - 1) / g->every;
  if (st == CLV_OK && hipMalloc((void **)&g->ckpt, (size_t)g->nckpt * 4 * g->arr * sizeof(double)) != hipSuccess)
    st = set_error(CLV_ERR_ALLOC, "clv_adjoint_create: checkpoint allocation failed");
  g->log_slots = (size_t)g->every * kSnapsPerStep;
  // kLeadPad doubles ahead of each slot's element (0, 0), as in the chunk arena, so
  // that a slot swapped in as a field keeps element (2, j) on a 128-byte line
  if (st == CLV_OK && hipMalloc((void **)&g->log, (g->log_slots * g->arr + 32) * sizeof(double)) != hipSuccess)
    st = set_error(CLV_ERR_ALLOC, "clv_adjoint_create: undo-log allocation failed");
  if (st == CLV_OK)
    for (size_t k = 0; k < g->log_slots; ++k) g->slot.push_back(g->log + kLeadPad + k * g->arr);
  if (st == CLV_OK) st = save_state(g, 0);  // checkpoint 0 = the initial state
  if (st == CLV_OK) st = hip_check(hipStreamSynchronize(g->p->stream), "clv_adjoint_create: sync");
  if (st != CLV_OK) {
    free_adjoint(g);
    return st;
  }
  *out = g;
  return CLV_OK;
}

clv_status clv_adjoint_destroy(clv_adjoint *g) {
  if (g) free_adjoint(g);
  return CLV_OK;
}

size_t clv_adjoint_bytes(const clv_adjoint *g) {
  if (!g) return 0;
  return clv_chunk_bytes(g->p) + clv_chunk_bytes(g->a) + ((size_t)g->nckpt * 4 + g->log_slots) * g->arr * sizeof(double);
}

clv_status clv_adjoint_set_initial(clv_adjoint *g, const double *density0, const double *energy0, int64_t pitch) {
  if (!g || !density0 || !energy0) return set_error(CLV_ERR_ARG, "clv_adjoint_set_initial: null argument");
  const int nx = g->p->d.nx, ny = g->p->d.ny;
  if (pitch < nx) return set_error(CLV_ERR_ARG, "clv_adjoint_set_initial: pitch < x_cells");
  clv_status st = load_state(g, 0);
  // interior cells [2, nx+1] x [2, ny+1] from the caller's (ny x pitch) arrays (host or device)
  for (int q = 0; q < 2 && st == CLV_OK; ++q) {
    double *dst = g->p->d.f[q == 0 ? CLV_DENSITY0 : CLV_ENERGY0] + 2 * g->p->d.pitch + 2;
    st = hip_check(hipMemcpy2DAsync(dst, g->p->d.pitch * 8, q == 0 ? density0 : energy0, pitch * 8, (size_t)nx * 8, ny, hipMemcpyDefault,
                                    g->p->stream),
                   "clv_adjoint_set_initial: copy");
  }
  // the start-up halo of the new fields (driver/start.cpp:117-126)
  int fields[CLV_NUM_HALO_FIELDS] = {0};
  fields[CLV_DENSITY0] = fields[CLV_ENERGY0] = 1;
  if (st == CLV_OK) st = k_update_halo(g->p, fields, 2);
  if (st == CLV_OK) st = save_state(g, 0);
  if (st == CLV_OK) st = hip_check(hipStreamSynchronize(g->p->stream), "clv_adjoint_set_initial: sync");
  g->forward_done = false;
  return st;
}

clv_status clv_adjoint_forward(clv_adjoint *g, double summary[5]) {
  if (!g) return set_error(CLV_ERR_ARG, "clv_adjoint_forward: null argument");
  clv_status st = load_state(g, 0);
  g->entries.clear();
  g->log_used = 0;
  for (int n = 0; n < g->steps && st == CLV_OK; ++n) {
    if (n > 0 && n % g->every == 0) st = save_state(g, n / g->every);
    if (st == CLV_OK) st = forward_step(g, n, false);
  }
  double s[5];
  if (st == CLV_OK) st = k_ideal_gas(g->p, 0);  // field_summary: ideal_gas, then the five sums (tree order)
  if (st == CLV_OK) st = k_field_summary(g->p, s);
  if (st != CLV_OK) return st;
  if (summary) std::memcpy(summary, s, sizeof s);
  g->forward_done = true;
  return CLV_OK;
}

clv_status clv_adjoint_backward(clv_adjoint *g, double ke_bar, double *grad_density0, double *grad_energy0, int64_t pitch) {
  if (!g) return set_error(CLV_ERR_ARG, "clv_adjoint_backward: null argument");
  if (!g->forward_done) return set_error(CLV_ERR_ARG, "clv_adjoint_backward: run clv_adjoint_forward first");
  const int nx = g->p->d.nx, ny = g->p->d.ny;
  if ((grad_density0 || grad_energy0) && pitch < nx) return set_error(CLV_ERR_ARG, "clv_adjoint_backward: pitch < x_cells");
  clv_chunk *p = g->p, *a = g->a;
  clv_status st = CLV_OK;
  for (int f = 0; f < CLV_NUM_FIELDS && st == CLV_OK; ++f)
    st = hip_check(hipMemsetAsync(base(a, f), 0, g->arr * sizeof(double), a->stream), "clv_adjoint_backward: clear");
  if (st == CLV_OK) st = hip_check(hipStreamSynchronize(a->stream), "clv_adjoint_backward: sync");
  // d ke / d (final state); p still holds the final state
  if (st == CLV_OK) st = adj_kinetic_energy(p, a, ke_bar);
  if (st == CLV_OK) st = hip_check(hipStreamSynchronize(a->stream), "clv_adjoint_backward: sync");
  for (int k = g->nckpt - 1; k >= 0 && st == CLV_OK; --k) {
    const int n0 = k * g->every, n1 = std::min(g->steps, n0 + g->every);
    st = load_state(g, k);
    g->entries.clear();
    g->log_used = 0;
    for (int n = n0; n < n1 && st == CLV_OK; ++n) st = forward_step(g, n, true);
    if (st == CLV_OK) st = hip_check(hipStreamSynchronize(p->stream), "clv_adjoint_backward: sync");
    for (auto it = g->entries.rbegin(); it != g->entries.rend() && st == CLV_OK; ++it) {
      // restore what the kernel overwrote: the snapshot becomes the field, the
      // field's buffer the (consumed) slot
      for (int q = 0; q < it->nsnap; ++q) std::swap(p->d.f[it->snap[q]], g->slot[it->slot + q]);
      if (st == CLV_OK && p->stream != a->stream) st = hip_check(hipStreamSynchronize(p->stream), "sync");
      if (st == CLV_OK) st = adjoint_of(g, *it);
      if (st == CLV_OK && p->stream != a->stream) st = hip_check(hipStreamSynchronize(a->stream), "sync");
    }
  }
  if (st == CLV_OK) st = hip_check(hipStreamSynchronize(a->stream), "clv_adjoint_backward: sync");
  // the gradient: interior cells of the initial density0 / energy0 adjoints
  for (int q = 0; q < 2 && st == CLV_OK; ++q) {
    double *dst = q == 0 ? grad_density0 : grad_energy0;
    if (!dst) continue;
    const double *src = a->d.f[q == 0 ? CLV_DENSITY0 : CLV_ENERGY0] + 2 * a->d.pitch + 2;
    st = hip_check(hipMemcpy2D(dst, pitch * 8, src, a->d.pitch * 8, (size_t)nx * 8, ny, hipMemcpyDefault), "clv_adjoint_backward: copy");
  }
  g->forward_done = false;  // the primal now holds checkpoint 0's segment
  return st;
}

clv_status clv_adjoint_chunks(clv_adjoint *g, clv_chunk **primal, clv_chunk **adjoint) {
  if (!g) return set_error(CLV_ERR_ARG, "clv_adjoint_chunks: null argument");
  if (primal) *primal = g->p;
  if (adjoint) *adjoint = g->a;
  return CLV_OK;
}

}  // extern "C"
