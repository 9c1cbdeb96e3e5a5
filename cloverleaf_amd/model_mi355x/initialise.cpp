// create_context / report_context (driver/initialise.h:160-162) and the tile ->
// chunk binding of the MI355X model: each tile of each chunk is wrapped once as a
// clover_mi355x chunk of the tile's size.
#include "initialise.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "mi355x_model.h"

model create_context(bool silent, const std::vector<std::string> &args) {
  const int n = clv_device_count();
  std::vector<std::string> devices;
  for (int d = 0; d < (n > 0 ? n : 1); ++d) devices.push_back("HIP device " + std::to_string(d) + " (gfx950, clover_mi355x)");
  auto [device, parsed] = list_and_parse<std::string>(
      silent, devices, [](const auto &d) { return d; }, args);
  if (n < 1) {
    std::fprintf(stderr, "mi355x model: no HIP device visible\n");
    std::exit(EXIT_FAILURE);
  }
  clover::context ctx{};
  ctx.device = 0;
  for (int d = 0; d < n; ++d)
    if (devices[d] == device) ctx.device = d;
  return {ctx, "MI355X (clover_mi355x, gfx950 HIP kernels)", true, parsed};
}

void report_context(const clover::context &ctx) {
  std::printf(" - Device:    %d (%s)\n", ctx.device, clv_version());
}

namespace mi355x {

static std::map<const void *, clv_chunk *> g_chunks;

[[noreturn]] void unsupported(const char *what) {
  std::fprintf(stderr, "mi355x model: %s is not implemented by this model\n", what);
  std::abort();
}

namespace {
clover::Buffer2D<double> *const *fields_of(field_type &f, clover::Buffer2D<double> *out[CLV_NUM_FIELDS]) {
  clover::Buffer2D<double> *b2[CLV_NUM_FIELDS] = {
      &f.density0,   &f.density1,    &f.energy0,     &f.energy1,     &f.pressure,    &f.viscosity,   &f.soundspeed,
      &f.xvel0,      &f.xvel1,       &f.yvel0,       &f.yvel1,       &f.vol_flux_x,  &f.vol_flux_y,  &f.mass_flux_x,
      &f.mass_flux_y, &f.work_array1, &f.work_array2, &f.work_array3, &f.work_array4, &f.work_array5, &f.work_array6,
      &f.work_array7, &f.volume,      &f.xarea,       &f.yarea};
  for (int k = 0; k < CLV_NUM_FIELDS; ++k) out[k] = b2[k];
  return out;
}
}  // namespace

clv_chunk *chunk_of(global_variables &globals, const double *field) {
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    clover::Buffer2D<double> *b2[CLV_NUM_FIELDS];
    fields_of(globals.chunk.tiles[tile].field, b2);
    for (int k = 0; k < CLV_NUM_FIELDS; ++k)
      if (b2[k]->data == field) return chunk(globals, tile);
  }
  std::fprintf(stderr, "mi355x model: a halo or message buffer that belongs to no tile\n");
  std::abort();
}

void external_faces(const global_variables &globals, int tile, int external[4]) {
  const tile_info &t = globals.chunk.tiles[tile].info;
  const chunk_type &c = globals.chunk;
  const bool at[4] = {t.t_left == c.left, t.t_right == c.right, t.t_bottom == c.bottom, t.t_top == c.top};
  for (int face = 0; face < 4; ++face) external[face] = at[face] && c.chunk_neighbours[face] == external_face;
}

namespace {
// The calls recorded for a one-launch chain (one tile, see mi355x_model.h), in order.
enum RecOp { R_GAS, R_P_HALO, R_VISC, R_Q_HALO, R_PDV_PREDICT, R_ACCEL, R_PDV_CORRECT, R_CELL, R_ADV_HALO, R_MOM_X, R_MOM_Y };
struct Rec {
  RecOp op;
  double arg;         // halo depth or dt
  int sweep = 0, dir = 0;  // advection
};
clv_chunk *g_rec_chunk = nullptr;
global_variables *g_rec_globals = nullptr;
std::vector<Rec> g_rec;
// The chunk whose pressure (with its depth-1 halo) and soundspeed the timestep chain
// left unwritten, since the Lagrangian chain rewrites them; anything else that
// touches the tile first has them written as ideal_gas and the halo would have.
clv_chunk *g_pc_owed = nullptr;

void settle_pressure() {
  if (!g_pc_owed) return;
  clv_chunk *c = g_pc_owed;
  g_pc_owed = nullptr;
  int f[CLV_NUM_HALO_FIELDS] = {0};
  f[CLV_PRESSURE] = 1;
  check(clv_ideal_gas(c, 0), "ideal_gas");
  check(clv_update_halo(c, f, 1), "update_halo");
}

bool chain_allowed(const global_variables &globals) {
  const char *e = std::getenv("CLOVER_MI355X_CHAINS");
  if (e && std::string(e) == "off") return false;
  if (!fused_entries(globals) || globals.profiler_on || globals.config.tiles_per_chunk != 1) return false;
  for (int face = 0; face < 4; ++face)
    if (globals.chunk.chunk_neighbours[face] != external_face) return false;
  const tile_info &t = globals.chunk.tiles[0].info;
  return t.t_xmax - t.t_xmin + 1 >= 4 && t.t_ymax - t.t_ymin + 1 >= 4;
}

bool recorded(std::initializer_list<RecOp> ops) {
  if (g_rec.size() != ops.size()) return false;
  size_t k = 0;
  for (RecOp op : ops)
    if (g_rec[k++].op != op) return false;
  return true;
}

// start recording on tile 0 (after running what an earlier record left; the
// Lagrangian chain keeps an owed pressure, which it rewrites)
bool start_record(global_variables &globals, RecOp op, double arg, int sweep = 0, int dir = 0) {
  if (!g_rec.empty() || op != R_PDV_PREDICT) flush_deferred();
  clv_chunk *c = chunk_keep(globals, 0);
  g_rec_chunk = c;
  g_rec_globals = &globals;
  g_rec.push_back({op, arg, sweep, dir});
  return true;
}

// the halo request between advec_cell and advec_mom (driver/advection.cpp:57-63, 87-93)
bool is_advection_halo(const int fields[CLV_NUM_HALO_FIELDS], int depth) {
  if (depth != 2) return false;
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f) {
    const bool want = f == CLV_DENSITY1 || f == CLV_ENERGY1 || f == CLV_XVEL1 || f == CLV_YVEL1 || f == CLV_MASS_FLUX_X ||
                      f == CLV_MASS_FLUX_Y;
    if ((fields[f] == 1) != want) return false;
  }
  return true;
}
}  // namespace

void flush_deferred() {
  settle_pressure();
  if (g_rec.empty()) return;
  const std::vector<Rec> rec = g_rec;
  clv_chunk *c = g_rec_chunk;
  g_rec.clear();
  if (rec.back().op == R_MOM_Y) {  // a whole recorded sweep: the pass as advec_mom(yvel) would have run it
    check(clv_advection_pass(c, rec[0].sweep, rec[0].dir), "advection_pass");
    refresh_buffers(*g_rec_globals, 0);
    return;
  }
  for (const Rec &r : rec) {
    int f[CLV_NUM_HALO_FIELDS] = {0};
    switch (r.op) {
      case R_GAS: check(clv_ideal_gas(c, 0), "ideal_gas"); break;
      case R_P_HALO:
        f[CLV_PRESSURE] = 1;
        check(clv_update_halo(c, f, (int)r.arg), "update_halo");
        break;
      case R_VISC: check(clv_viscosity(c), "viscosity"); break;
      case R_Q_HALO:
        f[CLV_VISCOSITY] = 1;
        check(clv_update_halo(c, f, 1), "update_halo");
        break;
      case R_PDV_PREDICT: check(clv_PdV_driver(c, 1, r.arg), "PdV"); break;
      case R_ACCEL: check(clv_accelerate(c, r.arg), "accelerate"); break;
      case R_PDV_CORRECT: check(clv_PdV_driver(c, 0, r.arg), "PdV"); break;
      case R_CELL:
        check(clv_advec_cell(c, r.sweep, r.dir), "advec_cell");
        refresh_buffers(*g_rec_globals, 0);
        break;
      case R_ADV_HALO:  // the velocities' part ran when it was recorded
        f[CLV_DENSITY1] = f[CLV_ENERGY1] = f[CLV_MASS_FLUX_X] = f[CLV_MASS_FLUX_Y] = 1;
        check(clv_update_halo(c, f, 2), "update_halo");
        break;
      case R_MOM_X:
        check(clv_advec_mom(c, 1, r.dir, r.sweep), "advec_mom");
        refresh_buffers(*g_rec_globals, 0);
        break;
      case R_MOM_Y: break;  // only ever last in a record, handled above
    }
  }
}

// ---- the timestep chain: ideal_gas, update_halo(pressure ...), viscosity, update_halo(viscosity), calc_dt
bool defer_ideal_gas(global_variables &globals, int tile, bool predict) {
  if (predict || tile != 0 || !chain_allowed(globals)) return false;
  return start_record(globals, R_GAS, 0);
}

bool defer_viscosity(global_variables &globals) {
  (void)globals;
  if (!recorded({R_GAS, R_P_HALO})) return false;
  g_rec.push_back({R_VISC, 0});
  return true;
}

void defer_halo(global_variables &globals, int fields[CLV_NUM_HALO_FIELDS], int depth) {
  (void)globals;
  if (g_rec.empty()) {
    settle_pressure();  // an owed pressure gets its halo before any other halo call
    return;
  }
  if (recorded({R_GAS}) && fields[CLV_PRESSURE] && !fields[CLV_VISCOSITY] && !fields[CLV_SOUNDSPEED]) {
    fields[CLV_PRESSURE] = 0;
    g_rec.push_back({R_P_HALO, (double)depth});
    return;
  }
  if (recorded({R_GAS, R_P_HALO, R_VISC}) && depth == 1 && fields[CLV_VISCOSITY] && !fields[CLV_PRESSURE] && !fields[CLV_SOUNDSPEED]) {
    fields[CLV_VISCOSITY] = 0;
    g_rec.push_back({R_Q_HALO, 1});
    return;
  }
  if (recorded({R_CELL}) && is_advection_halo(fields, depth)) {
    // the velocities' halo now (advec_cell does not change them), the rest recorded
    int v[CLV_NUM_HALO_FIELDS] = {0};
    v[CLV_XVEL1] = v[CLV_YVEL1] = 1;
    check(clv_update_halo(g_rec_chunk, v, 2), "update_halo");
    for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f) fields[f] = 0;
    g_rec.push_back({R_ADV_HALO, 2});
    return;
  }
  flush_deferred();  // a request the chain does not expect: the arrays first
}

bool deferred_calc_dt(global_variables &globals, int tile, double &dt, int &control, int &jldt, int &kldt, int &small) {
  if (tile != 0 || !recorded({R_GAS, R_P_HALO, R_VISC, R_Q_HALO})) {
    flush_deferred();
    return false;
  }
  const global_config &cfg = globals.config;
  const int p_depth = (int)g_rec[1].arg;
  g_rec.clear();
  // with the Lagrangian chain next (driver/hydro.cpp:59-63), pressure and soundspeed
  // are owed rather than stored: that chain rewrites both over the same cells
  const bool owe = p_depth == 1;
  check(clv_timestep_kernels(g_rec_chunk, cfg.dtmin, cfg.dtc_safe, cfg.dtu_safe, cfg.dtv_safe, cfg.dtdiv_safe, owe ? 0 : p_depth, &dt,
                             &control, &jldt, &kldt, &small),
        "timestep_kernels");
  if (owe) g_pc_owed = g_rec_chunk;
  return true;
}

// ---- the Lagrangian chain: PdV(predict), accelerate, PdV(correct), flux_calc
bool defer_pdv(global_variables &globals, bool predict) {
  if (predict) {
    if (!chain_allowed(globals)) return false;
    return start_record(globals, R_PDV_PREDICT, globals.dt);
  }
  if (!recorded({R_PDV_PREDICT, R_ACCEL}) || g_rec[0].arg != globals.dt) return false;
  g_rec.push_back({R_PDV_CORRECT, globals.dt});
  return true;
}

bool defer_accelerate(global_variables &globals) {
  if (!recorded({R_PDV_PREDICT}) || g_rec[0].arg != globals.dt) return false;
  g_rec.push_back({R_ACCEL, globals.dt});
  return true;
}

bool deferred_flux_calc(global_variables &globals) {
  if (!recorded({R_PDV_PREDICT, R_ACCEL, R_PDV_CORRECT}) || g_rec[0].arg != globals.dt) {
    flush_deferred();
    return false;
  }
  g_rec.clear();
  if (g_pc_owed && g_pc_owed != g_rec_chunk) settle_pressure();
  g_pc_owed = nullptr;  // the chain stores the predictor's pressure, its halo and soundspeed
  check(clv_lagrange_kernels(g_rec_chunk, globals.dt), "lagrange_kernels");
  return true;
}

// ---- one advection sweep: advec_cell, update_halo(density1, energy1, xvel1, yvel1, mass fluxes, 2),
// advec_mom(xvel), advec_mom(yvel)
bool defer_advec_cell(global_variables &globals, int tile, int sweep, int dir) {
  if (tile != 0 || !chain_allowed(globals)) return false;
  return start_record(globals, R_CELL, 0, sweep, dir);
}

bool defer_advec_mom(global_variables &globals, int tile, int which_vel, int dir, int sweep) {
  if (tile != 0 || g_rec.empty() || g_rec[0].op != R_CELL || g_rec[0].sweep != sweep || g_rec[0].dir != dir) return false;
  if (which_vel == 1 && recorded({R_CELL, R_ADV_HALO})) {
    g_rec.push_back({R_MOM_X, 0, sweep, dir});
    return true;
  }
  if (which_vel == 2 && recorded({R_CELL, R_ADV_HALO, R_MOM_X})) {
    if (sweep == 2) {  // reset_field comes next in driver/hydro.cpp:79-83
      g_rec.push_back({R_MOM_Y, 0, sweep, dir});
      return true;
    }
    g_rec.clear();
    check(clv_advection_pass(g_rec_chunk, sweep, dir), "advection_pass");
    refresh_buffers(globals, 0);  // density1, energy1 and the velocities are other arrays now
    return true;
  }
  return false;
}

bool deferred_reset_field(global_variables &globals, int tile) {
  if (tile != 0 || !recorded({R_CELL, R_ADV_HALO, R_MOM_X, R_MOM_Y})) return false;
  const Rec r = g_rec[0];
  g_rec.clear();
  check(clv_advection_pass_reset(g_rec_chunk, r.sweep, r.dir), "advection_pass_reset");
  refresh_buffers(globals, 0);
  return true;
}

clv_chunk *chunk(global_variables &globals, int tile) {
  flush_deferred();
  return chunk_keep(globals, tile);
}

clv_chunk *chunk_keep(global_variables &globals, int tile) {
  tile_type &t = globals.chunk.tiles[tile];
  field_type &f = t.field;
  auto it = g_chunks.find(f.density0.data);
  if (it != g_chunks.end()) return it->second;
  if (t.info.t_xmin != 1 || t.info.t_ymin != 1) unsupported("a tile whose local indices do not start at 1");
  clv_chunk_view v{};
  v.x_cells = t.info.t_xmax - t.info.t_xmin + 1;
  v.y_cells = t.info.t_ymax - t.info.t_ymin + 1;
  v.pitch = (int64_t)f.density0.pitch;
  clover::Buffer2D<double> *b2[CLV_NUM_FIELDS];
  fields_of(f, b2);
  for (int k = 0; k < CLV_NUM_FIELDS; ++k) {
    if ((int64_t)b2[k]->pitch != v.pitch) {
      std::fprintf(stderr, "mi355x model: field %d has pitch %zu, density0 %zu\n", k, b2[k]->pitch, f.density0.pitch);
      std::abort();
    }
    v.field[k] = b2[k]->data;
  }
  clover::Buffer1D<double> *b1[CLV_NUM_MESH] = {&f.cellx, &f.celly, &f.vertexx, &f.vertexy, &f.celldx, &f.celldy, &f.vertexdx, &f.vertexdy};
  for (int k = 0; k < CLV_NUM_MESH; ++k) v.mesh[k] = b1[k]->data;
  clv_chunk *c = nullptr;
  check(clv_chunk_wrap(&v, &c), "clv_chunk_wrap");
  if (const char *order = std::getenv("CLOVER_MI355X_SUMMARY"); order && std::string(order) == "serial")
    check(clv_chunk_set_summary_order(c, CLV_SUMMARY_SERIAL), "clv_chunk_set_summary_order");
  if (fused_entries(globals)) check(clv_chunk_set_entry_plan(c, CLV_ENTRY_FUSED), "clv_chunk_set_entry_plan");
  g_chunks[f.density0.data] = c;
  return c;
}

bool fused_entries(const global_variables &globals) {
  const char *entry = std::getenv("CLOVER_MI355X_ENTRY");
  return globals.config.dumpDir.empty() && !(entry && std::string(entry) == "reference");
}

void refresh_buffers(global_variables &globals, int tile) {
  clv_chunk_view v{};
  check(clv_chunk_get_view(chunk(globals, tile), &v), "clv_chunk_get_view");
  clover::Buffer2D<double> *b2[CLV_NUM_FIELDS];
  fields_of(globals.chunk.tiles[tile].field, b2);
  for (int k = 0; k < CLV_NUM_FIELDS; ++k) b2[k]->data = v.field[k];
}

void sync_if_profiling(global_variables &globals, int tile) {
  if (globals.profiler_on) check(clv_chunk_synchronize(chunk(globals, tile)), "clv_chunk_synchronize");
}

void release_chunks() {
  flush_deferred();
  for (auto &kv : g_chunks) clv_chunk_destroy(kv.second);
  g_chunks.clear();
}

}  // namespace mi355x

void mi355x_flush_deferred() { mi355x::flush_deferred(); }
