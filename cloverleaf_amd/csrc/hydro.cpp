// The host side of one hydro run: the deck reader (driver/read_input.cpp) and the
// step sequencing of driver/start.cpp, driver/timestep.cpp, driver/hydro.cpp,
// driver/advection.cpp and the predictor chain of src/serial/PdV.cpp:113-132,
// over the chunk's kernels.  Two kernel plans give bit-identical fields:
// CLV_PLAN_REFERENCE issues one launch per reference loop (what the plugin
// entry points do), CLV_PLAN_FUSED the fused sweeps of kernels_fused.hip.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "clv_internal.h"

namespace clv {
clv_status fused_timestep_kernels(clv_chunk *c, const clv_config *cfg, double *dt_min_val);
clv_status fused_lagrange(clv_chunk *c, double dt);
clv_status fused_advection(clv_chunk *c, int advect_x);
}  // namespace clv

using namespace clv;

namespace {

void mask(int fields[CLV_NUM_HALO_FIELDS], std::initializer_list<int> ids) {
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f) fields[f] = 0;
  for (int id : ids) fields[id] = 1;
}

#define TRY(x)                        \
  do {                                \
    clv_status st_ = (x);             \
    if (st_ != CLV_OK) return st_;    \
  } while (0)

// driver/advection.cpp:29-105
clv_status advection_reference(clv_chunk *c, int advect_x) {
  int fields[CLV_NUM_HALO_FIELDS];
  int direction = advect_x ? CLV_XDIR : CLV_YDIR;
  mask(fields, {CLV_ENERGY1, CLV_DENSITY1, CLV_VOL_FLUX_X, CLV_VOL_FLUX_Y});
  TRY(k_update_halo(c, fields, 2));
  TRY(k_advec_cell(c, 1, direction));
  mask(fields, {CLV_DENSITY1, CLV_ENERGY1, CLV_XVEL1, CLV_YVEL1, CLV_MASS_FLUX_X, CLV_MASS_FLUX_Y});
  TRY(k_update_halo(c, fields, 2));
  TRY(k_advec_mom(c, 1, direction, 1));
  TRY(k_advec_mom(c, 2, direction, 1));
  direction = advect_x ? CLV_YDIR : CLV_XDIR;
  TRY(k_advec_cell(c, 2, direction));
  TRY(k_update_halo(c, fields, 2));
  TRY(k_advec_mom(c, 1, direction, 2));
  TRY(k_advec_mom(c, 2, direction, 2));
  return CLV_OK;
}

}  // namespace

extern "C" {

// driver/read_input.cpp:37-277
clv_status clv_deck_parse(const char *text, clv_config *g) {
  if (!text || !g) return set_error(CLV_ERR_ARG, "clv_deck_parse: null argument");
  std::memset(g, 0, sizeof(*g));
  // defaults, :39-69
  g->test_problem = 0;
  g->xmin = g->ymin = g->xmax = g->ymax = 0.0;
  g->x_cells = g->y_cells = 10;
  g->end_time = 10.0;
  g->end_step = 640000;  // g_ibig, driver/definitions.h:31
  g->visit_frequency = 0;
  g->summary_frequency = 10;
  g->tiles_per_chunk = 1;
  g->dtinit = 0.1;
  g->dtmax = 1.0;
  g->dtmin = 0.0000001;
  g->dtrise = 1.5;
  g->dtc_safe = 0.7;
  g->dtu_safe = 0.5;
  g->dtv_safe = 0.5;
  g->dtdiv_safe = 0.7;
  g->profiler_on = 0;

  std::vector<std::string> lines;
  {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) lines.push_back(line);
  }
  // :94-106, count "state" lines
  int state_max = 0;
  for (const auto &line : lines) {
    std::istringstream iss(line);
    std::string w0, w1;
    if (!(iss >> w0)) continue;
    if (w0 == "state" && (iss >> w1)) state_max = std::max(state_max, std::atoi(w1.c_str()));
  }
  if (state_max < 1) return set_error(CLV_ERR_DECK, "read_input: No states defined.");
  if (state_max > CLV_MAX_STATES) return set_error(CLV_ERR_DECK, "read_input: more states than CLV_MAX_STATES");
  g->number_of_states = state_max;
  std::vector<bool> defined(state_max, false);
  // :125-256, tokens split on ' ' and '='
  for (const auto &line : lines) {
    if (line.empty()) continue;
    std::vector<std::string> w;
    {
      std::string cur;
      for (char ch : line) {
        if (ch == ' ' || ch == '=') {
          if (!cur.empty()) w.push_back(cur);
          cur.clear();
        } else {
          cur.push_back(ch);
        }
      }
      if (!cur.empty()) w.push_back(cur);
    }
    if (w.empty()) break;
    auto num = [&](size_t k) -> double { return k < w.size() ? std::strtod(w[k].c_str(), nullptr) : 0.0; };
    auto inum = [&](size_t k) -> int { return k < w.size() ? std::atoi(w[k].c_str()) : 0; };
    const std::string &k = w[0];
    if (k == "initial_timestep") g->dtinit = num(1);
    else if (k == "max_timestep") g->dtmax = num(1);
    else if (k == "timestep_rise") g->dtrise = num(1);
    else if (k == "end_time") g->end_time = num(1);
    else if (k == "end_step") g->end_step = inum(1);
    else if (k == "xmin") g->xmin = num(1);
    else if (k == "xmax") g->xmax = num(1);
    else if (k == "ymin") g->ymin = num(1);
    else if (k == "ymax") g->ymax = num(1);
    else if (k == "x_cells") g->x_cells = inum(1);
    else if (k == "y_cells") g->y_cells = inum(1);
    else if (k == "visit_frequency") g->visit_frequency = inum(1);
    else if (k == "summary_frequency") g->summary_frequency = inum(1);
    else if (k == "tiles_per_chunk") g->tiles_per_chunk = inum(1);
    else if (k == "tiles_per_problem") g->tiles_per_chunk = inum(1);  // one task
    else if (k == "profiler_on") g->profiler_on = 1;
    else if (k == "test_problem") g->test_problem = inum(1);
    else if (k == "state") {
      const int s = inum(1) - 1;
      if (s < 0 || s >= state_max) return set_error(CLV_ERR_DECK, "read_input: bad state number");
      if (defined[s]) return set_error(CLV_ERR_DECK, "read_input: State defined twice.");
      defined[s] = true;
      clv_state &st = g->states[s];
      for (size_t iw = 2; iw < w.size(); ++iw) {
        const std::string &t = w[iw];
        if (t == "xvel") st.xvel = num(++iw);
        else if (t == "yvel") st.yvel = num(++iw);
        else if (t == "xmin") st.xmin = num(++iw);
        else if (t == "ymin") st.ymin = num(++iw);
        else if (t == "xmax") st.xmax = num(++iw);
        else if (t == "ymax") st.ymax = num(++iw);
        else if (t == "radius") st.radius = num(++iw);
        else if (t == "density") st.density = num(++iw);
        else if (t == "energy") st.energy = num(++iw);
        else if (t == "geometry") {
          const std::string geo = ++iw < w.size() ? w[iw] : "";
          if (geo == "rectangle") st.geometry = 1;
          else if (geo == "circle") st.geometry = 2;
          else if (geo == "point") st.geometry = 3;
        }
      }
    }
  }
  if (g->x_cells < 2 || g->y_cells < 2) return set_error(CLV_ERR_DECK, "read_input: x_cells and y_cells must be >= 2");
  // :262-276, the 1/100-cell shrink (with the reference's (float) cast of the cell count)
  double dx = (g->xmax - g->xmin) / (float)g->x_cells;
  double dy = (g->ymax - g->ymin) / (float)g->y_cells;
  for (int n = 1; n < g->number_of_states; ++n) {
    g->states[n].xmin += dx / 100.0;
    g->states[n].ymin += dy / 100.0;
    g->states[n].xmax -= dx / 100.0;
    g->states[n].ymax -= dy / 100.0;
  }
  return CLV_OK;
}

clv_status clv_hydro_summary(clv_chunk *c, double out[5]) {
  if (!c || !out) return set_error(CLV_ERR_ARG, "clv_hydro_summary: null argument");
  return k_field_summary_gas(c, out);
}

// driver/start.cpp:69-127 and driver/definitions.cpp:53-55
clv_status clv_hydro_start(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, double summary0[5]) {
  if (!c || !cfg || !st) return set_error(CLV_ERR_ARG, "clv_hydro_start: null argument");
  if (cfg->x_cells != c->d.nx || cfg->y_cells != c->d.ny) return set_error(CLV_ERR_ARG, "clv_hydro_start: config does not match chunk");
  TRY(k_fill_zero(c));
  TRY(k_initialise_chunk(c, cfg, 1, 1));
  TRY(k_generate_chunk(c, cfg));
  TRY(k_ideal_gas(c, 0));
  int fields[CLV_NUM_HALO_FIELDS];
  mask(fields, {CLV_DENSITY0, CLV_ENERGY0, CLV_PRESSURE, CLV_VISCOSITY, CLV_DENSITY1, CLV_ENERGY1, CLV_XVEL0, CLV_YVEL0, CLV_XVEL1, CLV_YVEL1});
  TRY(k_update_halo(c, fields, 2));
  std::memset(st, 0, sizeof(*st));
  st->step = 0;
  st->advect_x = 1;
  st->time = 0.0;
  st->dt = cfg->dtinit;
  st->dtold = cfg->dtinit;
  if (summary0) TRY(clv_hydro_summary(c, summary0));
  return hip_check(hipStreamSynchronize(c->stream), "clv_hydro_start");
}

// driver/timestep.cpp:37-108
clv_status clv_timestep(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, int plan) {
  if (!c || !cfg || !st) return set_error(CLV_ERR_ARG, "clv_timestep: null argument");
  st->dt = g_big;
  double dtlp = 0.0;
  if (plan == CLV_PLAN_FUSED) {
    TRY(fused_timestep_kernels(c, cfg, &dtlp));
  } else {
    int fields[CLV_NUM_HALO_FIELDS];
    TRY(k_ideal_gas(c, 0));
    mask(fields, {CLV_PRESSURE, CLV_ENERGY0, CLV_DENSITY0, CLV_XVEL0, CLV_YVEL0});
    TRY(k_update_halo(c, fields, 1));
    TRY(k_viscosity(c));
    mask(fields, {CLV_VISCOSITY});
    TRY(k_update_halo(c, fields, 1));
    TRY(k_calc_dt(c, cfg->dtc_safe, cfg->dtu_safe, cfg->dtv_safe, cfg->dtdiv_safe, &dtlp));
  }
  int dt_control = 0, jldt = 0, kldt = 0;
  dt_location(c->d.nx, &dt_control, &jldt, &kldt);
  if (dtlp <= st->dt) {
    st->dt = dtlp;
    st->dt_control = dt_control;
    st->jdt = jldt;
    st->kdt = kldt;
  }
  // std::min(std::min(dt, dtold * dtrise), dtmax)
  const double a = st->dt, b = st->dtold * cfg->dtrise;
  const double m = (b < a) ? b : a;
  st->dt = (cfg->dtmax < m) ? cfg->dtmax : m;
  const bool small = st->dt < cfg->dtmin;
  st->dtold = st->dt;
  if (small) return set_error(CLV_ERR_SMALL_DT, "timestep: small timestep");
  return CLV_OK;
}

// One phase of driver/hydro.cpp:57-89 on the reference plan; CLV_PHASE_END_STEP is the
// bookkeeping after reset_field: flip advect_x, advance time, set complete (:86-89, :108).
clv_status clv_hydro_phase(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, int phase) {
  if (!c || !cfg || !st) return set_error(CLV_ERR_ARG, "clv_hydro_phase: null argument");
  const double dt = st->dt;
  switch (phase) {
    case CLV_PHASE_TIMESTEP:
      st->step += 1;
      return clv_timestep(c, cfg, st, CLV_PLAN_REFERENCE);
    case CLV_PHASE_PDV_PREDICT: return clv_PdV_driver(c, 1, dt);
    case CLV_PHASE_ACCELERATE: return k_accelerate(c, dt);
    case CLV_PHASE_PDV_CORRECT: return clv_PdV_driver(c, 0, dt);
    case CLV_PHASE_FLUX_CALC: return k_flux_calc(c, dt);
    case CLV_PHASE_ADVECTION: return advection_reference(c, st->advect_x);
    case CLV_PHASE_RESET_FIELD: return k_reset_field(c);
    case CLV_PHASE_END_STEP:
      st->advect_x = !st->advect_x;
      st->time += dt;
      st->complete = (st->time + g_small > cfg->end_time || st->step >= cfg->end_step) ? 1 : 0;
      return CLV_OK;
  }
  return set_error(CLV_ERR_ARG, "clv_hydro_phase: phase is one of CLV_PHASE_*");
}

// driver/hydro.cpp:57-89 and :108
clv_status clv_hydro_step(clv_chunk *c, const clv_config *cfg, clv_hydro_state *st, int plan) {
  if (!c || !cfg || !st) return set_error(CLV_ERR_ARG, "clv_hydro_step: null argument");
  if (plan != CLV_PLAN_FUSED) {
    for (int phase = CLV_PHASE_TIMESTEP; phase <= CLV_PHASE_END_STEP; ++phase) TRY(clv_hydro_phase(c, cfg, st, phase));
    return CLV_OK;
  }
  st->step += 1;
  TRY(clv_timestep(c, cfg, st, plan));
  const double dt = st->dt;
  TRY(fused_lagrange(c, dt));
  TRY(fused_advection(c, st->advect_x));
  st->advect_x = !st->advect_x;
  st->time += dt;
  st->complete = (st->time + g_small > cfg->end_time || st->step >= cfg->end_step) ? 1 : 0;
  return CLV_OK;
}

}  // extern "C"
