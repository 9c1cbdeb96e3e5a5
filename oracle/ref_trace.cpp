// Full-precision trace of the reference serial model -- TEST INFRASTRUCTURE ONLY.
//
// Links the reference's own driver (driver/*.cpp except clover_leaf.cpp and
// report.cpp) and serial model (src/serial/*.cpp), compiled from their sources
// by oracle/Makefile.  This file supplies main() and the three report
// functions of driver/report.h so that every field_summary (driver/hydro.cpp:92-94
// and :111) is recorded with 17 significant digits instead of the reference's 6.
// The summary frequency is forced to 1, which is numerically neutral: the
// ideal_gas call inside field_summary (src/serial/field_summary.cpp:52)
// recomputes exactly what the next timestep's ideal_gas (driver/timestep.cpp:48)
// overwrites, and no halo is touched.
//
// Besides the reference's own five sums (whose value depends on the model's
// summation order: left to right in serial, per-thread SIMD partials in omp),
// every record carries x_vol .. x_press: the same per-cell terms, formed with the
// reference's expressions (src/serial/field_summary.cpp:76-90), summed in
// double-double (TwoSum error-free transformation, Knuth TAOCP vol. 2, 4.2.2)
// and rounded once.  For these non-negative terms that is the exact sum of the
// reference's own state to within one rounding, independent of any order, so a
// reproduction of the state can be compared with it far below 1e-11.
//
// usage: ref_trace DECK OUT.json
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "comms.h"
#include "definitions.h"
#include "hydro.h"
#include "read_input.h"
#include "report.h"
#include "start.h"

std::ostream g_out(nullptr);

static FILE *g_json = nullptr;
static bool g_first = true;

void report_error(char *location, char *error) {
  std::fprintf(stderr, "reference error in %s: %s\n", location, error);
  if (g_json) {
    std::fprintf(g_json, "\n], \"error\": \"%s: %s\"}\n", location, error);
    std::fclose(g_json);
  }
  std::exit(3);
}

void clover_report_step_header(global_variables &, parallel_ &) {}

namespace {
// double-double accumulator: hi + lo carries the running sum with no rounding
// beyond ~2^-104 relative per addition
struct dd {
  double hi = 0.0, lo = 0.0;
  void add(double x) {
    const double s = hi + x;
    const double bb = s - hi;
    const double err = (hi - (s - bb)) + (x - bb);
    hi = s;
    lo += err;
  }
  void add(const dd &o) {
    add(o.hi);
    add(o.lo);
  }
  double value() const { return hi + lo; }
};

// the five sums of field_summary over the interior cells of every tile, each
// per-cell term as src/serial/field_summary.cpp:76-90 forms it
void exact_sums(global_variables &globals, double out[5]) {
  dd tot[5];
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    tile_type &t = globals.chunk.tiles[tile];
    const int xmin = t.info.t_xmin, xmax = t.info.t_xmax, ymin = t.info.t_ymin, ymax = t.info.t_ymax;
    field_type &field = t.field;
    const int rows = ymax - ymin + 1;
    std::vector<dd> part(5 * (size_t)rows);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; ++r) {
      const int k = ymin + 1 + r;
      dd *p = &part[5 * (size_t)r];
      for (int j = xmin + 1; j <= xmax + 1; ++j) {
        double vsqrd = 0.0;
        for (int kv = k; kv <= k + 1; ++kv) {
          for (int jv = j; jv <= j + 1; ++jv) {
            vsqrd += 0.25 * (field.xvel0(jv, kv) * field.xvel0(jv, kv) + field.yvel0(jv, kv) * field.yvel0(jv, kv));
          }
        }
        const double cell_vol = field.volume(j, k);
        const double cell_mass = cell_vol * field.density0(j, k);
        p[0].add(cell_vol);
        p[1].add(cell_mass);
        p[2].add(cell_mass * field.energy0(j, k));
        p[3].add(cell_mass * 0.5 * vsqrd);
        p[4].add(cell_vol * field.pressure(j, k));
      }
    }
    for (int r = 0; r < rows; ++r)
      for (int q = 0; q < 5; ++q) tot[q].add(part[5 * (size_t)r + q]);
  }
  for (int q = 0; q < 5; ++q) out[q] = tot[q].value();
}
} // namespace

void clover_report_step(global_variables &globals, parallel_ &, double vol, double mass, double ie, double ke, double press) {
  double x[5];
  exact_sums(globals, x);
  std::fprintf(g_json,
               "%s\n  {\"step\": %d, \"time\": %.17e, \"dt\": %.17e, \"vol\": %.17e, \"mass\": %.17e, \"ie\": %.17e, \"ke\": %.17e, "
               "\"press\": %.17e, \"x_vol\": %.17e, \"x_mass\": %.17e, \"x_ie\": %.17e, \"x_ke\": %.17e, \"x_press\": %.17e}",
               g_first ? "" : ",", globals.step, globals.time, globals.dt, vol, mass, ie, ke, press, x[0], x[1], x[2], x[3], x[4]);
  std::fflush(g_json);
  g_first = false;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s DECK OUT.json\n", argv[0]);
    return 2;
  }
  g_out.rdbuf(nullptr);
  parallel_ parallel;
  global_config config;
  std::ifstream g_in(argv[1]);
  if (g_in.fail()) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 2;
  }
  read_input(g_in, parallel, config);
  config.number_of_chunks = parallel.max_task;
  config.summary_frequency = 1;
  config.profiler_on = false;
  g_json = std::fopen(argv[2], "w");
  std::fprintf(g_json, "{\"deck\": \"%s\", \"x_cells\": %d, \"y_cells\": %d, \"summaries\": [", argv[1], config.grid.x_cells,
               config.grid.y_cells);
  clover::context ctx{};
  auto globals = start(parallel, config, ctx);
  hydro(globals, parallel);
  std::fprintf(g_json, "\n]}\n");
  std::fclose(g_json);
  return 0;
}
