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

void clover_report_step(global_variables &globals, parallel_ &, double vol, double mass, double ie, double ke, double press) {
  std::fprintf(g_json, "%s\n  {\"step\": %d, \"time\": %.17e, \"dt\": %.17e, \"vol\": %.17e, \"mass\": %.17e, \"ie\": %.17e, \"ke\": %.17e, \"press\": %.17e}",
               g_first ? "" : ",", globals.step, globals.time, globals.dt, vol, mass, ie, ke, press);
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
