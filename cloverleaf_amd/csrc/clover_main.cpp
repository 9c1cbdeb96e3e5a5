// clover_mi355x: the host driver over the C ABI.  Runs one deck on one GPU with
// the step sequence of driver/hydro.cpp and prints what the reference prints:
// the per-step dt line (driver/timestep.cpp:98-103), the field summary table
// (driver/report.cpp:48-71, to the --out file), the known-answer check
// (driver/report.cpp:73-111) and the grind-time lines (driver/hydro.cpp:208-221).
//
// usage: clover_mi355x [--file|--in DECK] [--out FILE] [--plan reference|fused] [--profile] [--quiet]
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "clover_mi355x.h"

namespace {

double wall() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec * 1.0e-6;
}

void die(const char *what, clv_status st) {
  std::cerr << "clover_mi355x: " << what << " failed (" << st << "): " << clv_last_error() << std::endl;
  std::exit(st == CLV_ERR_SMALL_DT ? 3 : 1);
}

// known answers of driver/report.cpp:77-92
bool expected_ke(int problem, double &ke) {
  switch (problem) {
    case 1: ke = 1.82280367310258; return true;
    case 2: ke = 1.19316898756307; return true;
    case 3: ke = 2.58984003503994; return true;
    case 4: ke = 0.307475452287895; return true;
    case 5: ke = 4.85350315783719; return true;
    case 487: ke = 6.088288e-01; return true;
    case 287: ke = 6.062609e-01; return true;
    case 168: ke = 2.465082e-02; return true;
  }
  return false;
}

void report_step(std::ostream &out, const clv_hydro_state &st, const double s[5]) {
  const double vol = s[0], mass = s[1], ie = s[2], ke = s[3], press = s[4];
  out << std::endl
      << "Time " << st.time << std::endl
      << "                "
      << "Volume          Mass            Density         Pressure        Internal Energy Kinetic Energy  Total Energy    " << std::endl;
  auto f = out.flags();
  out << " step: " << st.step << std::scientific << std::setw(15) << vol << std::setw(15) << mass << std::setw(15) << mass / vol
      << std::setw(15) << press / vol << std::setw(15) << ie << std::setw(15) << ke << std::setw(15) << ie + ke << std::endl
      << std::endl;
  out.flags(f);
}

}  // namespace

int main(int argc, char **argv) {
  std::string in = "clover.in", outname = "clover.out";
  int plan = CLV_PLAN_FUSED;
  bool profile = false, quiet = false;
  for (int a = 1; a < argc; ++a) {
    std::string arg = argv[a];
    auto need = [&](const char *msg) -> std::string {
      if (a + 1 >= argc) {
        std::cerr << msg << std::endl;
        std::exit(EXIT_FAILURE);
      }
      return argv[++a];
    };
    if (arg == "--in" || arg == "--file") in = need("--in,--file specified but no path was given");
    else if (arg == "--out") outname = need("--out specified but no path was given");
    else if (arg == "--plan") {
      std::string p = need("--plan specified but no plan was given");
      if (p == "reference") plan = CLV_PLAN_REFERENCE;
      else if (p == "fused") plan = CLV_PLAN_FUSED;
      else {
        std::cerr << "Illegal --plan option: " << p << std::endl;
        return EXIT_FAILURE;
      }
    } else if (arg == "--profile") profile = true;
    else if (arg == "--quiet") quiet = true;
    else if (arg == "-h" || arg == "--help") {
      std::cout << "Usage: clover_mi355x [--file|--in DECK] [--out FILE] [--plan reference|fused] [--profile] [--quiet]" << std::endl;
      return EXIT_SUCCESS;
    } else {
      std::cerr << "Unknown argument: " << arg << std::endl;
      return EXIT_FAILURE;
    }
  }
  std::ifstream fin(in);
  if (fin.fail()) {
    std::cerr << "Unable to open file: `" << in << "`" << std::endl;
    return 1;
  }
  std::stringstream ss;
  ss << fin.rdbuf();
  clv_config cfg;
  clv_status st = clv_deck_parse(ss.str().c_str(), &cfg);
  if (st != CLV_OK) die("read_input", st);
  if (cfg.tiles_per_chunk != 1) std::cerr << "# note: tiles_per_chunk " << cfg.tiles_per_chunk << " ignored, one tile per GPU" << std::endl;
  std::ofstream g_out(outname);
  g_out << "Clover Version 2.000 (clover_mi355x, gfx950)" << std::endl << "Task Count 1" << std::endl << std::endl;
  std::cout << " Output file " << outname << " opened. All output will go there." << std::endl;

  clv_chunk *c = nullptr;
  if ((st = clv_chunk_create(cfg.x_cells, cfg.y_cells, &c)) != CLV_OK) die("clv_chunk_create", st);
  if (profile && (st = clv_profile_enable(c, 1)) != CLV_OK) die("clv_profile_enable", st);
  clv_hydro_state hs;
  double s[5];
  if ((st = clv_hydro_start(c, &cfg, &hs, s)) != CLV_OK) die("start", st);
  std::cout << " Problem initialised and generated" << std::endl;
  report_step(g_out, hs, s);

  const double timerstart = wall();
  double first_step = 0, second_step = 0;
  const char *control_names[] = {"", "sound", "xvel", "yvel", "div"};
  while (true) {
    const double step_time = wall();
    const double t_before = hs.time;
    if ((st = clv_hydro_step(c, &cfg, &hs, plan)) != CLV_OK) die("hydro step", st);
    const char *ctl = (hs.dt_control >= 1 && hs.dt_control <= 4) ? control_names[hs.dt_control] : "";
    // the reference prints the time before this step's increment (driver/timestep.cpp:99)
    if (!quiet) {
      std::cout << " Step " << hs.step << " time " << t_before << " control " << ctl << " timestep  " << hs.dt << " " << hs.jdt
                << "," << hs.kdt << " x 0 y 0" << std::endl;
    }
    g_out << " Step " << hs.step << " time " << t_before << " control " << ctl << " timestep  " << hs.dt << " " << hs.jdt << ","
          << hs.kdt << " x 0 y 0" << std::endl;
    if (cfg.summary_frequency != 0 && hs.step % cfg.summary_frequency == 0) {
      if ((st = clv_hydro_summary(c, s)) != CLV_OK) die("field_summary", st);
      report_step(g_out, hs, s);
    }
    if (hs.step == 1) first_step = wall() - step_time;
    if (hs.step == 2) second_step = wall() - step_time;
    if (hs.complete) {
      if ((st = clv_hydro_summary(c, s)) != CLV_OK) die("field_summary", st);
      report_step(g_out, hs, s);
      const double wall_clock = wall() - timerstart;
      int rc = 0;
      double ke_exp;
      if (cfg.test_problem >= 1) {
        double qa_diff = 100;
        if (expected_ke(cfg.test_problem, ke_exp)) qa_diff = std::fabs((100.0 * (s[3] / ke_exp)) - 100.0);
        else std::cout << " WARNING: Unknown test problem " << cfg.test_problem << ", validation will fail" << std::endl;
        std::cout << " Test problem " << cfg.test_problem << " is within " << qa_diff << "% of the expected solution" << std::endl;
        g_out << "Test problem " << cfg.test_problem << " is within " << qa_diff << "% of the expected solution" << std::endl;
        const bool pass = !std::isnan(qa_diff) && qa_diff < 0.001;
        std::cout << (pass ? " This test is considered PASSED" : " This test is considered NOT PASSED") << std::endl;
        g_out << (pass ? "This test is considered PASSED" : "This test is considered NOT PASSED") << std::endl;
        rc = pass ? 0 : 1;
      }
      std::cout << " Wall clock " << wall_clock << std::endl << " First step overhead " << first_step - second_step << std::endl;
      g_out << std::endl << "Calculation complete" << std::endl << "Clover is finishing" << std::endl << "Wall clock " << wall_clock << std::endl;
      const double cells = (double)cfg.x_cells * cfg.y_cells;
      std::cout << " Average time per cell " << wall_clock / (hs.step * cells) << std::endl;
      std::cout << " Cell-steps per second " << (hs.step * cells) / wall_clock << std::endl;
      if (profile) {
        const char *names[64];
        int64_t launches[64];
        double ms[64];
        const int n = clv_profile_read(c, names, launches, ms, 64);
        std::cout << " Kernel profile (HIP events):" << std::endl;
        for (int k = 0; k < n && k < 64; ++k)
          std::cout << "  " << std::left << std::setw(24) << names[k] << std::right << std::setw(8) << launches[k] << std::setw(14)
                    << ms[k] << " ms" << std::endl;
      }
      clv_chunk_destroy(c);
      return rc;
    }
  }
}
