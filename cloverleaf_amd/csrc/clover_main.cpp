// clover_mi355x: the host driver over the C ABI.  Runs one deck on one GPU with
// the step sequence of driver/hydro.cpp and prints what the reference's driver
// prints, line for line:
//   stdout   the YAML run header (driver/clover_leaf.cpp:56-117, the device list of
//            driver/initialise.h:86-91), the input/launch lines (:136-206), the
//            per-step dt line (driver/timestep.cpp:98-103) and grind lines
//            (driver/hydro.cpp:205-218), the completion lines (:108-118), the
//            known-answer check (driver/report.cpp:73-111), the profiler table
//            (driver/hydro.cpp:125-200) and the result block (driver/clover_leaf.cpp:211-216);
//   --out    the version preamble (driver/clover_leaf.cpp:127-131), the deck echo of
//            driver/read_input.cpp:86-262, the set-up lines of driver/start.cpp and
//            driver/comms.cpp:148-153, the field summary table (driver/report.cpp:48-71)
//            and the same step, grind and completion lines;
//   --dump   the field dumps of driver/shared.cpp:96-189 after every phase of start
//            (driver/start.cpp:80-129) and of every step (driver/hydro.cpp:50-85).
// Model-specific lines are the device list and the Model block.  Timings are this
// run's.  --dump and decks under 4 cells a side run the reference plan (one launch
// per reference loop, which has the state between phases); otherwise --plan chooses.
//
// usage: clover_mi355x [--file|--in DECK] [--out FILE] [--dump DIR] [--profile]
//                      [--plan reference|fused] [--geometry arrays|mesh] [--list]
//                      [--device INDEX] [--staging-buffer true|false|auto] [--quiet]
#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "clover_mi355x.h"

namespace {

const char *kVersion = "2.000";  // driver/version.h:22

double wall() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec * 1.0e-6;
}

void die(const char *what, clv_status st) {
  std::cerr << "clover_mi355x: " << what << " failed (" << st << "): " << clv_last_error() << std::endl;
  std::exit(st == CLV_ERR_SMALL_DT ? 3 : 1);
}

#define CHECK(call, what)                   \
  do {                                      \
    const clv_status st_ = (call);          \
    if (st_ != CLV_OK) die((what), st_);    \
  } while (0)

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

struct Args {
  std::string in = "clover.in", out = "clover.out", dump;
  int plan = CLV_PLAN_FUSED;
  int geometry = CLV_GEOMETRY_ARRAYS;
  int profile = -1;  // -1: as the deck says
  int device = 0;
  bool quiet = false;
  std::vector<std::string> raw;
};

void help() {
  std::cout << std::endl
            << "Usage: clover_mi355x [OPTIONS]\n\n"
            << "Options:\n"
            << "  -h  --help                             Print this message\n"
            << "      --list                             List available devices with index and exit\n"
            << "      --device                <INDEX>    Use device at INDEX from output of --list\n"
            << "      --file,--in              <FILE>    Custom clover.in file FILE (defaults to clover.in if unspecified)\n"
            << "      --out                    <FILE>    Custom clover.out file FILE (defaults to clover.out if unspecified)\n"
            << "      --dump                    <DIR>    Dumps all field data in ASCII to ./DIR for debugging, DIR is created if missing\n"
            << "      --profile                          Enables kernel profiling, this takes precedence over the profiler_on in clover.in\n"
            << "      --staging-buffer <true|false|auto> Accepted for compatibility; one GPU exchanges no halos over MPI\n"
            << "      --plan      <reference|fused>      Kernel plan (default fused; --dump and decks under 4 cells use reference)\n"
            << "      --geometry      <arrays|mesh>      Volume and face areas read from their arrays (default) or computed\n"
            << "                                         from the 1D mesh spacings\n"
            << "      --quiet                            No per-step lines on stdout\n"
            << std::endl;
}

void list_devices() {
  std::cout << "Devices:" << std::endl;
  const int n = clv_device_count();
  for (int d = 0; d < (n > 0 ? n : 1); ++d) std::cout << " " << d << ": HIP device " << d << " (gfx950, clover_mi355x)" << std::endl;
}

Args parse_args(int argc, char **argv) {
  Args a;
  for (int k = 1; k < argc; ++k) a.raw.emplace_back(argv[k]);
  for (size_t k = 0; k < a.raw.size(); ++k) {
    const std::string &arg = a.raw[k];
    auto param = [&](const char *msg) -> std::string {
      if (k + 1 < a.raw.size()) return a.raw[++k];
      std::cerr << msg << std::endl;
      help();
      std::exit(EXIT_FAILURE);
    };
    if (arg == "--help" || arg == "-h") {
      help();
      std::exit(EXIT_SUCCESS);
    } else if (arg == "--dump") {
      a.dump = param("--dump specified but no dir was given");
    } else if (arg == "--list") {
      list_devices();
      std::exit(EXIT_SUCCESS);
    } else if (arg == "--profile") {
      a.profile = 1;
    } else if (arg == "--device") {
      const std::string p = param("--device specified but no size was given");
      char *end = nullptr;
      const long d = std::strtol(p.c_str(), &end, 10);
      if (end == p.c_str() || *end || d < 0 || d >= std::max(1, clv_device_count())) {
        std::cerr << "No matching devices, all devices:" << std::endl;
        list_devices();
        std::exit(EXIT_FAILURE);
      }
      a.device = (int)d;
    } else if (arg == "--in" || arg == "--file") {
      a.in = param("--in,--file specified but no path was given");
    } else if (arg == "--out") {
      a.out = param("--out specified but no path was given");
    } else if (arg == "--staging-buffer") {
      const std::string p = param("--staging-buffer specified but no option given, expecting <true|false|auto>");
      if (p != "true" && p != "false" && p != "auto") {
        std::cerr << "Illegal --staging-buffer option:" << p << std::endl;
        std::exit(EXIT_FAILURE);
      }
    } else if (arg == "--plan") {
      const std::string p = param("--plan specified but no plan was given");
      if (p == "reference") a.plan = CLV_PLAN_REFERENCE;
      else if (p == "fused") a.plan = CLV_PLAN_FUSED;
      else {
        std::cerr << "Illegal --plan option: " << p << std::endl;
        std::exit(EXIT_FAILURE);
      }
    } else if (arg == "--geometry") {
      const std::string p = param("--geometry specified but no mode was given");
      if (p == "arrays") a.geometry = CLV_GEOMETRY_ARRAYS;
      else if (p == "mesh") a.geometry = CLV_GEOMETRY_MESH;
      else {
        std::cerr << "Illegal --geometry option: " << p << std::endl;
        std::exit(EXIT_FAILURE);
      }
    } else if (arg == "--quiet") {
      a.quiet = true;
    } else {
      std::cerr << "Unknown argument: " << arg << std::endl;
      help();
      std::exit(EXIT_FAILURE);
    }
  }
  return a;
}

// The deck echo of driver/read_input.cpp:86-262: each line split on ' ' and '=',
// the keywords it knows printed with their parsed values (default stream format),
// one blank line after every non-empty line.  The values come from the same
// std::stod / std::stoi the reference applies.
void echo_deck(std::ostream &g_out, const std::string &text) {
  g_out << "Reading input file" << std::endl << std::endl;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    std::vector<std::string> w;
    std::string local = line;
    for (char *t = std::strtok(local.data(), " ="); t; t = std::strtok(nullptr, " =")) w.emplace_back(t);
    if (w.empty() || w[0].empty()) break;
    const std::string &k = w[0];
    auto num = [&](size_t i) { return i < w.size() ? std::stod(w[i]) : 0.0; };
    auto inum = [&](size_t i) { return i < w.size() ? std::stoi(w[i]) : 0; };
    static const char *dkeys[] = {"initial_timestep", "max_timestep", "timestep_rise", "end_time", "xmin", "xmax", "ymin", "ymax"};
    static const char *ikeys[] = {"end_step", "x_cells", "y_cells", "visit_frequency", "summary_frequency", "tiles_per_chunk"};
    bool done = false;
    for (const char *dk : dkeys)
      if (k == dk) {
        g_out << " " << k << " " << num(1) << std::endl;
        done = true;
      }
    for (const char *ik : ikeys)
      if (!done && k == ik) {
        g_out << " " << k << " " << inum(1) << std::endl;
        done = true;
      }
    if (done) {
    } else if (k == "tiles_per_problem") {
      g_out << " tiles_per_chunk " << inum(1) << std::endl;  // / max_task = 1
    } else if (k == "profiler_on") {
      g_out << " Profiler on" << std::endl;
    } else if (k == "test_problem") {
      g_out << " test_problem " << inum(1);
    } else if (k == "state") {
      g_out << "Reading specification for state " << inum(1) << std::endl << std::endl;
      for (size_t iw = 2; iw < w.size(); ++iw) {
        const std::string &sk = w[iw];
        if (sk.empty()) break;
        if (sk == "xvel" || sk == "yvel" || sk == "xmin" || sk == "ymin" || sk == "xmax" || sk == "ymax" || sk == "radius" ||
            sk == "density" || sk == "energy") {
          ++iw;
          g_out << " " << sk << " " << num(iw) << std::endl;
        } else if (sk == "geometry") {
          ++iw;
          const std::string g = iw < w.size() ? w[iw] : "";
          if (g == "rectangle") g_out << " state geometry rectangular" << std::endl;
          else if (g == "circle") g_out << " state geometry circular" << std::endl;
          else if (g == "point") g_out << " state geometry point" << std::endl;
        }
      }
    }
    g_out << std::endl;
  }
  g_out << std::endl << std::endl << "Input read finished." << std::endl << std::endl;
}

// driver/report.cpp:48-71
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

// ---- --dump (driver/shared.cpp:96-189) ------------------------------------------------------
// Every array at the reference's extent (driver/definitions.cpp:4-41), element (i, j)
// with i outermost, as the serial model's mirrored2() view orders it.
struct DumpArray {
  const char *name;
  int field;
  int xext, yext;  // extent beyond x_cells + 4 / y_cells + 4
};
const DumpArray kDump2D[] = {
    {"density0", CLV_DENSITY0, 0, 0},       {"density1", CLV_DENSITY1, 0, 0},       {"energy0", CLV_ENERGY0, 0, 0},
    {"energy1", CLV_ENERGY1, 0, 0},         {"pressure", CLV_PRESSURE, 0, 0},       {"viscosity", CLV_VISCOSITY, 0, 0},
    {"soundspeed", CLV_SOUNDSPEED, 0, 0},   {"xvel0", CLV_XVEL0, 1, 1},             {"xvel1", CLV_XVEL1, 1, 1},
    {"yvel0", CLV_YVEL0, 1, 1},             {"yvel1", CLV_YVEL1, 1, 1},             {"vol_flux_x", CLV_VOL_FLUX_X, 1, 0},
    {"vol_flux_y", CLV_VOL_FLUX_Y, 0, 1},   {"mass_flux_x", CLV_MASS_FLUX_X, 1, 0}, {"mass_flux_y", CLV_MASS_FLUX_Y, 0, 1},
    {"work_array1", CLV_WORK1, 1, 1},       {"work_array2", CLV_WORK2, 1, 1},       {"work_array3", CLV_WORK3, 1, 1},
    {"work_array4", CLV_WORK4, 1, 1},       {"work_array5", CLV_WORK5, 1, 1},       {"work_array6", CLV_WORK6, 1, 1},
    {"work_array7", CLV_WORK7, 1, 1},
};
const DumpArray kDumpGeom[] = {{"volume", CLV_VOLUME, 0, 0}, {"xarea", CLV_XAREA, 1, 0}, {"yarea", CLV_YAREA, 0, 1}};
const std::pair<const char *, int> kDump1D[] = {{"cellx", CLV_CELLX},       {"celldx", CLV_CELLDX},     {"celly", CLV_CELLY},
                                                {"celldy", CLV_CELLDY},     {"vertexx", CLV_VERTEXX},   {"vertexdx", CLV_VERTEXDX},
                                                {"vertexy", CLV_VERTEXY},   {"vertexdy", CLV_VERTEXDY}};

struct Dumper {
  std::string dir;
  bool announced = false;
  void operator()(clv_chunk *c, const clv_config &cfg, const clv_hydro_state &st, const std::string &tag) {
    if (dir.empty()) return;
    const std::string file = "0_" + std::to_string(st.step) + "_" + tag + ".txt";
    std::cout << "Dumping globals to " << file << std::endl;
    const std::string d = dir + "/";
    struct stat info {};
    if (stat(d.c_str(), &info) != 0) {
      std::cout << "Creating " << d << " for field dump" << std::endl;
      if (errno = 0; mkdir(d.c_str(), 0777) != 0) {
        std::cerr << "Cannot create " << d << ": " << std::strerror(errno) << ", skipping field dump" << std::endl;
        return;
      }
    } else if (!(info.st_mode & S_IFDIR)) {
      std::cerr << d << " already exists and is not a directory, skipping field dump" << std::endl;
      return;
    }
    std::ofstream out(d + file, std::ofstream::out | std::ofstream::trunc);
    const int nx = cfg.x_cells, ny = cfg.y_cells;
    out << "Dump(tileCount = 1)\n"
        << "error_condition=0\n"
        << "step=" << st.step << "\n"
        << "advect_x=" << (st.advect_x ? 1 : 0) << "\n"
        << "time=" << st.time << "\n"
        << "dt=" << st.dt << "\n"
        << "dtold=" << st.dtold << "\n"
        << "complete=" << (st.complete ? 1 : 0) << "\n"
        << "jdt=" << st.jdt << "\n"
        << "kdt=" << st.kdt << "\n"
        << "\tTile[ 0]:\n";
    // one tile, every face external (driver/comms.cpp:158-253)
    for (int l = 0; l < 4; ++l) out << "info.tile_neighbours[i]=-1\ninfo.external_tile_mask[i]=1\n";
    out << "info.t_xmin=1\ninfo.t_xmax=" << nx << "\ninfo.t_ymin=1\ninfo.t_ymax=" << ny << "\ninfo.t_left=1\ninfo.t_right=" << nx
        << "\ninfo.t_bottom=1\ninfo.t_top=" << ny << "\n";
    std::vector<double> a((size_t)(nx + 5) * (ny + 5));
    auto show2 = [&](const DumpArray &f) {
      CHECK(clv_field_download(c, f.field, a.data(), nx + 5), "dump download");
      const int X = nx + 4 + f.xext, Y = ny + 4 + f.yext;
      out << f.name << "(2) [" << X << "x" << Y << "]\n\t";
      bool zero = true;
      for (int j = 0; j < Y && zero; ++j)
        for (int i = 0; i < X; ++i)
          if (a[(size_t)j * (nx + 5) + i] != 0.0) {
            zero = false;
            break;
          }
      if (zero) {
        out << "\t(0.0)";
      } else {
        for (int i = 0; i < X; ++i) {
          for (int j = 0; j < Y; ++j) out << a[(size_t)j * (nx + 5) + i] << ", ";
          out << "\t\n";
        }
      }
      out << "\n";
    };
    for (const auto &f : kDump2D) show2(f);
    for (const auto &m : kDump1D) {
      const size_t n = clv_mesh_length(nx, ny, m.second);
      std::vector<double> v(n);
      CHECK(clv_mesh_download(c, m.second, v.data()), "dump download");
      out << m.first << "(1) [" << n << "]\n";
      if (std::all_of(v.begin(), v.end(), [](double x) { return x == 0.0; })) {
        out << "\t(0.0)";
      } else {
        out << "\t";
        for (double x : v) out << x << ", ";
      }
      out << "\n";
    }
    for (const auto &f : kDumpGeom) show2(f);
    out << std::flush;
  }
};

// ---- visit output (driver/visit.cpp:38-196) ------------------------------------------------
// Every visit_frequency steps, after start-up and at the end: pressure, the pressure
// and velocity halos and viscosity brought up to date on the device, then clover.visit
// (the block count once, a file name each call) and one VTK rectilinear grid per call
// in the working directory, with the reference's file names and number formats.
struct Visit {
  bool first = true;
  void operator()(clv_chunk *c, const clv_config &cfg, const clv_hydro_state &st) {
    if (first) {
      std::ofstream u("clover.visit");
      u << "!NBLOCKS " << 1 << std::endl;  // chunks x tiles (one of each)
      first = false;
    }
    CHECK(clv_ideal_gas(c, 0), "visit ideal_gas");
    int fields[CLV_NUM_HALO_FIELDS] = {0};
    fields[CLV_PRESSURE] = fields[CLV_XVEL0] = fields[CLV_YVEL0] = 1;
    CHECK(clv_update_halo(c, fields, 1), "visit update_halo");
    CHECK(clv_viscosity(c), "visit viscosity");
    std::ostringstream name;
    name << "clover." << std::setfill('0') << std::setw(5) << 0 << "." << std::setw(5) << 1 << "." << std::setw(5) << st.step << ".vtk";
    {
      std::ofstream u("clover.visit", std::ios::app);
      u << name.str() << std::endl;
    }
    const int nx = cfg.x_cells, ny = cfg.y_cells, nxv = nx + 1, nyv = ny + 1;
    const int64_t P = nx + 5;
    std::vector<double> vx(clv_mesh_length(nx, ny, CLV_VERTEXX)), vy(clv_mesh_length(nx, ny, CLV_VERTEXY));
    CHECK(clv_mesh_download(c, CLV_VERTEXX, vx.data()), "visit download");
    CHECK(clv_mesh_download(c, CLV_VERTEXY, vy.data()), "visit download");
    std::ofstream u(name.str());
    u << "# vtk DataFile Version 3.0" << std::endl << "vtk output" << std::endl << "ASCII" << std::endl
      << "DATASET RECTILINEAR_GRID" << std::endl << "DIMENSIONS " << nxv << " " << nyv << " 1" << std::endl
      << "X_COORDINATES " << nxv << " double" << std::endl;
    for (int i = 2; i <= nx + 2; ++i) u << vx[i] << std::endl;
    u << "Y_COORDINATES " << nyv << " double" << std::endl;
    for (int j = 2; j <= ny + 2; ++j) u << vy[j] << std::endl;
    u << "Z_COORDINATES 1 double" << std::endl << "0" << std::endl;
    std::vector<double> a((size_t)P * (ny + 5));
    // cells [2, nx + 1] x [2, ny + 1] or nodes [2, nx + 2] x [2, ny + 2], x fastest;
    // small: values below 1e-8 in magnitude written as 0 (viscosity, velocities)
    auto block = [&](const char *what, int field, int n, int ext, bool small) {
      CHECK(clv_field_download(c, field, a.data(), P), "visit download");
      u << what << " 1 " << n << " double" << std::endl;
      for (int j = 2; j <= ny + 1 + ext; ++j)
        for (int i = 2; i <= nx + 1 + ext; ++i) {
          const double v = a[(size_t)j * P + i];
          u << std::scientific << std::setprecision(3) << (small && !(std::fabs(v) > 0.00000001) ? 0.0 : v) << std::endl;
        }
    };
    u << "CELL_DATA " << nx * ny << std::endl << "FIELD FieldData 4" << std::endl;
    block("density", CLV_DENSITY0, nx * ny, 0, false);
    block("energy", CLV_ENERGY0, nx * ny, 0, false);
    block("pressure", CLV_PRESSURE, nx * ny, 0, false);
    block("viscosity", CLV_VISCOSITY, nx * ny, 0, true);
    u << "POINT_DATA " << nxv * nyv << std::endl << "FIELD FieldData 2" << std::endl;
    block("x_vel", CLV_XVEL0, nxv * nyv, 1, true);
    block("y_vel", CLV_YVEL0, nxv * nyv, 1, true);
  }
};

// ---- --profile (driver/hydro.cpp:125-200) ---------------------------------------------------
// Kernel times from HIP events, grouped under the reference's profiler rows.  A fused
// kernel counts under the first reference phase it contains: f_timestep under Timestep
// (ideal_gas, viscosity and calc_dt), f_lagrange under PdV (with accelerate and
// flux_calc), the advection passes under Cell Advection (with the momentum remap).
enum Row { R_TIMESTEP, R_IDEAL_GAS, R_VISCOSITY, R_PDV, R_REVERT, R_ACCEL, R_FLUX, R_CELL, R_MOM, R_RESET, R_SUMMARY, R_VISIT,
           R_TILE_HALO, R_SELF_HALO, R_MPI_HALO, R_COUNT };
int row_of(const std::string &k) {
  auto starts = [&](const char *p) { return k.rfind(p, 0) == 0; };
  if (starts("f_timestep") || starts("calc_dt")) return R_TIMESTEP;
  if (k == "ideal_gas") return R_IDEAL_GAS;
  if (k == "viscosity") return R_VISCOSITY;
  if (starts("PdV") || k == "f_lagrange") return R_PDV;
  if (k == "revert") return R_REVERT;
  if (k == "accelerate") return R_ACCEL;
  if (k == "flux_calc") return R_FLUX;
  if (starts("advec_cell") || starts("f_cm")) return R_CELL;
  if (starts("advec_mom")) return R_MOM;
  if (k == "reset_field") return R_RESET;
  if (starts("field_summary")) return R_SUMMARY;
  if (k == "update_tile_halo") return R_TILE_HALO;
  if (k == "update_halo") return R_SELF_HALO;
  if (k == "pack_message") return R_MPI_HALO;
  return -1;
}

void profile_table(std::ostream &s, const double p[R_COUNT], double kernel_total, double wall_clock) {
  static const char *names[R_COUNT] = {" Timestep              :", " Ideal Gas             :", " Viscosity             :",
                                       " PdV                   :", " Revert                :", " Acceleration          :",
                                       " Fluxes                :", " Cell Advection        :", " Momentum Advection    :",
                                       " Reset                 :", " Summary               :", " Visit                 :",
                                       " Tile Halo Exchange    :", " Self Halo Exchange    :", " MPI Halo Exchange     :"};
  s << std::fixed << std::endl << " Profiler Output        Time     Percentage" << std::endl;
  for (int r = 0; r < R_COUNT; ++r) s << names[r] << p[r] << " " << 100.0 * (p[r] / wall_clock) << std::endl;
  s << " Total                 :" << kernel_total << " " << 100.0 * (kernel_total / wall_clock) << std::endl
    << " The Rest              :" << wall_clock - kernel_total << " " << 100.0 * (wall_clock - kernel_total) / wall_clock << std::endl
    << std::endl;
  s.unsetf(std::ios_base::floatfield);
}

}  // namespace

// field_summary with its ideal_gas: one pass (clv_hydro_summary), or, with the
// profiler on, the two entry points, so that its table keeps the summary's
// ideal_gas on the ideal_gas line as the reference's does
static clv_status summary(clv_chunk *c, int profiling, double s[5]) {
  if (!profiling) return clv_hydro_summary(c, s);
  const clv_status st = clv_ideal_gas(c, 0);
  return st == CLV_OK ? clv_field_summary(c, s) : st;
}

int main(int argc, char **argv) {
  Args args = parse_args(argc, argv);
  std::cout << "---" << std::endl;
  list_devices();
  if (clv_device_count() < 1) {
    std::cerr << "clover_mi355x: no HIP device visible" << std::endl;
    return EXIT_FAILURE;
  }
  // driver/clover_leaf.cpp:96-117
  std::cout << "CloverLeaf:\n"
            << " - Ver.:     " << kVersion << "\n"
            << " - Deck:     " << args.in << "\n"
            << " - Out:      " << args.out << "\n"
            << " - Profiler: " << (args.profile < 0 ? "deck-specified" : (args.profile ? "true" : "false")) << "\n"
            << "MPI:\n"
            << " - Enabled:     false\n"
            << " - Total ranks: 1\n"
            << " - Header device-awareness (CUDA-awareness):  unknown\n"
            << " - Runtime device-awareness (CUDA-awareness): unknown\n"
            << " - Host-Device halo exchange staging buffer:  true\n"
            << "Model:\n"
            << " - Name:      MI355X\n"
            << " - Execution: Offload (device)" << std::endl;
  std::cout << " - Device:    " << args.device << " (" << clv_version() << ")" << std::endl;
  std::cout << "# ---- " << std::endl << "Output: |+1" << std::endl;

  std::cout << " Output file " << args.out << " opened. All output will go there." << std::endl;
  std::ofstream g_out(args.out);
  if (!g_out.is_open()) {
    std::cerr << std::endl << "Error from initialise:" << std::endl << "Error opening clover.out file." << std::endl
              << "CLOVER is terminating." << std::endl << std::endl;
    return EXIT_FAILURE;
  }
  g_out << "Clover Version " << kVersion << std::endl << "Task Count 1" << std::endl << std::endl;
  g_out << "Clover will run from the following input:-" << std::endl << std::endl;
  if (!args.raw.empty()) {
    std::cout << " Args:";
    for (const auto &a : args.raw) std::cout << " " << a;
    std::cout << std::endl;
  }
  std::cout << " Using input: `" << args.in << "`" << std::endl;
  std::ifstream fin(args.in);
  if (fin.fail()) {
    std::cerr << "Unable to open file: `" << args.in << "`" << std::endl;
    return 1;
  }
  std::stringstream ss;
  ss << fin.rdbuf();
  g_out << std::endl << "Initialising and generating" << std::endl << std::endl;
  echo_deck(g_out, ss.str());
  clv_config cfg;
  clv_status st = clv_deck_parse(ss.str().c_str(), &cfg);
  if (st != CLV_OK) die("read_input", st);
  if (args.profile >= 0) cfg.profiler_on = args.profile;
  if (cfg.tiles_per_chunk != 1) std::cerr << "# note: tiles_per_chunk " << cfg.tiles_per_chunk << " ignored, one tile per GPU" << std::endl;

  int plan = args.plan;
  if (!args.dump.empty() || cfg.x_cells < 4 || cfg.y_cells < 4) plan = CLV_PLAN_REFERENCE;

  // driver/start.cpp:44-127 (one chunk of one tile)
  g_out << "Setting up initial geometry" << std::endl << std::endl;
  g_out << std::endl
        << "Mesh ratio of " << (double)cfg.x_cells / (double)cfg.y_cells << std::endl
        << "Decomposing the mesh into 1 by 1 chunks" << std::endl
        << "Decomposing the chunk with " << cfg.tiles_per_chunk << " tiles" << std::endl
        << std::endl;
  clv_chunk *c = nullptr;
  CHECK(clv_chunk_create(cfg.x_cells, cfg.y_cells, &c), "clv_chunk_create");
  CHECK(clv_chunk_set_geometry_mode(c, args.geometry), "clv_chunk_set_geometry_mode");
  if (cfg.profiler_on) CHECK(clv_profile_enable(c, 1), "clv_profile_enable");
  Dumper dump{args.dump};
  clv_hydro_state hs;
  std::memset(&hs, 0, sizeof(hs));
  hs.advect_x = 1;
  hs.dt = hs.dtold = cfg.dtinit;  // driver/definitions.cpp:53-55
  double s[5];
  CHECK(clv_build_field(c), "build_field");
  g_out << "Generating chunks" << std::endl;
  dump(c, cfg, hs, "00_build_field");
  CHECK(clv_initialise_chunk(c, &cfg), "initialise_chunk");
  CHECK(clv_generate_chunk(c, &cfg), "generate_chunk");
  dump(c, cfg, hs, "01_generate_chunk");
  CHECK(clv_ideal_gas(c, 0), "ideal_gas");
  dump(c, cfg, hs, "02_ideal_gas");
  const int halo0[CLV_NUM_HALO_FIELDS] = {1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 0, 0, 0};  // driver/start.cpp:104-116
  CHECK(clv_update_halo(c, halo0, 2), "update_halo");
  dump(c, cfg, hs, "03_update_halo");
  std::cout << " Problem initialised and generated" << std::endl;
  g_out << std::endl << "Problem initialised and generated" << std::endl;
  CHECK(summary(c, cfg.profiler_on, s), "field_summary");
  report_step(g_out, hs, s);
  dump(c, cfg, hs, "04_field_summary");
  Visit visit;
  if (cfg.visit_frequency != 0) visit(c, cfg, hs);  // driver/start.cpp:131
  g_out << "Starting the calculation" << std::endl;
  std::cout << " Launching hydro" << std::endl;
  if (cfg.profiler_on) CHECK(clv_profile_reset(c), "clv_profile_reset");

  // driver/hydro.cpp:46-222
  const double timerstart = wall();
  dump(c, cfg, hs, "05_hydro");
  const char *control_names[] = {"", "sound", "xvel", "yvel", "div"};
  static const char *phase_tags[] = {"1_timestep", "2_PdV", "3_accelerate", "4_PdV", "5_flux_calc", "6_advection", "7_reset_field"};
  bool failed = false;
  while (true) {
    const double step_time = wall();
    const double t_before = hs.time;
    if (plan == CLV_PLAN_REFERENCE) {
      for (int ph = CLV_PHASE_TIMESTEP; ph <= CLV_PHASE_END_STEP; ++ph) {
        if ((st = clv_hydro_phase(c, &cfg, &hs, ph)) != CLV_OK) die("hydro step", st);
        if (ph == CLV_PHASE_END_STEP) break;
        if (ph == CLV_PHASE_TIMESTEP) {
          const char *ctl = (hs.dt_control >= 1 && hs.dt_control <= 4) ? control_names[hs.dt_control] : "";
          // the reference prints the time before this step's increment (driver/timestep.cpp:99)
          g_out << " Step " << hs.step << " time " << t_before << " control " << ctl << " timestep  " << hs.dt << " " << hs.jdt << ","
                << hs.kdt << " x 0 y 0" << std::endl;
          if (!args.quiet)
            std::cout << " Step " << hs.step << " time " << t_before << " control " << ctl << " timestep  " << hs.dt << " " << hs.jdt
                      << "," << hs.kdt << " x 0 y 0" << std::endl;
        }
        dump(c, cfg, hs, phase_tags[ph]);
      }
    } else {
      if ((st = clv_hydro_step(c, &cfg, &hs, plan)) != CLV_OK) die("hydro step", st);
      const char *ctl = (hs.dt_control >= 1 && hs.dt_control <= 4) ? control_names[hs.dt_control] : "";
      g_out << " Step " << hs.step << " time " << t_before << " control " << ctl << " timestep  " << hs.dt << " " << hs.jdt << ","
            << hs.kdt << " x 0 y 0" << std::endl;
      if (!args.quiet)
        std::cout << " Step " << hs.step << " time " << t_before << " control " << ctl << " timestep  " << hs.dt << " " << hs.jdt << ","
                  << hs.kdt << " x 0 y 0" << std::endl;
    }
    if (cfg.summary_frequency != 0 && hs.step % cfg.summary_frequency == 0) {
      CHECK(summary(c, cfg.profiler_on, s), "field_summary");
      report_step(g_out, hs, s);
    }
    if (cfg.visit_frequency != 0 && hs.step % cfg.visit_frequency == 0) visit(c, cfg, hs);  // driver/hydro.cpp:95-97
    // the reference resets these every step, so its completion line prints 0 unless
    // the run ends at step 1 or 2 (driver/hydro.cpp:103-106)
    double first_step = 0, second_step = 0;
    if (hs.step == 1) first_step = wall() - step_time;
    if (hs.step == 2) second_step = wall() - step_time;
    if (hs.complete) {
      CHECK(summary(c, cfg.profiler_on, s), "field_summary");
      report_step(g_out, hs, s);
      if (cfg.visit_frequency != 0) visit(c, cfg, hs);  // driver/hydro.cpp:112
      if (cfg.test_problem >= 1) {  // driver/report.cpp:73-111
        double qa_diff = 0.0, ke_exp = 0.0;
        if (expected_ke(cfg.test_problem, ke_exp)) {
          qa_diff = std::fabs((100.0 * (s[3] / ke_exp)) - 100.0);
        } else {
          std::cout << " WARNING: Unknown test problem " << cfg.test_problem << ", validation will fail" << std::endl;
          g_out << " WARNING: Unknown test problem " << cfg.test_problem << ", validation will fail" << std::endl;
          qa_diff = 100.0;
        }
        std::cout << " Test problem " << cfg.test_problem << " is within " << qa_diff << "% of the expected solution" << std::endl;
        g_out << "Test problem " << cfg.test_problem << " is within " << qa_diff << "% of the expected solution" << std::endl;
        if (qa_diff < 0.001) {
          std::cout << " This test is considered PASSED" << std::endl;
          g_out << "This test is considered PASSED" << std::endl;
        } else {
          std::cout << " This test is considered NOT PASSED" << std::endl;
          g_out << "This test is considered NOT PASSED" << std::endl;
          failed = true;
        }
      }
      const double wall_clock = wall() - timerstart;
      g_out << std::endl
            << "Calculation complete" << std::endl
            << "Clover is finishing" << std::endl
            << "Wall clock " << wall_clock << std::endl
            << "First step overhead " << first_step - second_step << std::endl;
      std::cout << " Wall clock " << wall_clock << std::endl << " First step overhead " << first_step - second_step << std::endl;
      if (cfg.profiler_on) {
        const char *names[128];
        int64_t launches[128];
        double ms[128];
        const int n = std::min(clv_profile_read(c, names, launches, ms, 128), 128);
        double p[R_COUNT] = {0};
        double total = 0.0;
        for (int k = 0; k < n; ++k) {
          const int r = row_of(names[k]);
          if (r >= 0) p[r] += ms[k] * 1e-3;
          total += ms[k] * 1e-3;
        }
        profile_table(g_out, p, total, wall_clock);
        profile_table(std::cout, p, total, wall_clock);
        std::cout << " Kernel profile (HIP events, plan " << (plan == CLV_PLAN_FUSED ? "fused" : "reference") << "):" << std::endl;
        for (int k = 0; k < n; ++k)
          std::cout << "  " << std::left << std::setw(24) << names[k] << std::right << std::setw(8) << launches[k] << std::setw(14) << ms[k]
                    << " ms" << std::endl;
      }
      break;
    }
    // driver/hydro.cpp:205-218
    const double wall_clock = wall() - timerstart;
    const double step_clock = wall() - step_time;
    g_out << "Wall clock " << wall_clock << std::endl;
    std::cout << " Wall clock " << wall_clock << std::endl;
    const double cells = (double)cfg.x_cells * cfg.y_cells;
    const double grind_time = wall_clock / ((double)hs.step * cells), step_grind = step_clock / cells;
    std::cout << " Average time per cell " << grind_time << std::endl;
    g_out << "Average time per cell " << grind_time << std::endl;
    std::cout << "  Step time per cell    " << step_grind << std::endl;
    g_out << "Step time per cell    " << step_grind << std::endl;
  }
  clv_chunk_destroy(c);
  // driver/clover_leaf.cpp:211-216
  std::cout << "Result:\n"
            << " - Problem: " << (cfg.test_problem == 0 ? "none" : std::to_string(cfg.test_problem)) << "\n"
            << " - Outcome: " << (failed ? "FAILED" : "PASSED") << std::endl;
  return failed ? EXIT_FAILURE : EXIT_SUCCESS;
}
