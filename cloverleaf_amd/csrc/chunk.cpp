// C ABI layer: chunk lifetime, copies, errors, per-launch profiling, and the
// per-kernel entry points that forward to the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <utility>

#include "clv_internal.h"

namespace clv {

static thread_local std::string g_last_error;

clv_status set_error(clv_status code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

clv_status hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return CLV_OK;
  return set_error(CLV_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Launch-level timing with HIP events on the chunk's stream.  Event pairs are
// folded into per-name totals once the stream has passed them.
struct Profiler {
  struct Pending {
    int name;
    hipEvent_t a, b;
  };
  std::vector<std::string> names;
  std::map<std::string, int> index;
  std::vector<int64_t> launches;
  std::vector<double> total_ms;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t open = nullptr;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  int id(const char *n) {
    auto it = index.find(n);
    if (it != index.end()) return it->second;
    int k = (int)names.size();
    names.emplace_back(n);
    index[n] = k;
    launches.push_back(0);
    total_ms.push_back(0.0);
    return k;
  }
  void fold(bool all) {
    size_t keep = 0;
    for (size_t p = 0; p < pending.size(); ++p) {
      Pending &q = pending[p];
      if (!all && hipEventQuery(q.b) != hipSuccess) {
        pending[keep++] = q;
        continue;
      }
      if (all) (void)hipEventSynchronize(q.b);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, q.a, q.b);
      launches[q.name] += 1;
      total_ms[q.name] += ms;
      pool.push_back(q.a);
      pool.push_back(q.b);
    }
    pending.resize(keep);
  }
  ~Profiler() {
    fold(true);
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

void before_launch(clv_chunk *c, const char *name) {
  (void)name;
  if (!c->prof) return;
  c->prof->open = c->prof->get();
  (void)hipEventRecord(c->prof->open, c->stream);
}

clv_status after_launch(clv_chunk *c, const char *name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(CLV_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
  if (c->prof) {
    Profiler &p = *c->prof;
    hipEvent_t b = p.get();
    (void)hipEventRecord(b, c->stream);
    p.pending.push_back({p.id(name), p.open, b});
    if (p.pending.size() > 4096) p.fold(false);
  }
  return CLV_OK;
}

static clv_status alloc_scratch(clv_chunk *c) {
  // block partials of the reductions: 5 sums x 2048 blocks, or one minimum per
  // wave of the fused timestep kernel, whichever is larger
  size_t n = 5 * 2048;
  const size_t nf = fused_partials_needed(c->d.nx, c->d.ny);
  if (nf > n) n = nf;
  clv_status st = hip_check(hipMalloc(&c->partials, sizeof(double) * n), "hipMalloc partials");
  if (st != CLV_OK) return st;
  c->partials_len = n;
  st = hip_check(hipMalloc(&c->result_dev, sizeof(double) * 8), "hipMalloc result");
  if (st != CLV_OK) return st;
  return hip_check(hipHostMalloc((void **)&c->result_host, sizeof(double) * 8, hipHostMallocDefault), "hipHostMalloc result");
}

static bool valid_field(int f) { return f >= 0 && f < CLV_NUM_FIELDS; }
static bool valid_mesh(int m) { return m >= 0 && m < CLV_NUM_MESH; }
static size_t mesh_len(const clv_chunk *c, int m) { return clv_mesh_length(c->d.nx, c->d.ny, m); }

}  // namespace clv

using namespace clv;

#define CHECK_CHUNK(c) \
  if (!(c)) return set_error(CLV_ERR_ARG, "null chunk")

extern "C" {

const char *clv_last_error(void) { return g_last_error.c_str(); }
const char *clv_version(void) { return "clover_mi355x 0.1 (gfx950, FP64)"; }

int clv_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

clv_status clv_chunk_create(int x_cells, int y_cells, clv_chunk **out) {
  if (!out || x_cells < 2 || y_cells < 2) return set_error(CLV_ERR_ARG, "clv_chunk_create: need x_cells, y_cells >= 2 and an out pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(CLV_ERR_HIP, "clv_chunk_create: no HIP device");
  auto *c = new clv_chunk();
  c->d.nx = x_cells;
  c->d.ny = y_cells;
  c->d.pitch = pitch_for(x_cells);
  (void)hipGetDevice(&c->device);
  // one arena: 25 2D arrays, then the 8 mesh arrays; every array starts on a 256 B boundary
  const size_t per2d = ((size_t)kLeadPad + (size_t)(y_cells + 5) * (size_t)c->d.pitch + 31) / 32 * 32;
  const size_t per1d = ((size_t)(x_cells > y_cells ? x_cells : y_cells) + 5 + 31) / 32 * 32;
  c->arena_bytes = sizeof(double) * (per2d * CLV_NUM_FIELDS + per1d * CLV_NUM_MESH);
  hipError_t e = hipMalloc(&c->arena, c->arena_bytes);
  if (e != hipSuccess) {
    delete c;
    return set_error(CLV_ERR_ALLOC, std::string("clv_chunk_create: hipMalloc of arena failed: ") + hipGetErrorString(e));
  }
  c->owns = true;
  double *base = static_cast<double *>(c->arena);
  for (int f = 0; f < CLV_NUM_FIELDS; ++f) c->d.f[f] = base + (size_t)f * per2d + kLeadPad;
  for (int m = 0; m < CLV_NUM_MESH; ++m) c->d.m[m] = base + per2d * CLV_NUM_FIELDS + (size_t)m * per1d;
  clv_status st = alloc_scratch(c);
  if (st != CLV_OK) {
    clv_chunk_destroy(c);
    return st;
  }
  *out = c;
  return CLV_OK;
}

clv_status clv_chunk_wrap(const clv_chunk_view *v, clv_chunk **out) {
  if (!v || !out) return set_error(CLV_ERR_ARG, "clv_chunk_wrap: null argument");
  if (v->x_cells < 2 || v->y_cells < 2 || v->pitch < v->x_cells + 5) return set_error(CLV_ERR_ARG, "clv_chunk_wrap: bad sizes or pitch");
  for (int f = 0; f < CLV_NUM_FIELDS; ++f)
    if (!v->field[f]) return set_error(CLV_ERR_ARG, "clv_chunk_wrap: null field pointer");
  for (int m = 0; m < CLV_NUM_MESH; ++m)
    if (!v->mesh[m]) return set_error(CLV_ERR_ARG, "clv_chunk_wrap: null mesh pointer");
  auto *c = new clv_chunk();
  c->d.nx = v->x_cells;
  c->d.ny = v->y_cells;
  c->d.pitch = v->pitch;
  for (int f = 0; f < CLV_NUM_FIELDS; ++f) c->d.f[f] = v->field[f];
  for (int m = 0; m < CLV_NUM_MESH; ++m) c->d.m[m] = v->mesh[m];
  (void)hipGetDevice(&c->device);
  clv_status st = alloc_scratch(c);
  if (st != CLV_OK) {
    clv_chunk_destroy(c);
    return st;
  }
  *out = c;
  return CLV_OK;
}

clv_status clv_chunk_destroy(clv_chunk *c) {
  if (!c) return CLV_OK;
  (void)hipStreamSynchronize(c->stream);
  delete c->prof;
  if (c->owns && c->arena) (void)hipFree(c->arena);
  if (c->partials) (void)hipFree(c->partials);
  if (c->result_dev) (void)hipFree(c->result_dev);
  if (c->result_host) (void)hipHostFree(c->result_host);
  delete c;
  return CLV_OK;
}

clv_status clv_chunk_get_view(const clv_chunk *c, clv_chunk_view *out) {
  CHECK_CHUNK(c);
  if (!out) return set_error(CLV_ERR_ARG, "null view");
  // the caller may write any array through the view: read the geometry arrays from now on
  const_cast<clv_chunk *>(c)->d.geom_mesh = 0;
  out->x_cells = c->d.nx;
  out->y_cells = c->d.ny;
  out->pitch = c->d.pitch;
  for (int f = 0; f < CLV_NUM_FIELDS; ++f) out->field[f] = c->d.f[f];
  for (int m = 0; m < CLV_NUM_MESH; ++m) out->mesh[m] = c->d.m[m];
  return CLV_OK;
}

int clv_chunk_geometry_from_mesh(const clv_chunk *c) { return c ? c->d.geom_mesh : -1; }

clv_status clv_chunk_set_geometry_mode(clv_chunk *c, int mode) {
  CHECK_CHUNK(c);
  if (mode != CLV_GEOMETRY_ARRAYS && mode != CLV_GEOMETRY_MESH)
    return set_error(CLV_ERR_ARG, "clv_chunk_set_geometry_mode: mode is CLV_GEOMETRY_ARRAYS or CLV_GEOMETRY_MESH");
  // wrapped chunks: the caller keeps write access to every array, so never assume them
  if (mode == CLV_GEOMETRY_MESH && !c->owns)
    return set_error(CLV_ERR_ARG, "clv_chunk_set_geometry_mode: CLV_GEOMETRY_MESH needs a chunk made by clv_chunk_create");
  c->geom_mode = mode;
  if (mode == CLV_GEOMETRY_ARRAYS) c->d.geom_mesh = 0;
  return CLV_OK;
}

clv_status clv_chunk_set_stream(clv_chunk *c, void *s) {
  CHECK_CHUNK(c);
  c->stream = static_cast<hipStream_t>(s);
  return CLV_OK;
}

clv_status clv_chunk_synchronize(clv_chunk *c) {
  CHECK_CHUNK(c);
  return hip_check(hipStreamSynchronize(c->stream), "clv_chunk_synchronize");
}

size_t clv_chunk_bytes(const clv_chunk *c) { return c && c->owns ? c->arena_bytes : 0; }

clv_status clv_device_alloc(size_t bytes, void **out) {
  if (!out) return set_error(CLV_ERR_ARG, "clv_device_alloc: null out");
  *out = nullptr;
  if (bytes == 0) return CLV_OK;
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess) return set_error(CLV_ERR_ALLOC, std::string("clv_device_alloc: ") + hipGetErrorString(e));
  return CLV_OK;
}

clv_status clv_device_free(void *p) { return p ? hip_check(hipFree(p), "clv_device_free") : CLV_OK; }

clv_status clv_copy_to_host(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return CLV_OK;
  if (!dst || !src) return set_error(CLV_ERR_ARG, "clv_copy_to_host: null pointer");
  return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "clv_copy_to_host");
}

clv_status clv_copy_to_device(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return CLV_OK;
  if (!dst || !src) return set_error(CLV_ERR_ARG, "clv_copy_to_device: null pointer");
  return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "clv_copy_to_device");
}

clv_status clv_field_upload(clv_chunk *c, int field, const double *host, int64_t host_pitch) {
  CHECK_CHUNK(c);
  if (!valid_field(field) || !host || host_pitch < c->d.nx + 5) return set_error(CLV_ERR_ARG, "clv_field_upload: bad argument");
  if (field == CLV_VOLUME || field == CLV_XAREA || field == CLV_YAREA) c->d.geom_mesh = 0;
  clv_status st = hip_check(hipMemcpy2DAsync(c->d.f[field], c->d.pitch * sizeof(double), host, host_pitch * sizeof(double),
                                             (c->d.nx + 5) * sizeof(double), c->d.ny + 5, hipMemcpyHostToDevice, c->stream),
                            "clv_field_upload");
  if (st != CLV_OK) return st;
  return hip_check(hipStreamSynchronize(c->stream), "clv_field_upload sync");
}

clv_status clv_field_download(clv_chunk *c, int field, double *host, int64_t host_pitch) {
  CHECK_CHUNK(c);
  if (!valid_field(field) || !host || host_pitch < c->d.nx + 5) return set_error(CLV_ERR_ARG, "clv_field_download: bad argument");
  clv_status st = hip_check(hipMemcpy2DAsync(host, host_pitch * sizeof(double), c->d.f[field], c->d.pitch * sizeof(double),
                                             (c->d.nx + 5) * sizeof(double), c->d.ny + 5, hipMemcpyDeviceToHost, c->stream),
                            "clv_field_download");
  if (st != CLV_OK) return st;
  return hip_check(hipStreamSynchronize(c->stream), "clv_field_download sync");
}

clv_status clv_mesh_upload(clv_chunk *c, int mesh, const double *host) {
  CHECK_CHUNK(c);
  if (!valid_mesh(mesh) || !host) return set_error(CLV_ERR_ARG, "clv_mesh_upload: bad argument");
  // the areas were derived from the old spacings
  if (mesh == CLV_CELLDX || mesh == CLV_CELLDY) c->d.geom_mesh = 0;
  clv_status st = hip_check(hipMemcpyAsync(c->d.m[mesh], host, mesh_len(c, mesh) * sizeof(double), hipMemcpyHostToDevice, c->stream),
                            "clv_mesh_upload");
  if (st != CLV_OK) return st;
  return hip_check(hipStreamSynchronize(c->stream), "clv_mesh_upload sync");
}

clv_status clv_mesh_download(clv_chunk *c, int mesh, double *host) {
  CHECK_CHUNK(c);
  if (!valid_mesh(mesh) || !host) return set_error(CLV_ERR_ARG, "clv_mesh_download: bad argument");
  clv_status st = hip_check(hipMemcpyAsync(host, c->d.m[mesh], mesh_len(c, mesh) * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                            "clv_mesh_download");
  if (st != CLV_OK) return st;
  return hip_check(hipStreamSynchronize(c->stream), "clv_mesh_download sync");
}

clv_status clv_build_field(clv_chunk *c) {
  CHECK_CHUNK(c);
  return k_fill_zero(c);
}

clv_status clv_initialise_chunk(clv_chunk *c, const clv_config *cfg) {
  CHECK_CHUNK(c);
  if (!cfg || cfg->x_cells != c->d.nx || cfg->y_cells != c->d.ny) return set_error(CLV_ERR_ARG, "clv_initialise_chunk: config does not match chunk");
  return k_initialise_chunk(c, cfg, 1, 1);
}

clv_status clv_generate_chunk(clv_chunk *c, const clv_config *cfg) {
  CHECK_CHUNK(c);
  if (!cfg) return set_error(CLV_ERR_ARG, "clv_generate_chunk: null config");
  return k_generate_chunk(c, cfg);
}

clv_status clv_ideal_gas(clv_chunk *c, int predict) {
  CHECK_CHUNK(c);
  return k_ideal_gas(c, predict);
}

clv_status clv_viscosity(clv_chunk *c) {
  CHECK_CHUNK(c);
  return k_viscosity(c);
}

clv_status clv_calc_dt(clv_chunk *c, double dtmin, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe, double *dt_min_val,
                       int *dt_control, int *jldt, int *kldt, int *small) {
  CHECK_CHUNK(c);
  if (!dt_min_val) return set_error(CLV_ERR_ARG, "clv_calc_dt: null dt_min_val");
  clv_status st = k_calc_dt(c, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, dt_min_val);
  if (st != CLV_OK) return st;
  // constant bookkeeping of src/serial/calc_dt.cpp:40,80-83
  double jk_control = 1.1;
  const int x_max = c->d.nx;
  const int l_control = (int)(10.01 * (jk_control - (int)(jk_control)));
  jk_control = jk_control - (jk_control - (int)(jk_control));
  if (dt_control) *dt_control = l_control;
  if (jldt) *jldt = ((int)jk_control) % x_max;
  if (kldt) *kldt = (int)(1.f + (jk_control / x_max));
  if (small) *small = *dt_min_val < dtmin ? 1 : 0;
  return CLV_OK;
}

clv_status clv_timestep_kernels(clv_chunk *c, double dtmin, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe,
                                int pressure_halo_depth, double *dt_min_val, int *dt_control, int *jldt, int *kldt, int *small) {
  CHECK_CHUNK(c);
  if (!dt_min_val) return set_error(CLV_ERR_ARG, "clv_timestep_kernels: null dt_min_val");
  if (pressure_halo_depth < 0 || pressure_halo_depth > 2) return set_error(CLV_ERR_ARG, "clv_timestep_kernels: pressure_halo_depth must be 0, 1 or 2");
  clv_status st = fused_timestep_entry(c, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, pressure_halo_depth, dt_min_val);
  if (st != CLV_OK) return st;
  // the bookkeeping of clv_calc_dt (src/serial/calc_dt.cpp:40,80-83)
  double jk_control = 1.1;
  const int x_max = c->d.nx;
  const int l_control = (int)(10.01 * (jk_control - (int)(jk_control)));
  jk_control = jk_control - (jk_control - (int)(jk_control));
  if (dt_control) *dt_control = l_control;
  if (jldt) *jldt = ((int)jk_control) % x_max;
  if (kldt) *kldt = (int)(1.f + (jk_control / x_max));
  if (small) *small = *dt_min_val < dtmin ? 1 : 0;
  return CLV_OK;
}

clv_status clv_advection_pass(clv_chunk *c, int sweep_number, int direction) {
  CHECK_CHUNK(c);
  if ((sweep_number != 1 && sweep_number != 2) || (direction != CLV_XDIR && direction != CLV_YDIR))
    return set_error(CLV_ERR_ARG, "clv_advection_pass: sweep_number must be 1 or 2, direction CLV_XDIR or CLV_YDIR");
  return fused_advection_entry(c, sweep_number, direction, false);
}

clv_status clv_advection_pass_reset(clv_chunk *c, int sweep_number, int direction) {
  CHECK_CHUNK(c);
  if ((sweep_number != 1 && sweep_number != 2) || (direction != CLV_XDIR && direction != CLV_YDIR))
    return set_error(CLV_ERR_ARG, "clv_advection_pass_reset: sweep_number must be 1 or 2, direction CLV_XDIR or CLV_YDIR");
  return fused_advection_entry(c, sweep_number, direction, true);
}

clv_status clv_lagrange_kernels(clv_chunk *c, double dt) {
  CHECK_CHUNK(c);
  return fused_lagrange_entry(c, dt);
}

clv_status clv_PdV(clv_chunk *c, int predict, double dt) {
  CHECK_CHUNK(c);
  return k_PdV(c, predict, dt);
}

clv_status clv_PdV_driver(clv_chunk *c, int predict, double dt) {
  CHECK_CHUNK(c);
  if (predict && c->entry_fused) return e_PdV_predict_chain(c, dt);
  clv_status st = k_PdV(c, predict, dt);
  if (st != CLV_OK || !predict) return st;
  if ((st = k_ideal_gas(c, 1)) != CLV_OK) return st;
  int fields[CLV_NUM_HALO_FIELDS] = {0};
  fields[CLV_PRESSURE] = 1;
  if ((st = k_update_halo(c, fields, 1)) != CLV_OK) return st;
  return k_revert(c);
}

clv_status clv_revert(clv_chunk *c) {
  CHECK_CHUNK(c);
  return k_revert(c);
}

clv_status clv_accelerate(clv_chunk *c, double dt) {
  CHECK_CHUNK(c);
  return k_accelerate(c, dt);
}

clv_status clv_flux_calc(clv_chunk *c, double dt) {
  CHECK_CHUNK(c);
  return k_flux_calc(c, dt);
}

clv_status clv_advec_cell(clv_chunk *c, int sweep_number, int direction) {
  CHECK_CHUNK(c);
  if ((sweep_number != 1 && sweep_number != 2) || (direction != CLV_XDIR && direction != CLV_YDIR))
    return set_error(CLV_ERR_ARG, "clv_advec_cell: sweep_number must be 1|2 and direction 1|2");
  if (!c->entry_fused) return k_advec_cell(c, sweep_number, direction);
  // the new density1 / energy1 go to work arrays 6 / 7, which then take their place
  clv_status st = e_advec_cell(c, sweep_number, direction, c->d.f[CLV_WORK6], c->d.f[CLV_WORK7]);
  if (st != CLV_OK) return st;
  std::swap(c->d.f[CLV_DENSITY1], c->d.f[CLV_WORK6]);
  std::swap(c->d.f[CLV_ENERGY1], c->d.f[CLV_WORK7]);
  return CLV_OK;
}

clv_status clv_advec_mom(clv_chunk *c, int which_vel, int direction, int sweep_number) {
  CHECK_CHUNK(c);
  if ((which_vel != 1 && which_vel != 2) || (sweep_number != 1 && sweep_number != 2) || (direction != CLV_XDIR && direction != CLV_YDIR))
    return set_error(CLV_ERR_ARG, "clv_advec_mom: which_vel, direction and sweep_number must each be 1|2");
  if (!c->entry_fused) return k_advec_mom(c, which_vel, direction, sweep_number);
  // the new velocity goes to work array 5, which then takes its place
  clv_status st = e_advec_mom(c, which_vel, direction, sweep_number, c->d.f[CLV_WORK5]);
  if (st != CLV_OK) return st;
  std::swap(c->d.f[which_vel == 1 ? CLV_XVEL1 : CLV_YVEL1], c->d.f[CLV_WORK5]);
  return CLV_OK;
}

clv_status clv_chunk_set_entry_plan(clv_chunk *c, int plan) {
  CHECK_CHUNK(c);
  if (plan != CLV_ENTRY_REFERENCE && plan != CLV_ENTRY_FUSED)
    return set_error(CLV_ERR_ARG, "clv_chunk_set_entry_plan: plan is CLV_ENTRY_REFERENCE or CLV_ENTRY_FUSED");
  c->entry_fused = plan == CLV_ENTRY_FUSED;
  return CLV_OK;
}

clv_status clv_reset_field(clv_chunk *c) {
  CHECK_CHUNK(c);
  return k_reset_field(c);
}

clv_status clv_update_halo(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth) {
  CHECK_CHUNK(c);
  if (!fields || (depth != 1 && depth != 2)) return set_error(CLV_ERR_ARG, "clv_update_halo: fields required and depth must be 1 or 2");
  return k_update_halo(c, fields, depth);
}

clv_status clv_initialise_tile(clv_chunk *c, const clv_config *cfg, int t_left, int t_bottom) {
  CHECK_CHUNK(c);
  if (!cfg) return set_error(CLV_ERR_ARG, "clv_initialise_tile: null config");
  if (t_left < 1 || t_bottom < 1 || t_left + c->d.nx - 1 > cfg->x_cells || t_bottom + c->d.ny - 1 > cfg->y_cells)
    return set_error(CLV_ERR_ARG, "clv_initialise_tile: the tile does not lie inside the grid");
  return k_initialise_chunk(c, cfg, t_left, t_bottom);
}

static bool faces_ok(int face) { return face >= CLV_FACE_LEFT && face <= CLV_FACE_TOP; }

clv_status clv_update_halo_faces(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth, const int external[4]) {
  CHECK_CHUNK(c);
  if (!fields || !external || (depth != 1 && depth != 2))
    return set_error(CLV_ERR_ARG, "clv_update_halo_faces: fields and external required, depth must be 1 or 2");
  return k_update_halo_faces(c, fields, depth, external);
}

clv_status clv_update_tile_halo(clv_chunk *tile, clv_chunk *neighbour, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth) {
  CHECK_CHUNK(tile);
  CHECK_CHUNK(neighbour);
  if (!fields || !faces_ok(face) || (depth != 1 && depth != 2))
    return set_error(CLV_ERR_ARG, "clv_update_tile_halo: fields required, face CLV_FACE_*, depth 1 or 2");
  if (tile == neighbour) return set_error(CLV_ERR_ARG, "clv_update_tile_halo: a tile is not its own neighbour");
  if (face <= CLV_FACE_RIGHT ? neighbour->d.ny != tile->d.ny : neighbour->d.nx != tile->d.nx)
    return set_error(CLV_ERR_ARG, "clv_update_tile_halo: neighbours across a left/right face share y_cells, across bottom/top x_cells");
  if (neighbour->device != tile->device) return set_error(CLV_ERR_ARG, "clv_update_tile_halo: tiles on different devices");
  // the neighbour's interior must be final before the copy on the tile's stream
  if (neighbour->stream != tile->stream) {
    clv_status st = hip_check(hipStreamSynchronize(neighbour->stream), "clv_update_tile_halo: neighbour sync");
    if (st != CLV_OK) return st;
  }
  return k_tile_halo(tile, neighbour, face, fields, depth);
}

// one past the largest message element a face's message of one field touches
static int64_t message_extent(const clv_chunk *c, int face, int x_inc, int y_inc, int depth) {
  const int64_t len = face <= CLV_FACE_RIGHT ? c->d.ny + y_inc : c->d.nx + x_inc;
  return (len + depth + 1) * depth + depth;
}

static clv_status pack_array(clv_chunk *c, double *field, int field_type, int face, int depth, double *buffer, size_t buffer_len,
                             int64_t offset, int unpack, const char *who) {
  CHECK_CHUNK(c);
  if (!field || !buffer || !faces_ok(face) || (depth != 1 && depth != 2) || field_type < CLV_CELL_DATA || field_type > CLV_Y_FACE_DATA)
    return set_error(CLV_ERR_ARG, std::string(who) + ": field and buffer required, face CLV_FACE_*, depth 1 or 2, field_type CLV_*_DATA");
  const int x_inc = (field_type == CLV_VERTEX_DATA || field_type == CLV_X_FACE_DATA) ? 1 : 0;
  const int y_inc = (field_type == CLV_VERTEX_DATA || field_type == CLV_Y_FACE_DATA) ? 1 : 0;
  if (offset < 0 || offset + message_extent(c, face, x_inc, y_inc, depth) > (int64_t)buffer_len)
    return set_error(CLV_ERR_ARG, std::string(who) + ": the message does not fit in buffer_len");
  // ghosts written into the geometry arrays: read them from now on
  if (unpack && (field == c->d.f[CLV_VOLUME] || field == c->d.f[CLV_XAREA] || field == c->d.f[CLV_YAREA])) c->d.geom_mesh = 0;
  return k_pack_array(c, field, field_type, face, depth, buffer, offset, unpack);
}

clv_status clv_pack_message(clv_chunk *c, const double *field, int field_type, int face, int depth, double *buffer, size_t buffer_len,
                            int64_t offset) {
  return pack_array(c, const_cast<double *>(field), field_type, face, depth, buffer, buffer_len, offset, 0, "clv_pack_message");
}

clv_status clv_unpack_message(clv_chunk *c, double *field, int field_type, int face, int depth, const double *buffer, size_t buffer_len,
                              int64_t offset) {
  return pack_array(c, field, field_type, face, depth, const_cast<double *>(buffer), buffer_len, offset, 1, "clv_unpack_message");
}

size_t clv_message_length(const int fields[CLV_NUM_HALO_FIELDS], int depth, int chunk_cells) {
  if (!fields || depth < 1 || chunk_cells < 1) return 0;
  size_t n = 0;
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f) n += fields[f] == 1;
  return n * (size_t)depth * (size_t)(chunk_cells + 5);
}

static clv_status pack_fields(clv_chunk *c, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth, int chunk_cells, int64_t tile_offset,
                              double *buffer, size_t buffer_len, int unpack, const char *who) {
  CHECK_CHUNK(c);
  if (!fields || !buffer || !faces_ok(face) || (depth != 1 && depth != 2))
    return set_error(CLV_ERR_ARG, std::string(who) + ": fields and buffer required, face CLV_FACE_*, depth 1 or 2");
  const int cells = face <= CLV_FACE_RIGHT ? c->d.ny : c->d.nx;
  if (chunk_cells < cells || tile_offset < 0 || tile_offset + (int64_t)cells * depth > (int64_t)chunk_cells * depth)
    return set_error(CLV_ERR_ARG, std::string(who) + ": the tile does not lie inside the chunk's face");
  if (clv_message_length(fields, depth, chunk_cells) > buffer_len)
    return set_error(CLV_ERR_ARG, std::string(who) + ": buffer_len is shorter than clv_message_length");
  return k_pack_fields(c, face, fields, depth, (int64_t)depth * (chunk_cells + 5), tile_offset, buffer, unpack);
}

clv_status clv_pack_fields(clv_chunk *c, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth, int chunk_cells, int64_t tile_offset,
                           double *buffer, size_t buffer_len) {
  return pack_fields(c, face, fields, depth, chunk_cells, tile_offset, buffer, buffer_len, 0, "clv_pack_fields");
}

clv_status clv_unpack_fields(clv_chunk *c, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth, int chunk_cells, int64_t tile_offset,
                             const double *buffer, size_t buffer_len) {
  return pack_fields(c, face, fields, depth, chunk_cells, tile_offset, const_cast<double *>(buffer), buffer_len, 1, "clv_unpack_fields");
}

clv_status clv_field_summary(clv_chunk *c, double out[5]) {
  CHECK_CHUNK(c);
  if (!out) return set_error(CLV_ERR_ARG, "clv_field_summary: null out");
  return k_field_summary(c, out);
}

clv_status clv_chunk_set_summary_order(clv_chunk *c, int order) {
  CHECK_CHUNK(c);
  if (order != CLV_SUMMARY_TREE && order != CLV_SUMMARY_SERIAL) return set_error(CLV_ERR_ARG, "clv_chunk_set_summary_order: bad order");
  c->summary_serial = order == CLV_SUMMARY_SERIAL;
  return CLV_OK;
}

clv_status clv_profile_enable(clv_chunk *c, int on) {
  CHECK_CHUNK(c);
  if (on && !c->prof) c->prof = new Profiler();
  if (!on && c->prof) {
    delete c->prof;
    c->prof = nullptr;
  }
  return CLV_OK;
}

int clv_profile_read(clv_chunk *c, const char **names, int64_t *launches, double *total_ms, int max) {
  if (!c || !c->prof) return 0;
  (void)hipStreamSynchronize(c->stream);
  c->prof->fold(true);
  const int n = (int)c->prof->names.size();
  for (int k = 0; k < n && k < max; ++k) {
    if (names) names[k] = c->prof->names[k].c_str();
    if (launches) launches[k] = c->prof->launches[k];
    if (total_ms) total_ms[k] = c->prof->total_ms[k];
  }
  return n;
}

clv_status clv_profile_reset(clv_chunk *c) {
  CHECK_CHUNK(c);
  if (!c->prof) return CLV_OK;
  (void)hipStreamSynchronize(c->stream);
  c->prof->fold(true);
  for (auto &x : c->prof->launches) x = 0;
  for (auto &x : c->prof->total_ms) x = 0.0;
  return CLV_OK;
}

}  // extern "C"
