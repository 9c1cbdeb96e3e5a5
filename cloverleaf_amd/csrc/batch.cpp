// Batched form: many independent problems, each a chunk of its own, advanced
// together.  One batch step is the fused step of clv_hydro_step (kernels_fused.hip)
// for every member that is not complete:
//
//   batch_timestep_k  (all members' tiles)   -> per-member timestep minima
//   batch_dt_k        (one block per member) -> dt, time, step, complete on the device
//   batch_lagrange_k  (q recomputed, as the timestep pass does not store it)
//   batch_halo_k      energy1, density1, vol_flux_x/y, xvel1, yvel1, depth 2
//   batch_advec_k     advection pass 1 (advec_cell + advec_mom), dir1
//   batch_halo_k      density1, energy1, xvel1, yvel1, depth 2
//   batch_advec_k     advection pass 2, dir2
//   batch_halo_k      density1, energy1, depth 2
//
// (the advection passes leave the mass fluxes unstored, as the plain fused step does)
//
// with no host synchronisation: dt never leaves the device, so the eight
// launches are captured once as a hipGraph and replayed per step.
//
// Buffer renaming.  The fused step double-buffers the advected fields and ends
// with reset_field as a rename (kernels_fused.hip).  Over one step the renames
// compose to the 0 <-> 1 swap of density, energy, xvel and yvel, so the views a
// step's kernels see depend only on the parity of the step: two tables of three
// stage views per member ([parity][stage][member]) cover every step, and a graph
// per parity is replayed alternately.  A member that completes stops being
// stepped; its host view is reconciled from the number of steps it ran.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "clv_internal.h"

using namespace clv;

namespace {

constexpr int kStages = 3;  // views before advection pass 1, before pass 2, after pass 2
constexpr int kMinSeg = 8, kMaxSeg = 4096;
constexpr int kCheckEvery = 16;  // batch steps between completion checks

void swap_in(Dev &d, int a, int b) { std::swap(d.f[a], d.f[b]); }
Dev after_reset(Dev d) {
  swap_in(d, CLV_DENSITY0, CLV_DENSITY1);
  swap_in(d, CLV_ENERGY0, CLV_ENERGY1);
  swap_in(d, CLV_XVEL0, CLV_XVEL1);
  swap_in(d, CLV_YVEL0, CLV_YVEL1);
  return d;
}
Dev after_pass(Dev d) {
  swap_in(d, CLV_DENSITY1, CLV_WORK1);
  swap_in(d, CLV_ENERGY1, CLV_WORK2);
  swap_in(d, CLV_XVEL1, CLV_WORK3);
  swap_in(d, CLV_YVEL1, CLV_WORK4);
  return d;
}
// the same arrays and shape (the geometry flag is checked on its own)
bool same_view(const Dev &a, const Dev &b) {
  return std::memcmp(a.f, b.f, sizeof(a.f)) == 0 && std::memcmp(a.m, b.m, sizeof(a.m)) == 0 && a.nx == b.nx && a.ny == b.ny &&
         a.pitch == b.pitch;
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }
// host mirror of the device strips_of
int strips(int fam, int nx) {
  switch (fam) {
    case CLV_BF_TIMESTEP: return ceil_div(nx, kOwnTimestep);
    case CLV_BF_LAGRANGE: return ceil_div(nx, kOwnLagrange);
    case CLV_BF_CM_X: return ceil_div(nx + 1, kOwnCmX);
    default: return ceil_div(nx + 1, kOwnCmY);
  }
}
// the advection passes' tiles cover node rows
int tile_rows(int fam, int ny) { return (fam == CLV_BF_CM_X || fam == CLV_BF_CM_Y) ? ny + 1 : ny; }
constexpr int kWaves = 4;  // strips per 256-thread block

// the kernel each family is timed with by the autotuner
int probe_kind(int fam) {
  static const int k[CLV_BF_COUNT] = {BK_TIMESTEP, BK_LAGRANGE, BK_CM_X1, BK_CM_Y1};
  return k[fam];
}

// splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): a 64-bit Weyl sequence through a mixing function.
uint64_t splitmix64(uint64_t &s) {
  s += 0x9e3779b97f4a7c15ULL;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
double unit(uint64_t &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }  // [0, 1)

}  // namespace

struct clv_batch {
  int n = 0;
  int device = 0;
  int cus = 0;
  hipStream_t stream = nullptr;
  std::vector<clv_chunk *> chunks;
  std::vector<clv_config> cfgs;
  std::vector<clv_hydro_state> st0;  // member states at creation
  std::vector<Dev> d0;               // member views at creation
  int mesh = 0;                      // every member's geometry from its mesh at creation (Dev::geom_mesh)
  int advect_x0 = 1;
  long long bstep = 0;  // batch steps issued since creation
  int hist_len = 0;
  clv_batch_schedule sched{};
  // device memory
  Dev *tabs = nullptr;  // [2][kStages][n]
  BatchConst *pc = nullptr;
  BatchState *ps = nullptr;
  BatchState *ps_host = nullptr;  // pinned
  double *partials = nullptr;
  double *hist = nullptr;
  std::vector<size_t> part_off;  // per member, in doubles
  BatchItem *items[CLV_BF_COUNT] = {};
  size_t items_cap[CLV_BF_COUNT] = {};
  int nitems[CLV_BF_COUNT] = {};
  BatchItem *probe_items = nullptr;
  size_t probe_cap = 0;
  hipGraphExec_t exec[2] = {nullptr, nullptr};

  const Dev *tab(int parity, int stage) const { return tabs + ((size_t)parity * kStages + stage) * n; }
};

namespace {

void drop_graphs(clv_batch *b) {
  for (auto &e : b->exec) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
}

// The tile list of one family: members in the schedule's order, each member's
// tiles row segment by row segment, strip blocks innermost (the order of the
// plain launch's linear block index).
std::vector<BatchItem> build_items(const clv_batch *b, int fam, const clv_batch_family &fs) {
  std::vector<int> order(b->n);
  std::iota(order.begin(), order.end(), 0);
  if (fs.order == CLV_BATCH_ORDER_LARGEST_FIRST)
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return (long long)b->d0[x].nx * b->d0[x].ny > (long long)b->d0[y].nx * b->d0[y].ny;
    });
  std::vector<BatchItem> v;
  for (int p : order) {
    const int nbx = ceil_div(strips(fam, b->d0[p].nx), kWaves);
    const int nseg = ceil_div(tile_rows(fam, b->d0[p].ny), fs.seg);
    for (int k = 0; k < nseg; ++k)
      for (int bx = 0; bx < nbx; ++bx) v.push_back(BatchItem{(uint16_t)p, (uint16_t)bx, (uint16_t)k, (uint16_t)fs.seg});
  }
  return v;
}

int grid_for(const clv_batch *b, const clv_batch_family &fs, int nitems) {
  if (fs.blocks_per_cu <= 0) return nitems;
  return std::max(1, std::min(nitems, fs.blocks_per_cu * b->cus));
}

clv_status upload_items(BatchItem **buf, size_t *cap, const std::vector<BatchItem> &v) {
  if (v.size() > *cap) {
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    if (hipMalloc((void **)buf, v.size() * sizeof(BatchItem)) != hipSuccess)
      return set_error(CLV_ERR_ALLOC, "batch: tile list allocation failed");
    *cap = v.size();
  }
  return hip_check(hipMemcpy(*buf, v.data(), v.size() * sizeof(BatchItem), hipMemcpyHostToDevice), "batch: tile list upload");
}

clv_status check_schedule(const clv_batch *b, const clv_batch_schedule &s) {
  for (int f = 0; f < CLV_BF_COUNT; ++f) {
    const clv_batch_family &fs = s.fam[f];
    if (fs.seg < kMinSeg || fs.seg > kMaxSeg) return set_error(CLV_ERR_ARG, "batch schedule: seg must be in [8, 4096]");
    if (fs.blocks_per_cu < 0 || fs.blocks_per_cu > 64) return set_error(CLV_ERR_ARG, "batch schedule: blocks_per_cu must be in [0, 64]");
    if (fs.order != CLV_BATCH_ORDER_GIVEN && fs.order != CLV_BATCH_ORDER_LARGEST_FIRST)
      return set_error(CLV_ERR_ARG, "batch schedule: unknown member order");
    for (int p = 0; p < b->n; ++p)
      if (ceil_div(tile_rows(f, b->d0[p].ny), fs.seg) > 65535) return set_error(CLV_ERR_ARG, "batch schedule: too many row segments");
  }
  return CLV_OK;
}

// Tile lists and the per-member constants (the timestep minima count depends on
// the timestep family's rows per tile).
clv_status apply_schedule(clv_batch *b, const clv_batch_schedule &s) {
  clv_status st = check_schedule(b, s);
  if (st != CLV_OK) return st;
  for (int f = 0; f < CLV_BF_COUNT; ++f) {
    const std::vector<BatchItem> v = build_items(b, f, s.fam[f]);
    if ((st = upload_items(&b->items[f], &b->items_cap[f], v)) != CLV_OK) return st;
    b->nitems[f] = (int)v.size();
  }
  std::vector<BatchConst> pc(b->n);
  for (int p = 0; p < b->n; ++p) {
    const clv_config &c = b->cfgs[p];
    BatchConst &k = pc[p];
    k.dtc_safe = c.dtc_safe;
    k.dtu_safe = c.dtu_safe;
    k.dtv_safe = c.dtv_safe;
    k.dtdiv_safe = c.dtdiv_safe;
    k.dtrise = c.dtrise;
    k.dtmax = c.dtmax;
    k.dtmin = c.dtmin;
    k.end_time = c.end_time;
    k.partials = b->partials + b->part_off[p];
    k.dt_hist = b->hist + (size_t)p * b->hist_len;
    k.nparts = strips(CLV_BF_TIMESTEP, b->d0[p].nx) * ceil_div(b->d0[p].ny, s.fam[CLV_BF_TIMESTEP].seg);
    k.end_step = c.end_step;
    k.hist_len = b->hist_len;
  }
  if ((st = hip_check(hipMemcpy(b->pc, pc.data(), pc.size() * sizeof(BatchConst), hipMemcpyHostToDevice), "batch: constants upload")) != CLV_OK)
    return st;
  b->sched = s;
  drop_graphs(b);
  return CLV_OK;
}

clv_status launch_family(clv_batch *b, int kind, int fam, const Dev *tab) {
  const int grid = grid_for(b, b->sched.fam[fam], b->nitems[fam]);
  return hip_check(batch_launch_tiles(kind, b->stream, grid, tab, b->pc, b->ps, b->items[fam], b->nitems[fam], b->mesh), "batch tile launch");
}

clv_status launch_halo(clv_batch *b, const Dev *tab, std::initializer_list<int> ids, int depth) {
  std::vector<int> v(ids);
  return hip_check(batch_launch_halo(b->stream, b->n, tab, b->ps, v.data(), (int)v.size(), depth), "batch halo launch");
}

// One batch step of parity q (driver/hydro.cpp:57-89 with the fused kernels; the
// sequence of fused_timestep_kernels, fused_lagrange and fused_advection).
clv_status emit_step(clv_batch *b, int q) {
  const int ax = q == 0 ? b->advect_x0 : !b->advect_x0;
  const bool x_first = ax != 0;
  clv_status st;
#define TRY_B(x)                      \
  do {                                \
    if ((st = (x)) != CLV_OK) return st; \
  } while (0)
  const Dev *s0 = b->tab(q, 0), *s1 = b->tab(q, 1), *s2 = b->tab(q, 2);
  const int fam1 = x_first ? CLV_BF_CM_X : CLV_BF_CM_Y, fam2 = x_first ? CLV_BF_CM_Y : CLV_BF_CM_X;
  TRY_B(launch_family(b, BK_TIMESTEP, CLV_BF_TIMESTEP, s0));
  TRY_B(hip_check(batch_launch_dt(b->stream, b->n, b->pc, b->ps), "batch dt launch"));
  TRY_B(launch_family(b, BK_LAGRANGE, CLV_BF_LAGRANGE, s0));
  TRY_B(launch_halo(b, s0, {CLV_ENERGY1, CLV_DENSITY1, CLV_VOL_FLUX_X, CLV_VOL_FLUX_Y, CLV_XVEL1, CLV_YVEL1}, 2));
  TRY_B(launch_family(b, x_first ? BK_CM_X1 : BK_CM_Y1, fam1, s0));
  TRY_B(launch_halo(b, s1, {CLV_DENSITY1, CLV_ENERGY1, CLV_XVEL1, CLV_YVEL1}, 2));
  TRY_B(launch_family(b, x_first ? BK_CM_Y2 : BK_CM_X2, fam2, s1));
  TRY_B(launch_halo(b, s2, {CLV_DENSITY1, CLV_ENERGY1}, 2));
#undef TRY_B
  return CLV_OK;
}

clv_status capture(clv_batch *b, int q) {
  hipGraph_t g = nullptr;
  clv_status st = hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "batch: begin capture");
  if (st != CLV_OK) return st;
  const clv_status emitted = emit_step(b, q);
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (emitted != CLV_OK) {
    if (g) (void)hipGraphDestroy(g);
    return emitted;
  }
  if ((st = hip_check(e, "batch: end capture")) != CLV_OK) return st;
  st = hip_check(hipGraphInstantiate(&b->exec[q], g, nullptr, nullptr, 0), "batch: graph instantiate");
  (void)hipGraphDestroy(g);
  return st;
}

// member views after the steps they ran; the members' own chunks are updated
void reconcile(clv_batch *b) {
  for (int p = 0; p < b->n; ++p) {
    const int geom_mesh = b->chunks[p]->d.geom_mesh;  // the chunk's own, possibly cleared since creation
    b->chunks[p]->d = (b->ps_host[p].ran & 1) ? after_reset(b->d0[p]) : b->d0[p];
    b->chunks[p]->d.geom_mesh = geom_mesh;
  }
}

clv_status pull_states(clv_batch *b) {
  clv_status st = hip_check(hipMemcpyAsync(b->ps_host, b->ps, (size_t)b->n * sizeof(BatchState), hipMemcpyDeviceToHost, b->stream),
                            "batch: state copy");
  if (st != CLV_OK) return st;
  return hip_check(hipStreamSynchronize(b->stream), "batch: sync");
}

void free_batch(clv_batch *b) {
  drop_graphs(b);
  for (auto &p : b->items)
    if (p) (void)hipFree(p);
  if (b->probe_items) (void)hipFree(b->probe_items);
  void *bufs[] = {b->tabs, b->pc, b->ps, b->partials, b->hist};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (b->ps_host) (void)hipHostFree(b->ps_host);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace

extern "C" {

clv_status clv_batch_deck(uint64_t seed, int index, char *text, size_t len) {
  if (!text || index < 0) return set_error(CLV_ERR_ARG, "clv_batch_deck: null text or negative index");
  uint64_t s = seed, key = 0;
  for (int k = 0; k <= index; ++k) key = splitmix64(s);  // member streams: successive outputs of the seed's stream
  uint64_t r = key;
  const int side = 256 + 128 * (int)(unit(r) * 15.0);
  double xa = 10.0 * unit(r), xb = 10.0 * unit(r), ya = 10.0 * unit(r), yb = 10.0 * unit(r);
  if (xb < xa) std::swap(xa, xb);
  if (yb < ya) std::swap(ya, yb);
  const double rho = 0.5 + 1.5 * unit(r), e = 1.5 + 2.0 * unit(r);
  const int w = std::snprintf(text, len,
                              "*clover\n"
                              " state 1 density=0.2 energy=1.0\n"
                              " state 2 density=%.17g energy=%.17g geometry=rectangle xmin=%.17g xmax=%.17g ymin=%.17g ymax=%.17g\n"
                              " x_cells=%d\n y_cells=%d\n"
                              " xmin=0.0\n ymin=0.0\n xmax=10.0\n ymax=10.0\n"
                              " initial_timestep=0.04\n timestep_rise=1.5\n max_timestep=0.04\n end_step=100\n"
                              "*endclover\n",
                              rho, e, xa, xb, ya, yb, side, side);
  if (w < 0 || (size_t)w >= len) return set_error(CLV_ERR_ARG, "clv_batch_deck: text buffer too small");
  return CLV_OK;
}

clv_status clv_batch_create(clv_chunk *const *chunks, const clv_config *cfgs, const clv_hydro_state *states, int n, int hist_len,
                            clv_batch **out) {
  if (!chunks || !cfgs || !states || !out) return set_error(CLV_ERR_ARG, "clv_batch_create: null argument");
  if (n < 1 || n > 65535) return set_error(CLV_ERR_ARG, "clv_batch_create: n must be in [1, 65535]");
  if (hist_len < 0) return set_error(CLV_ERR_ARG, "clv_batch_create: negative hist_len");
  *out = nullptr;
  for (int p = 0; p < n; ++p) {
    const clv_chunk *c = chunks[p];
    if (!c) return set_error(CLV_ERR_ARG, "clv_batch_create: null chunk");
    if (!c->owns) return set_error(CLV_ERR_ARG, "clv_batch_create: members must be chunks created by clv_chunk_create");
    if (c->d.nx < 4 || c->d.ny < 4 || c->d.nx > 65000) return set_error(CLV_ERR_ARG, "clv_batch_create: member sizes must be in [4, 65000]");
    if (c->device != chunks[0]->device) return set_error(CLV_ERR_ARG, "clv_batch_create: members on different devices");
    if ((states[p].advect_x != 0) != (states[0].advect_x != 0))
      return set_error(CLV_ERR_ARG, "clv_batch_create: members must share advect_x (step parity)");
    if (cfgs[p].x_cells != c->d.nx || cfgs[p].y_cells != c->d.ny) return set_error(CLV_ERR_ARG, "clv_batch_create: config does not match chunk");
    for (int q = 0; q < p; ++q)
      if (chunks[q] == c) return set_error(CLV_ERR_ARG, "clv_batch_create: a chunk appears twice");
  }
  clv_batch *b = new clv_batch();
  b->n = n;
  b->device = chunks[0]->device;
  b->chunks.assign(chunks, chunks + n);
  b->cfgs.assign(cfgs, cfgs + n);
  b->st0.assign(states, states + n);
  b->advect_x0 = states[0].advect_x ? 1 : 0;
  b->hist_len = hist_len;
  clv_status st = CLV_OK;
  auto fail = [&](clv_status s) {
    free_batch(b);
    return s;
  };
  if ((st = hip_check(hipSetDevice(b->device), "clv_batch_create: set device")) != CLV_OK) return fail(st);
  if ((st = hip_check(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device), "clv_batch_create: CU count")) != CLV_OK)
    return fail(st);
  for (int p = 0; p < n; ++p)
    if ((st = hip_check(hipStreamSynchronize(chunks[p]->stream), "clv_batch_create: member sync")) != CLV_OK) return fail(st);
  if ((st = hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "clv_batch_create: stream")) != CLV_OK) return fail(st);
  for (int p = 0; p < n; ++p) b->d0.push_back(chunks[p]->d);
  b->mesh = 1;
  for (int p = 0; p < n; ++p) b->mesh = b->mesh && chunks[p]->d.geom_mesh;
  // stage views of both parities
  std::vector<Dev> tabs((size_t)2 * kStages * n);
  for (int p = 0; p < n; ++p) {
    Dev s = b->d0[p];
    for (int q = 0; q < 2; ++q) {
      tabs[((size_t)q * kStages + 0) * n + p] = s;
      tabs[((size_t)q * kStages + 1) * n + p] = after_pass(s);
      tabs[((size_t)q * kStages + 2) * n + p] = after_pass(after_pass(s));
      s = after_reset(after_pass(after_pass(s)));
    }
  }
  size_t nparts = 0;
  for (int p = 0; p < n; ++p) {
    b->part_off.push_back(nparts);
    nparts += (size_t)strips(CLV_BF_TIMESTEP, b->d0[p].nx) * ceil_div(b->d0[p].ny, kMinSeg);
  }
  auto dmalloc = [&](void **ptr, size_t bytes) {
    return hipMalloc(ptr, bytes ? bytes : 8) == hipSuccess ? CLV_OK : set_error(CLV_ERR_ALLOC, "clv_batch_create: device allocation failed");
  };
  if ((st = dmalloc((void **)&b->tabs, tabs.size() * sizeof(Dev))) != CLV_OK) return fail(st);
  if ((st = dmalloc((void **)&b->pc, n * sizeof(BatchConst))) != CLV_OK) return fail(st);
  if ((st = dmalloc((void **)&b->ps, n * sizeof(BatchState))) != CLV_OK) return fail(st);
  if ((st = dmalloc((void **)&b->partials, nparts * sizeof(double))) != CLV_OK) return fail(st);
  if ((st = dmalloc((void **)&b->hist, (size_t)n * hist_len * sizeof(double))) != CLV_OK) return fail(st);
  if (hipHostMalloc((void **)&b->ps_host, n * sizeof(BatchState), hipHostMallocDefault) != hipSuccess)
    return fail(set_error(CLV_ERR_ALLOC, "clv_batch_create: pinned allocation failed"));
  if ((st = hip_check(hipMemcpy(b->tabs, tabs.data(), tabs.size() * sizeof(Dev), hipMemcpyHostToDevice), "clv_batch_create: views")) != CLV_OK)
    return fail(st);
  for (int p = 0; p < n; ++p) {
    BatchState &s = b->ps_host[p];
    s = BatchState{};
    s.dt = states[p].dt;
    s.dtold = states[p].dtold;
    s.time = states[p].time;
    s.step = states[p].step;
    s.complete = states[p].complete ? 1 : 0;
    s.run = s.complete ? 0 : 1;
  }
  if ((st = hip_check(hipMemcpy(b->ps, b->ps_host, n * sizeof(BatchState), hipMemcpyHostToDevice), "clv_batch_create: states")) != CLV_OK)
    return fail(st);
  if (hist_len > 0 &&
      (st = hip_check(hipMemset(b->hist, 0, (size_t)n * hist_len * sizeof(double)), "clv_batch_create: history clear")) != CLV_OK)
    return fail(st);
  // the plain launches' segments (measured at 15360^2), one block per tile
  clv_batch_schedule s{};
  static const int seg0[CLV_BF_COUNT] = {16, 16, 8, 128};
  for (int f = 0; f < CLV_BF_COUNT; ++f) s.fam[f] = clv_batch_family{seg0[f], 0, CLV_BATCH_ORDER_GIVEN};
  if ((st = apply_schedule(b, s)) != CLV_OK) return fail(st);
  *out = b;
  return CLV_OK;
}

clv_status clv_batch_destroy(clv_batch *b) {
  if (!b) return CLV_OK;
  (void)hipStreamSynchronize(b->stream);
  free_batch(b);
  return CLV_OK;
}

clv_status clv_batch_get_schedule(const clv_batch *b, clv_batch_schedule *out) {
  if (!b || !out) return set_error(CLV_ERR_ARG, "clv_batch_get_schedule: null argument");
  *out = b->sched;
  return CLV_OK;
}

clv_status clv_batch_set_schedule(clv_batch *b, const clv_batch_schedule *s) {
  if (!b || !s) return set_error(CLV_ERR_ARG, "clv_batch_set_schedule: null argument");
  clv_status st = hip_check(hipStreamSynchronize(b->stream), "clv_batch_set_schedule: sync");
  if (st != CLV_OK) return st;
  return apply_schedule(b, *s);
}

clv_status clv_batch_autotune(clv_batch *b, int reps, clv_batch_schedule *out, double family_ms[CLV_BF_COUNT]) {
  if (!b) return set_error(CLV_ERR_ARG, "clv_batch_autotune: null batch");
  if (reps < 1) reps = 1;
  clv_status st = hip_check(hipSetDevice(b->device), "clv_batch_autotune: set device");
  if (st != CLV_OK) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if ((st = hip_check(hipEventCreate(&e0), "autotune event")) != CLV_OK) return st;
  if ((st = hip_check(hipEventCreate(&e1), "autotune event")) != CLV_OK) {
    (void)hipEventDestroy(e0);
    return st;
  }
  // The family kernels only write outputs that the next step recomputes before
  // reading them (p, q, the timestep minima, the 1-fields and fluxes, work
  // arrays), and the views are not renamed here: timing them leaves the
  // members' state unchanged.
  const Dev *tab = b->tab((int)(b->bstep & 1), 0);
  auto time_of = [&](int fam, const clv_batch_family &fs, double *ms) -> clv_status {
    clv_batch_schedule trial = b->sched;
    trial.fam[fam] = fs;
    clv_status s = check_schedule(b, trial);
    if (s != CLV_OK) return s;
    const std::vector<BatchItem> v = build_items(b, fam, fs);
    if ((s = upload_items(&b->probe_items, &b->probe_cap, v)) != CLV_OK) return s;
    const int grid = grid_for(b, fs, (int)v.size());
    const int kind = probe_kind(fam);
    if ((s = hip_check(batch_launch_tiles(kind, b->stream, grid, tab, b->pc, b->ps, b->probe_items, (int)v.size(), b->mesh), "autotune warm-up")) != CLV_OK)
      return s;
    (void)hipEventRecord(e0, b->stream);
    for (int r = 0; r < reps; ++r)
      if ((s = hip_check(batch_launch_tiles(kind, b->stream, grid, tab, b->pc, b->ps, b->probe_items, (int)v.size(), b->mesh), "autotune launch")) != CLV_OK)
        return s;
    (void)hipEventRecord(e1, b->stream);
    if ((s = hip_check(hipEventSynchronize(e1), "autotune sync")) != CLV_OK) return s;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    return CLV_OK;
  };
  static const int segs[] = {8, 16, 32, 64, 128, 256};
  static const int bpcs[] = {0, 1, 2, 4, 8};
  static const int orders[] = {CLV_BATCH_ORDER_GIVEN, CLV_BATCH_ORDER_LARGEST_FIRST};
  clv_batch_schedule best = b->sched;
  double best_ms[CLV_BF_COUNT] = {0};
  for (int f = 0; f < CLV_BF_COUNT && st == CLV_OK; ++f) {
    clv_batch_family cur = best.fam[f];
    double cur_ms = 0;
    if ((st = time_of(f, cur, &cur_ms)) != CLV_OK) break;
    auto consider = [&](clv_batch_family cand) -> clv_status {
      double ms = 0;
      clv_status s = time_of(f, cand, &ms);
      if (s == CLV_OK && ms < cur_ms) {
        cur = cand;
        cur_ms = ms;
      }
      return s;
    };
    for (int sg : segs) {
      clv_batch_family c = cur;
      c.seg = sg;
      if ((st = consider(c)) != CLV_OK) break;
    }
    for (int g : bpcs) {
      if (st != CLV_OK) break;
      clv_batch_family c = cur;
      c.blocks_per_cu = g;
      st = consider(c);
    }
    for (int o : orders) {
      if (st != CLV_OK) break;
      clv_batch_family c = cur;
      c.order = o;
      st = consider(c);
    }
    best.fam[f] = cur;
    best_ms[f] = cur_ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != CLV_OK) return st;
  if ((st = apply_schedule(b, best)) != CLV_OK) return st;
  if (out) *out = best;
  if (family_ms)
    for (int f = 0; f < CLV_BF_COUNT; ++f) family_ms[f] = best_ms[f];
  return CLV_OK;
}

clv_status clv_batch_run(clv_batch *b, int max_steps, int use_graph, int *steps_done) {
  if (!b) return set_error(CLV_ERR_ARG, "clv_batch_run: null batch");
  if (steps_done) *steps_done = 0;
  clv_status st = hip_check(hipSetDevice(b->device), "clv_batch_run: set device");
  if (st != CLV_OK) return st;
  // members must still hold the views the batch left them with
  if ((st = pull_states(b)) != CLV_OK) return st;
  for (int p = 0; p < b->n; ++p) {
    const Dev want = (b->ps_host[p].ran & 1) ? after_reset(b->d0[p]) : b->d0[p];
    if (!same_view(b->chunks[p]->d, want)) return set_error(CLV_ERR_ARG, "clv_batch_run: a member was stepped outside the batch");
    if (b->mesh && !b->chunks[p]->d.geom_mesh)
      return set_error(CLV_ERR_ARG, "clv_batch_run: a member's volume or face areas may have changed since the batch was created");
    if ((st = hip_check(hipStreamSynchronize(b->chunks[p]->stream), "clv_batch_run: member sync")) != CLV_OK) return st;
  }
  // no member runs past its end_step: stop there at the latest
  int remaining = 0;
  for (int p = 0; p < b->n; ++p)
    if (!b->ps_host[p].complete) remaining = std::max(remaining, b->cfgs[p].end_step - b->ps_host[p].step);
  max_steps = std::min(max_steps, remaining);
  int done = 0;
  for (int s = 0; s < max_steps; ++s) {
    const int q = (int)(b->bstep & 1);
    if (use_graph) {
      if (!b->exec[q] && (st = capture(b, q)) != CLV_OK) return st;
      if ((st = hip_check(hipGraphLaunch(b->exec[q], b->stream), "clv_batch_run: graph launch")) != CLV_OK) return st;
    } else if ((st = emit_step(b, q)) != CLV_OK) {
      return st;
    }
    ++b->bstep;
    ++done;
    if (done % kCheckEvery == 0 && s + 1 < max_steps) {
      if ((st = pull_states(b)) != CLV_OK) return st;
      bool all = true;
      for (int p = 0; p < b->n; ++p) all = all && b->ps_host[p].complete;
      if (all) break;
    }
  }
  if ((st = pull_states(b)) != CLV_OK) return st;
  reconcile(b);
  if (steps_done) *steps_done = done;
  for (int p = 0; p < b->n; ++p)
    if (b->ps_host[p].failed_step) {
      char msg[128];
      std::snprintf(msg, sizeof msg, "batch: member %d: small timestep at step %d", p, b->ps_host[p].failed_step);
      return set_error(CLV_ERR_SMALL_DT, msg);
    }
  return CLV_OK;
}

clv_status clv_batch_states(clv_batch *b, clv_hydro_state *out) {
  if (!b || !out) return set_error(CLV_ERR_ARG, "clv_batch_states: null argument");
  clv_status st = pull_states(b);
  if (st != CLV_OK) return st;
  for (int p = 0; p < b->n; ++p) {
    const BatchState &s = b->ps_host[p];
    clv_hydro_state h = b->st0[p];
    h.step = s.step;
    h.time = s.time;
    h.dt = s.dt;
    h.dtold = s.dtold;
    h.complete = s.complete;
    h.advect_x = (b->advect_x0 ^ (s.ran & 1)) ? 1 : 0;
    if (s.step != b->st0[p].step) dt_location(b->d0[p].nx, &h.dt_control, &h.jdt, &h.kdt);
    out[p] = h;
  }
  return CLV_OK;
}

clv_status clv_batch_dt_history(clv_batch *b, int member, double *dt, int max, int *count) {
  if (!b || !dt || !count) return set_error(CLV_ERR_ARG, "clv_batch_dt_history: null argument");
  if (member < 0 || member >= b->n) return set_error(CLV_ERR_ARG, "clv_batch_dt_history: member out of range");
  clv_status st = pull_states(b);
  if (st != CLV_OK) return st;
  const BatchState &s = b->ps_host[member];
  int k = s.step - (s.failed_step ? 1 : 0);
  k = std::min(std::min(k, b->hist_len), std::max(max, 0));
  *count = k;
  if (k <= 0) return CLV_OK;
  return hip_check(hipMemcpy(dt, b->hist + (size_t)member * b->hist_len, (size_t)k * sizeof(double), hipMemcpyDeviceToHost),
                   "clv_batch_dt_history: copy");
}

}  // extern "C"
