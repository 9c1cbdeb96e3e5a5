// gfx950 kernels for the self halo, the two reductions and problem set-up.
#include <mutex>
#include <vector>

#include "clv_internal.h"
#include "eos.h"
#include "geom.h"

namespace clv {
namespace {

#define AT(fid, i, j) d.f[fid][(int64_t)(j) * d.pitch + (i)]
#define M(mid, i) d.m[mid][i]

// ---- update_halo ------------------------------------------------------------------------
// The reflective external halo of src/serial/update_halo.cpp:33-684 for one chunk
// and one tile.  The reference updates each field with four passes, bottom, top,
// left, right; the left/right passes also cover the ghost rows the bottom/top
// passes just wrote, so a corner ghost ends up as sx * sy * (interior mirror).
// With both mirrors resolved per ghost cell, every ghost value is a function of
// interior values only and one launch writes all faces of all requested fields.
// Per-field offsets (including the reference's irregular ones, e.g. the x-face
// top source y_max - k at :530/:573 and the y-face right source x_max - j at
// :636/:679):
struct HaloRule {
  signed char bsrc, bstep;  // bottom: dst row 1 - k <- row bsrc + bstep * k
  signed char tdst, tsrc;   // top:    dst row y_max + tdst + k <- row y_max + tsrc - k
  signed char lsrc;         // left:   dst col 1 - j <- col lsrc + j
  signed char rdst, rsrc;   // right:  dst col x_max + rdst + j <- col x_max + rsrc - j
  signed char xext, yext;   // node-in-x / node-in-y extension of the pass ranges
  signed char sy, sx;       // signs of the bottom/top and left/right reflections
};
// soundspeed's bottom pass reads rows 0 + k (src/serial/update_halo.cpp:305): row 1
// takes the old row 0 and row 0 is rewritten with that same value -> bstep 0.
__constant__ HaloRule kRules[CLV_NUM_HALO_FIELDS] = {
    {2, 1, 2, 1, 2, 2, 1, 0, 0, 1, 1},    // density0
    {2, 1, 2, 1, 2, 2, 1, 0, 0, 1, 1},    // density1
    {2, 1, 2, 1, 2, 2, 1, 0, 0, 1, 1},    // energy0
    {2, 1, 2, 1, 2, 2, 1, 0, 0, 1, 1},    // energy1
    {2, 1, 2, 1, 2, 2, 1, 0, 0, 1, 1},    // pressure
    {2, 1, 2, 1, 2, 2, 1, 0, 0, 1, 1},    // viscosity
    {0, 0, 2, 1, 2, 2, 1, 0, 0, 1, 1},    // soundspeed
    {3, 1, 3, 1, 3, 3, 1, 1, 1, 1, -1},   // xvel0
    {3, 1, 3, 1, 3, 3, 1, 1, 1, 1, -1},   // xvel1
    {3, 1, 3, 1, 3, 3, 1, 1, 1, -1, 1},   // yvel0
    {3, 1, 3, 1, 3, 3, 1, 1, 1, -1, 1},   // yvel1
    {3, 1, 2, 0, 3, 3, 1, 1, 0, 1, -1},   // vol_flux_x
    {3, 1, 3, 1, 3, 2, 0, 0, 1, -1, 1},   // vol_flux_y
    {3, 1, 2, 0, 3, 3, 1, 1, 0, 1, -1},   // mass_flux_x
    {3, 1, 3, 1, 3, 2, 0, 0, 1, -1, 1},   // mass_flux_y
};

struct HaloList {
  int n;
  int ids[CLV_NUM_HALO_FIELDS];
};

// Source coordinate along y for storage row y (identity if y is not a ghost row of
// the bottom/top passes) and the accumulated sign.
__device__ __forceinline__ int halo_src_y(const HaloRule &r, int y, int ny, int depth, double &sgn) {
  if (y <= 1) {  // bottom band rows 1 - k, k < depth
    sgn *= r.sy;
    return r.bsrc + r.bstep * (1 - y);
  }
  if (y >= ny + r.tdst) {  // top band rows y_max + tdst + k
    sgn *= r.sy;
    return ny + r.tsrc - (y - ny - r.tdst);
  }
  return y;
}
__device__ __forceinline__ int halo_src_x(const HaloRule &r, int x, int nx, double &sgn) {
  if (x <= 1) {
    sgn *= r.sx;
    return r.lsrc + (1 - x);
  }
  if (x >= nx + r.rdst) {
    sgn *= r.sx;
    return nx + r.rsrc - (x - nx - r.rdst);
  }
  return x;
}

// Work item t of field fid: t in [0, nrow) walks the bottom/top bands column by
// column (coalesced), t in [nrow, nrow + ncol) the left/right bands of the rows
// between them.  Every ghost cell is written exactly once.
__device__ __forceinline__ int halo_items(const Dev &d, const HaloRule &r, int depth) {
  return (d.nx + r.xext + 2 * depth) + (d.ny + r.yext);
}
__device__ __forceinline__ void halo_item(const Dev &d, const HaloRule &r, int fid, int depth, int t) {
  const int nrow = d.nx + r.xext + 2 * depth;  // columns [2 - depth, nx + xext + depth + 1]
  const int ncol = d.ny + r.yext;              // rows [2, ny + tdst - 1]
  double *a = d.f[fid];
  const int64_t P = d.pitch;
  if (t < nrow) {
    const int x = 2 - depth + t;
    double sx = 1.0;
    const int xs = halo_src_x(r, x, d.nx, sx);
    for (int k = 0; k < depth; ++k) {
      const int yb = 1 - k, yt = d.ny + r.tdst + k;
      double sb = sx, st = sx;
      const int ysb = halo_src_y(r, yb, d.ny, depth, sb);
      const int yst = halo_src_y(r, yt, d.ny, depth, st);
      // sign products are exact (+-1), as are the reference's sequential negations
      a[(int64_t)yb * P + x] = sb * a[(int64_t)ysb * P + xs];
      a[(int64_t)yt * P + x] = st * a[(int64_t)yst * P + xs];
    }
  } else if (t < nrow + ncol) {
    const int y = 2 + (t - nrow);
    for (int k = 0; k < depth; ++k) {
      const int xl = 1 - k, xr = d.nx + r.rdst + k;
      double sl = 1.0, sr = 1.0;
      const int xsl = halo_src_x(r, xl, d.nx, sl);
      const int xsr = halo_src_x(r, xr, d.nx, sr);
      a[(int64_t)y * P + xl] = sl * a[(int64_t)y * P + xsl];
      a[(int64_t)y * P + xr] = sr * a[(int64_t)y * P + xsr];
    }
  }
}

// blockIdx.y = field slot
__global__ __launch_bounds__(256) void update_halo_k(Dev d, HaloList list, int depth) {
  const int fid = list.ids[blockIdx.y];
  const HaloRule r = kRules[fid];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < halo_items(d, r, depth)) halo_item(d, r, fid, depth, t);
}

// Batched: blockIdx.x = problem, blockIdx.y = field slot, blockIdx.z = one of
// kBatchHaloSplit blocks that share the problem's ghost columns and rows of that
// field (4100 on a 2048^2 problem).  Problems not running this step are skipped.
constexpr int kBatchHaloSplit = 8;
__global__ __launch_bounds__(256) void batch_halo_k(const Dev *tab, const BatchState *ps, HaloList list, int depth) {
  const int p = blockIdx.x;
  if (!ps[p].run) return;
  const Dev d = tab[p];
  const int fid = list.ids[blockIdx.y];
  const HaloRule r = kRules[fid];
  const int n = halo_items(d, r, depth);
  for (int t = blockIdx.z * 256 + threadIdx.x; t < n; t += 256 * kBatchHaloSplit) halo_item(d, r, fid, depth, t);
}

// Small grids (nx or ny < 4): a mirror source can itself be a ghost written by an
// earlier pass, so the four passes run as four ordered launches, exactly the
// reference sequence.  pass 0 bottom, 1 top, 2 left, 3 right.
__global__ __launch_bounds__(256) void update_halo_pass_k(Dev d, HaloList list, int depth, int pass) {
  const int fid = list.ids[blockIdx.y];
  const HaloRule r = kRules[fid];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double *a = d.f[fid];
  const int64_t P = d.pitch;
  if (pass < 2) {
    const int x = 2 - depth + t;
    if (x > d.nx + r.xext + depth + 1) return;
    for (int k = 0; k < depth; ++k) {
      if (pass == 0) a[(int64_t)(1 - k) * P + x] = r.sy * a[(int64_t)(r.bsrc + r.bstep * k) * P + x];
      else a[(int64_t)(d.ny + r.tdst + k) * P + x] = r.sy * a[(int64_t)(d.ny + r.tsrc - k) * P + x];
    }
  } else {
    const int y = 2 - depth + t;
    if (y > d.ny + r.yext + depth + 1) return;
    for (int k = 0; k < depth; ++k) {
      if (pass == 2) a[(int64_t)y * P + (1 - k)] = r.sx * a[(int64_t)y * P + (r.lsrc + k)];
      else a[(int64_t)y * P + (d.nx + r.rdst + k)] = r.sx * a[(int64_t)y * P + (d.nx + r.rsrc - k)];
    }
  }
}

// ---- decomposed halos -------------------------------------------------------------------
// Tile-to-tile ghost copies, src/serial/update_tile_halo_kernel.cpp:30-804: the
// ghosts of d on `face` (0 left, 1 right, 2 bottom, 3 top) receive the neighbour
// n's interior layers next to the shared face; node and face data normal to it
// skip the shared boundary node (xext / yext), which both tiles compute.  One
// thread per row (left/right, rows 2-depth .. ny+yext+depth+1, corner ghosts
// included) or column (bottom/top, 2-depth .. nx+xext+depth+1); one field per
// blockIdx.y.  d and n are distinct chunks of equal height (left/right) or
// width (bottom/top).
__global__ __launch_bounds__(256) void tile_halo_k(Dev d, Dev n, HaloList list, int face, int depth) {
  const int fid = list.ids[blockIdx.y];
  const HaloRule r = kRules[fid];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double *a = d.f[fid];
  const double *b = n.f[fid];
  if (face <= 1) {
    const int y = 2 - depth + t;
    if (y > d.ny + r.yext + depth + 1) return;
    for (int j = 0; j < depth; ++j) {
      if (face == 0) a[(int64_t)y * d.pitch + (1 - j)] = b[(int64_t)y * n.pitch + (n.nx + 1 - j)];
      else a[(int64_t)y * d.pitch + (d.nx + 2 + r.xext + j)] = b[(int64_t)y * n.pitch + (2 + r.xext + j)];
    }
  } else {
    const int x = 2 - depth + t;
    if (x > d.nx + r.xext + depth + 1) return;
    for (int k = 0; k < depth; ++k) {
      if (face == 2) a[(int64_t)(1 - k) * d.pitch + x] = b[(int64_t)(n.ny + 1 - k) * n.pitch + x];
      else a[(int64_t)(d.ny + 2 + r.yext + k) * d.pitch + x] = b[(int64_t)(2 + r.yext + k) * n.pitch + x];
    }
  }
}

// One field's exchange message, src/serial/pack_kernel.cpp:27-303: left/right
// messages hold element offset + j + k*depth for rows k in 2-depth ..
// ny+y_inc+depth+1 and layers j < depth, bottom/top messages offset + k + j*depth
// for columns j in 2-depth .. nx+x_inc+depth+1.  Packing reads the interior layers
// next to the face (left and bottom skip the shared boundary node of node / face
// data), unpacking writes the ghost layers.  One thread per row / column.
__device__ __forceinline__ void pack_line(double *a, int64_t pitch, int nx, int ny, int x_inc, int y_inc, int face, int depth, double *buf,
                                          int64_t offset, bool unpack, int t) {
  if (face <= 1) {
    const int k = 2 - depth + t;
    if (k > ny + y_inc + depth + 1) return;
    for (int j = 0; j < depth; ++j) {
      const int64_t index = offset + j + (int64_t)k * depth;
      const int x = face == 0 ? (unpack ? 1 - j : 2 + x_inc + j) : (unpack ? nx + x_inc + 2 + j : nx + 1 - j);
      double &v = a[(int64_t)k * pitch + x];
      if (unpack) v = buf[index];
      else buf[index] = v;
    }
  } else {
    const int j = 2 - depth + t;
    if (j > nx + x_inc + depth + 1) return;
    for (int k = 0; k < depth; ++k) {
      const int64_t index = offset + k + (int64_t)j * depth;
      const int y = face == 2 ? (unpack ? 1 - k : 2 + y_inc + k) : (unpack ? ny + y_inc + 2 + k : ny + 1 - k);
      double &v = a[(int64_t)y * pitch + j];
      if (unpack) v = buf[index];
      else buf[index] = v;
    }
  }
}

// one array of the chunk's layout, field_type 1 cell, 2 vertex, 3 x face, 4 y face
__global__ __launch_bounds__(256) void pack_array_k(double *a, int64_t pitch, int nx, int ny, int x_inc, int y_inc, int face, int depth,
                                                    double *buf, int64_t offset, int unpack) {
  pack_line(a, pitch, nx, ny, x_inc, y_inc, face, depth, buf, offset, unpack != 0, blockIdx.x * blockDim.x + threadIdx.x);
}

// every requested field of a chunk, one per blockIdx.y, each at its own offset
struct PackList {
  int n;
  int ids[CLV_NUM_HALO_FIELDS];
  int64_t off[CLV_NUM_HALO_FIELDS];
};
__global__ __launch_bounds__(256) void pack_fields_k(Dev d, PackList list, int face, int depth, double *buf, int unpack) {
  const int fid = list.ids[blockIdx.y];
  const HaloRule r = kRules[fid];
  pack_line(d.f[fid], d.pitch, d.nx, d.ny, r.xext, r.yext, face, depth, buf, list.off[blockIdx.y], unpack != 0,
            blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- reductions -------------------------------------------------------------------------
constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 2048;

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// First row of a reduction block; the block then takes every gridDim.x-th row, so the
// grid streams one band of gridDim.x rows at a time.  Blocks are dealt to the 8 XCDs
// round-robin (block b on XCD b % 8); XCD x takes the x-th eighth of each band, so the
// node rows j and j + 1 that cell row j reads are read by blocks of one XCD (one L2)
// instead of being fetched twice.
__device__ __forceinline__ int band_row(int G) {
  const int b = blockIdx.x;
  if (G % 8) return b;
  return (b % 8) * (G / 8) + b / 8;
}

// calc_dt: src/serial/calc_dt.cpp:50-77
__global__ __launch_bounds__(RED_THREADS) void calc_dt_k(Dev d, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe,
                                                          double *partials) {
  double dt_min_val0 = g_big;
  for (int j = 2 + band_row(gridDim.x); j < d.ny + 2; j += gridDim.x) {
    const double dsy = M(CLV_CELLDY, j);
    for (int i = 2 + threadIdx.x; i < d.nx + 2; i += RED_THREADS) {
      double dsx = M(CLV_CELLDX, i);
      const double ss = AT(CLV_SOUNDSPEED, i, j);
      double cc = ss * ss;
      // 2 q / rho is +-0 when q is +-0 and rho > 0, and cc + (+-0) is cc
      const double q = AT(CLV_VISCOSITY, i, j);
      if (q != 0.0) {
        cc = cc + 2.0 * q / AT(CLV_DENSITY0, i, j);
      } else {
        const double rho = AT(CLV_DENSITY0, i, j);
        if (!(rho > 0.0)) cc = cc + 2.0 * q / rho;
      }
      cc = fmax(sqrt(cc), g_small);
      double dtct = dtc_safe * fmin(dsx, dsy) / cc;
      double div = 0.0;
      const double vol = geom_volume(d, i, j);
      double dv1 = (AT(CLV_XVEL0, i, j) + AT(CLV_XVEL0, i, j + 1)) * geom_xarea(d, i, j);
      double dv2 = (AT(CLV_XVEL0, i + 1, j) + AT(CLV_XVEL0, i + 1, j + 1)) * geom_xarea(d, i + 1, j);
      div = div + dv2 - dv1;
      double dtut = dtu_safe * 2.0 * vol / fmax(fmax(fabs(dv1), fabs(dv2)), g_small * vol);
      dv1 = (AT(CLV_YVEL0, i, j) + AT(CLV_YVEL0, i + 1, j)) * geom_yarea(d, i, j);
      dv2 = (AT(CLV_YVEL0, i, j + 1) + AT(CLV_YVEL0, i + 1, j + 1)) * geom_yarea(d, i, j + 1);
      div = div + dv2 - dv1;
      double dtvt = dtv_safe * 2.0 * vol / fmax(fmax(fabs(dv1), fabs(dv2)), g_small * vol);
      // div only meets `div < -g_small`: a +-0 div compares the same divided or not
      if (div != 0.0) div = div / (2.0 * vol);
      double dtdivt;
      if (div < -g_small) {
        dtdivt = dtdiv_safe * (-1.0 / div);
      } else {
        dtdivt = g_big;
      }
      double mins = fmin(dtct, fmin(dtut, fmin(dtvt, fmin(dtdivt, g_big))));
      dt_min_val0 = fmin(mins, dt_min_val0);
    }
  }
  __shared__ double lds[RED_THREADS / 64];
  double v = wave_min(dt_min_val0);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lds[0];
    for (int w = 1; w < RED_THREADS / 64; ++w) m = fmin(m, lds[w]);
    partials[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(RED_THREADS) void reduce_min_k(const double *partials, int n, double *out) {
  double m = g_big;
  for (int b = threadIdx.x; b < n; b += RED_THREADS) m = fmin(m, partials[b]);
  __shared__ double lds[RED_THREADS / 64];
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = lds[0];
    for (int w = 1; w < RED_THREADS / 64; ++w) r = fmin(r, lds[w]);
    out[0] = r;
  }
}

// field_summary: src/serial/field_summary.cpp:76-92.  The per-cell terms keep the
// reference's expressions; the sum order is a fixed tree (thread stride within a
// row, rows by block, wave xor-shuffle, waves, blocks in index order), so results
// are reproducible run to run and differ from the serial order only by rounding.
__global__ __launch_bounds__(RED_THREADS) void field_summary_k(Dev d, double *partials) {
  double vol = 0.0, mass = 0.0, ie = 0.0, ke = 0.0, press = 0.0;
  for (int k = 2 + band_row(gridDim.x); k < d.ny + 2; k += gridDim.x) {
    for (int j = 2 + threadIdx.x; j < d.nx + 2; j += RED_THREADS) {
      double vsqrd = 0.0;
      for (int kv = k; kv <= k + 1; ++kv) {
        for (int jv = j; jv <= j + 1; ++jv) {
          const double u = AT(CLV_XVEL0, jv, kv), v = AT(CLV_YVEL0, jv, kv);
          vsqrd += 0.25 * (u * u + v * v);
        }
      }
      double cell_vol = geom_volume(d, j, k);
      double cell_mass = cell_vol * AT(CLV_DENSITY0, j, k);
      vol += cell_vol;
      mass += cell_mass;
      ie += cell_mass * AT(CLV_ENERGY0, j, k);
      ke += cell_mass * 0.5 * vsqrd;
      press += cell_vol * AT(CLV_PRESSURE, j, k);
    }
  }
  __shared__ double lds[5][RED_THREADS / 64];
  double s[5] = {vol, mass, ie, ke, press};
  for (int q = 0; q < 5; ++q) {
    double v = wave_sum(s[q]);
    if ((threadIdx.x & 63) == 0) lds[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double r = lds[threadIdx.x][0];
    for (int w = 1; w < RED_THREADS / 64; ++w) r += lds[threadIdx.x][w];
    partials[(size_t)threadIdx.x * RED_MAX_BLOCKS + blockIdx.x] = r;
  }
}

// ideal_gas (the predict = false call of src/serial/field_summary.cpp:52) and the sums
// of field_summary_k in one pass: each cell's pressure and sound speed are formed
// with ideal_gas's expressions and written, and the pressure term of the sum uses
// the value just written, so arrays and sums equal ideal_gas_k followed by
// field_summary_k bit for bit; density and energy are read once instead of twice and
// pressure is not read back.
__global__ __launch_bounds__(RED_THREADS) void field_summary_gas_k(Dev d, double *partials) {
  double vol = 0.0, mass = 0.0, ie = 0.0, ke = 0.0, press = 0.0;
  for (int k = 2 + band_row(gridDim.x); k < d.ny + 2; k += gridDim.x) {
    for (int j = 2 + threadIdx.x; j < d.nx + 2; j += RED_THREADS) {
      double vsqrd = 0.0;
      for (int kv = k; kv <= k + 1; ++kv) {
        for (int jv = j; jv <= j + 1; ++jv) {
          const double u = AT(CLV_XVEL0, jv, kv), v = AT(CLV_YVEL0, jv, kv);
          vsqrd += 0.25 * (u * u + v * v);
        }
      }
      const double rho = AT(CLV_DENSITY0, j, k), e = AT(CLV_ENERGY0, j, k);
      double p, c;
      ideal_gas_cell(rho, e, p, c);
      __builtin_nontemporal_store(p, &AT(CLV_PRESSURE, j, k));  // streaming: read next by another launch
      __builtin_nontemporal_store(c, &AT(CLV_SOUNDSPEED, j, k));
      double cell_vol = geom_volume(d, j, k);
      double cell_mass = cell_vol * rho;
      vol += cell_vol;
      mass += cell_mass;
      ie += cell_mass * e;
      ke += cell_mass * 0.5 * vsqrd;
      press += cell_vol * p;
    }
  }
  __shared__ double lds[5][RED_THREADS / 64];
  double s[5] = {vol, mass, ie, ke, press};
  for (int q = 0; q < 5; ++q) {
    double v = wave_sum(s[q]);
    if ((threadIdx.x & 63) == 0) lds[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double r = lds[threadIdx.x][0];
    for (int w = 1; w < RED_THREADS / 64; ++w) r += lds[threadIdx.x][w];
    partials[(size_t)threadIdx.x * RED_MAX_BLOCKS + blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(RED_THREADS) void reduce_sum5_k(const double *partials, int n, double *out) {
  __shared__ double lds[5][RED_THREADS / 64];
  for (int q = 0; q < 5; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < n; b += RED_THREADS) s += partials[(size_t)q * RED_MAX_BLOCKS + b];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) lds[q][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double r = lds[threadIdx.x][0];
    for (int w = 1; w < RED_THREADS / 64; ++w) r += lds[threadIdx.x][w];
    out[threadIdx.x] = r;
  }
}

// The serial model's own summation order (src/serial/field_summary.cpp:76-91: one
// left-to-right sum over the linear cell index, x fastest), reproduced exactly:
// the block computes the per-cell terms of 1024 consecutive cells into LDS, then
// one lane folds them in index order.  The fold is sequential by definition, so
// this runs at a few hundred M cells/s; it exists to reproduce the reference's
// printed summaries bit-for-bit, not for throughput.
constexpr int SER_CHUNK = 1024;
__global__ __launch_bounds__(RED_THREADS) void field_summary_serial_k(Dev d, double *out) {
  __shared__ double terms[5][SER_CHUNK];
  const int64_t width = d.nx, n = (int64_t)d.nx * d.ny;
  double vol = 0.0, mass = 0.0, ie = 0.0, ke = 0.0, press = 0.0;
  for (int64_t base = 0; base < n; base += SER_CHUNK) {
    for (int t = threadIdx.x; t < SER_CHUNK; t += RED_THREADS) {
      const int64_t idx = base + t;
      if (idx >= n) break;
      const int j = 2 + (int)(idx % width);
      const int k = 2 + (int)(idx / width);
      double vsqrd = 0.0;
      for (int kv = k; kv <= k + 1; ++kv) {
        for (int jv = j; jv <= j + 1; ++jv) {
          const double u = AT(CLV_XVEL0, jv, kv), v = AT(CLV_YVEL0, jv, kv);
          vsqrd += 0.25 * (u * u + v * v);
        }
      }
      const double cell_vol = geom_volume(d, j, k);
      const double cell_mass = cell_vol * AT(CLV_DENSITY0, j, k);
      terms[0][t] = cell_vol;
      terms[1][t] = cell_mass;
      terms[2][t] = cell_mass * AT(CLV_ENERGY0, j, k);
      terms[3][t] = cell_mass * 0.5 * vsqrd;
      terms[4][t] = cell_vol * AT(CLV_PRESSURE, j, k);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int m = (int)((n - base) < SER_CHUNK ? (n - base) : SER_CHUNK);
      for (int t = 0; t < m; ++t) {
        vol += terms[0][t];
        mass += terms[1][t];
        ie += terms[2][t];
        ke += terms[3][t];
        press += terms[4][t];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = vol;
    out[1] = mass;
    out[2] = ie;
    out[3] = ke;
    out[4] = press;
  }
}

// ---- HBM probe ------------------------------------------------------------------------
// A plain copy of n doubles (16-byte vector accesses, grid-stride): the
// bandwidth reference for the roofline and the known-size kernel that scales
// the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void stream_copy_k(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n2) {
  // one 16-byte element per thread and many blocks: the access pattern of the
  // streaming kernels (one element per thread) at full occupancy
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n2) b[t] = a[t];
}

// The fused kernels' access pattern without their arithmetic: R arrays read and W
// written per element, one wave64 per 64-column strip marching a segment of rows,
// strips x-fastest and each XCD's workgroups on one contiguous band of tiles, with
// streaming stores (the faster store form: 2 % at 8 + 6 arrays).  Its time is the
// ceiling of a fused kernel with the same arrays.
struct MarchProbe {
  const double *in[9];
  double *out[7];
  int64_t pitch;
  int nx, ny, seg;
};
template <int R, int W>
__global__ __launch_bounds__(256) void march_probe_k(MarchProbe p) {
  const int n = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
  const int xcd = b % 8, k = b / 8, per = n / 8, rem = n % 8;
  const int t = xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
  const int bx = t % gridDim.x, by = t / gridDim.x;
  const int i = 2 + (bx * 4 + (int)(threadIdx.x >> 6)) * 64 + (int)(threadIdx.x & 63);
  if (i > p.nx + 1) return;
  const int j0 = 2 + by * p.seg, j1 = min(j0 + p.seg, p.ny + 2);
  double acc = 0.0;
  for (int j = j0; j < j1; ++j) {
    const int64_t o = (int64_t)j * p.pitch + i;
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < R; ++a) s += p.in[a][o];
#pragma unroll
    for (int w = 0; w < W; ++w) __builtin_nontemporal_store(s + (double)w, &p.out[w][o]);  // as lagrange_k / cm_x_k store
    if (W == 0) acc += s;
  }
  // a shape that writes nothing keeps its loads through a store that the zero-filled
  // inputs never reach
  if (W == 0 && acc == 1.0) const_cast<double *>(p.in[0])[(int64_t)j0 * p.pitch + i] = 0.0;
}

// ---- ghost frame copy ---------------------------------------------------------------------
// dst's ghost frame (depth 2 of the node extent: rows 0, 1, ny + 3, ny + 4 and columns
// 0, 1, nx + 3, nx + 4) <- src's; the interior is untouched.
__global__ __launch_bounds__(256) void copy_frame_k(Dev d, int src, int dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = d.nx + 5, nrow = 4 * w, ncol = 4 * (d.ny + 1);
  int i, j;
  if (t < nrow) {
    const int r = t / w;
    i = t % w;
    j = r < 2 ? r : d.ny + 1 + r;  // 0, 1, ny + 3, ny + 4
  } else if (t < nrow + ncol) {
    const int u = t - nrow, q = u / (d.ny + 1);
    j = 2 + u % (d.ny + 1);
    i = q < 2 ? q : d.nx + 1 + q;  // 0, 1, nx + 3, nx + 4
  } else {
    return;
  }
  AT(dst, i, j) = AT(src, i, j);
}

// ---- set-up -----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_zero_k(double *p, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = 0.0;
}

// src/serial/initialise_chunk.cpp:51-75 (one tile: t_left = t_bottom = 1)
__global__ __launch_bounds__(256) void init_mesh_k(Dev d, double xmin, double ymin, double dx, double dy) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int x_min = 1, y_min = 1;
  if (t < d.nx + 5) {
    M(CLV_VERTEXX, t) = xmin + dx * (t - 1 - x_min);
    M(CLV_VERTEXDX, t) = dx;
  }
  if (t < d.ny + 5) {
    M(CLV_VERTEXY, t) = ymin + dy * (t - 1 - y_min);
    M(CLV_VERTEXDY, t) = dy;
  }
  if (t < d.nx + 4) {
    // recomputed vertex pair, same arithmetic as the stored vertexx (:52, :67)
    const double v0 = xmin + dx * (t - 1 - x_min), v1 = xmin + dx * (t + 1 - 1 - x_min);
    M(CLV_CELLX, t) = 0.5 * (v0 + v1);
    M(CLV_CELLDX, t) = dx;
  }
  if (t < d.ny + 4) {
    const double v0 = ymin + dy * (t - 1 - y_min), v1 = ymin + dy * (t + 1 - 1 - y_min);
    M(CLV_CELLY, t) = 0.5 * (v0 + v1);
    M(CLV_CELLDY, t) = dy;
  }
}

// src/serial/initialise_chunk.cpp:78-84
__global__ __launch_bounds__(256) void init_geom_k(Dev d, double dxdy) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx + 4 || j >= d.ny + 4) return;
  AT(CLV_VOLUME, i, j) = dxdy;
  AT(CLV_XAREA, i, j) = M(CLV_CELLDY, j);
  AT(CLV_YAREA, i, j) = M(CLV_CELLDX, i);
}

struct StateTable {
  int n;
  clv_state s[CLV_MAX_STATES];
};

// The last state that covers cell (i, j) (src/serial/generate_chunk.cpp:84-130:
// states are painted in order, so the last hit wins); 0 = background.
__device__ int cell_state(const Dev &d, const StateTable &tab, int i, int j) {
  int hit = 0;
  for (int s = 1; s < tab.n; ++s) {
    const clv_state &st = tab.s[s];
    const double x_cent = st.xmin, y_cent = st.ymin;
    bool h = false;
    if (st.geometry == 1) {
      h = (M(CLV_VERTEXX, i + 1) >= st.xmin && M(CLV_VERTEXX, i) < st.xmax) && (M(CLV_VERTEXY, j + 1) >= st.ymin && M(CLV_VERTEXY, j) < st.ymax);
    } else if (st.geometry == 2) {
      const double cx = M(CLV_CELLX, i), cy = M(CLV_CELLY, j);
      double radius = sqrt((cx - x_cent) * (cx - x_cent) + (cy - y_cent) * (cy - y_cent));
      h = radius <= st.radius;
    } else if (st.geometry == 3) {
      h = M(CLV_VERTEXX, i) == x_cent && M(CLV_VERTEXY, j) == y_cent;
    }
    if (h) hit = s;
  }
  return hit;
}

// Cells: energy/density of the covering state (:77-78, :93-94).
__global__ __launch_bounds__(256) void generate_cells_k(Dev d, StateTable tab) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx + 4 || j >= d.ny + 4) return;
  const int s = cell_state(d, tab, i, j);
  AT(CLV_ENERGY0, i, j) = tab.s[s].energy;
  AT(CLV_DENSITY0, i, j) = tab.s[s].density;
}

// Nodes: a state that covers cell (i, j) also paints nodes (i..i+1, j..j+1)
// (:95-100), so node (x, y) takes the highest state among its (up to) four
// cells; untouched nodes inside [0, nx+4) x [0, ny+4) keep the background
// (:79-80), the last node row/column keeps build_field's zero.
__global__ __launch_bounds__(256) void generate_nodes_k(Dev d, StateTable tab) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= d.nx + 5 || y >= d.ny + 5) return;
  int best = 0;
  for (int cj = y - 1; cj <= y; ++cj) {
    for (int ci = x - 1; ci <= x; ++ci) {
      if (ci < 0 || cj < 0 || ci >= d.nx + 4 || cj >= d.ny + 4) continue;
      const int s = cell_state(d, tab, ci, cj);
      if (s > best) best = s;
    }
  }
  if (best > 0) {
    AT(CLV_XVEL0, x, y) = tab.s[best].xvel;
    AT(CLV_YVEL0, x, y) = tab.s[best].yvel;
  } else if (x < d.nx + 4 && y < d.ny + 4) {
    AT(CLV_XVEL0, x, y) = tab.s[0].xvel;
    AT(CLV_YVEL0, x, y) = tab.s[0].yvel;
  }
}

}  // namespace

clv_status k_copy_frame(clv_chunk *c, int src, int dst) {
  const int n = 4 * (c->d.nx + 5) + 4 * (c->d.ny + 1);
  before_launch(c, "copy_frame");
  hipLaunchKernelGGL(copy_frame_k, dim3((unsigned)((n + 255) / 256), 1, 1), dim3(256, 1, 1), 0, c->stream, c->d, src, dst);
  return after_launch(c, "copy_frame");
}


// ---- host wrappers ------------------------------------------------------------------------
#define LAUNCH_G(name, kern, grid, block, ...)                                 \
  do {                                                                         \
    before_launch(c, name);                                                    \
    hipLaunchKernelGGL(kern, grid, block, 0, c->stream, __VA_ARGS__);          \
    clv_status st_ = after_launch(c, name);                                    \
    if (st_ != CLV_OK) return st_;                                             \
  } while (0)

clv_status k_update_halo(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth) {
  const Dev &d = c->d;
  HaloList list{};
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f)
    if (fields[f] == 1) list.ids[list.n++] = f;
  if (list.n == 0) return CLV_OK;
  if (d.nx >= 4 && d.ny >= 4) {
    const int work = (d.nx + 1 + 2 * depth) + (d.ny + 1);
    LAUNCH_G("update_halo", update_halo_k, dim3((work + 255) / 256, list.n, 1), dim3(256, 1, 1), d, list, depth);
  } else {
    const int work = (d.nx > d.ny ? d.nx : d.ny) + 2 * depth + 2;
    for (int pass = 0; pass < 4; ++pass)
      LAUNCH_G("update_halo", update_halo_pass_k, dim3((work + 255) / 256, list.n, 1), dim3(256, 1, 1), d, list, depth, pass);
  }
  return CLV_OK;
}

static HaloList halo_list(const int fields[CLV_NUM_HALO_FIELDS]) {
  HaloList list{};
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f)
    if (fields[f] == 1) list.ids[list.n++] = f;
  return list;
}

// faces not flagged external keep the ghosts their tile / chunk exchange wrote, so
// the passes run in the reference's order over the flagged faces only
clv_status k_update_halo_faces(clv_chunk *c, const int fields[CLV_NUM_HALO_FIELDS], int depth, const int external[4]) {
  if (external[0] && external[1] && external[2] && external[3]) return k_update_halo(c, fields, depth);
  const Dev &d = c->d;
  const HaloList list = halo_list(fields);
  if (list.n == 0) return CLV_OK;
  const int work = (d.nx > d.ny ? d.nx : d.ny) + 2 * depth + 2;
  const int pass_face[4] = {2, 3, 0, 1};  // passes bottom, top, left, right
  for (int pass = 0; pass < 4; ++pass)
    if (external[pass_face[pass]])
      LAUNCH_G("update_halo", update_halo_pass_k, dim3((work + 255) / 256, list.n, 1), dim3(256, 1, 1), d, list, depth, pass);
  return CLV_OK;
}

clv_status k_tile_halo(clv_chunk *c, const clv_chunk *n, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth) {
  const HaloList list = halo_list(fields);
  if (list.n == 0) return CLV_OK;
  const int work = (face <= 1 ? c->d.ny : c->d.nx) + 2 * depth + 2;
  LAUNCH_G("update_tile_halo", tile_halo_k, dim3((work + 255) / 256, list.n, 1), dim3(256, 1, 1), c->d, n->d, list, face, depth);
  return CLV_OK;
}

clv_status k_pack_array(clv_chunk *c, double *a, int field_type, int face, int depth, double *buf, int64_t offset, int unpack) {
  const int x_inc = (field_type == 2 || field_type == 3) ? 1 : 0, y_inc = (field_type == 2 || field_type == 4) ? 1 : 0;
  const int work = (face <= 1 ? c->d.ny : c->d.nx) + 2 * depth + 2;
  LAUNCH_G("pack_message", pack_array_k, dim3((work + 255) / 256, 1, 1), dim3(256, 1, 1), a, c->d.pitch, c->d.nx, c->d.ny, x_inc, y_inc,
           face, depth, buf, offset, unpack);
  return CLV_OK;
}

clv_status k_pack_fields(clv_chunk *c, int face, const int fields[CLV_NUM_HALO_FIELDS], int depth, int64_t stride, int64_t tile_offset,
                         double *buf, int unpack) {
  PackList list{};
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f)
    if (fields[f] == 1) {
      list.off[list.n] = (int64_t)list.n * stride + tile_offset;
      list.ids[list.n++] = f;
    }
  if (list.n == 0) return CLV_OK;
  const int work = (face <= 1 ? c->d.ny : c->d.nx) + 2 * depth + 2;
  LAUNCH_G("pack_message", pack_fields_k, dim3((work + 255) / 256, list.n, 1), dim3(256, 1, 1), c->d, list, face, depth, buf, unpack);
  return CLV_OK;
}

// every problem of a batch has x_cells, y_cells >= 4: the single-pass form holds
hipError_t batch_launch_halo(hipStream_t s, int nprob, const Dev *tab, const BatchState *ps, const int *ids, int nids, int depth) {
  HaloList list{};
  for (int k = 0; k < nids && k < CLV_NUM_HALO_FIELDS; ++k) list.ids[list.n++] = ids[k];
  if (list.n == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_halo_k, dim3(nprob, list.n, kBatchHaloSplit), dim3(256, 1, 1), 0, s, tab, ps, list, depth);
  return hipGetLastError();
}

// Blocks of a reduction launch (calc_dt_k, field_summary_k, field_summary_gas_k): one
// per row, up to the number of blocks the device holds at once (CUs x resident blocks
// per CU, the least of the three kernels, so that the two summaries tree-sum in the
// same order).  More blocks than fit make a second, partly filled round: 2048 blocks
// at 6 per CU on 256 CUs kept calc_dt_k 25 % above its array traffic's time.
static int red_capacity(int device) {
  static std::mutex mu;
  static int cap[64] = {0};
  if (device < 0 || device >= 64) return RED_MAX_BLOCKS;
  std::lock_guard<std::mutex> lock(mu);
  if (!cap[device]) {
    int cus = 0, least = 1 << 30;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) return RED_MAX_BLOCKS;
    const void *kerns[] = {(const void *)calc_dt_k, (const void *)field_summary_k, (const void *)field_summary_gas_k};
    for (const void *k : kerns) {
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, RED_THREADS, 0) != hipSuccess || n < 1) return RED_MAX_BLOCKS;
      least = n < least ? n : least;
    }
    const int c = cus * least;
    cap[device] = c < RED_MAX_BLOCKS ? c : RED_MAX_BLOCKS;
  }
  return cap[device];
}
static int red_blocks(int ny, int device) {
  const int cap = red_capacity(device);
  return ny < cap ? ny : cap;
}

clv_status k_calc_dt(clv_chunk *c, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe, double *dt_min_val) {
  const int nb = red_blocks(c->d.ny, c->device);
  LAUNCH_G("calc_dt", calc_dt_k, dim3(nb, 1, 1), dim3(RED_THREADS, 1, 1), c->d, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, c->partials);
  LAUNCH_G("calc_dt_final", reduce_min_k, dim3(1, 1, 1), dim3(RED_THREADS, 1, 1), (const double *)c->partials, nb, c->result_dev);
  clv_status st = hip_check(hipMemcpyAsync(c->result_host, c->result_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream), "calc_dt copy");
  if (st != CLV_OK) return st;
  st = hip_check(hipStreamSynchronize(c->stream), "calc_dt sync");
  if (st != CLV_OK) return st;
  *dt_min_val = c->result_host[0];
  return CLV_OK;
}

// with_gas: ideal_gas first (src/serial/field_summary.cpp:52), in the same pass as the
// sums unless the serial summation order is asked for
static clv_status summary(clv_chunk *c, double out[5], bool with_gas) {
  const int nb = red_blocks(c->d.ny, c->device);
  if (c->summary_serial) {
    if (with_gas) {
      clv_status st = k_ideal_gas(c, 0);
      if (st != CLV_OK) return st;
    }
    LAUNCH_G("field_summary_serial", field_summary_serial_k, dim3(1, 1, 1), dim3(RED_THREADS, 1, 1), c->d, c->result_dev);
  } else {
    if (with_gas) LAUNCH_G("field_summary_gas", field_summary_gas_k, dim3(nb, 1, 1), dim3(RED_THREADS, 1, 1), c->d, c->partials);
    else LAUNCH_G("field_summary", field_summary_k, dim3(nb, 1, 1), dim3(RED_THREADS, 1, 1), c->d, c->partials);
    LAUNCH_G("field_summary_final", reduce_sum5_k, dim3(1, 1, 1), dim3(RED_THREADS, 1, 1), (const double *)c->partials, nb,
             c->result_dev);
  }
  clv_status st = hip_check(hipMemcpyAsync(c->result_host, c->result_dev, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "summary copy");
  if (st != CLV_OK) return st;
  st = hip_check(hipStreamSynchronize(c->stream), "summary sync");
  if (st != CLV_OK) return st;
  for (int q = 0; q < 5; ++q) out[q] = c->result_host[q];
  return CLV_OK;
}

clv_status k_field_summary(clv_chunk *c, double out[5]) { return summary(c, out, false); }
clv_status k_field_summary_gas(clv_chunk *c, double out[5]) { return summary(c, out, true); }

}  // namespace clv

extern "C" clv_status clv_bandwidth_probe(size_t bytes, int iters, double *copy_gbps) {
  using namespace clv;
  if (!copy_gbps || bytes < 1024 || iters < 1) return set_error(CLV_ERR_ARG, "clv_bandwidth_probe: bad argument");
  const size_t n2 = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  clv_status st = hip_check(hipMalloc(&a, n2 * 16), "probe alloc");
  if (st == CLV_OK) st = hip_check(hipMalloc(&b, n2 * 16), "probe alloc");
  if (st == CLV_OK) st = hip_check(hipMemset(a, 0, n2 * 16), "probe memset");
  if (st == CLV_OK) st = hip_check(hipEventCreate(&e0), "probe event");
  if (st == CLV_OK) st = hip_check(hipEventCreate(&e1), "probe event");
  const dim3 grid((unsigned)((n2 + 255) / 256)), block(256);
  if (st == CLV_OK) {
    hipLaunchKernelGGL(stream_copy_k, grid, block, 0, nullptr, (const double2 *)a, b, n2);
    (void)hipEventRecord(e0, nullptr);
    for (int k = 0; k < iters; ++k) hipLaunchKernelGGL(stream_copy_k, grid, block, 0, nullptr, (const double2 *)a, b, n2);
    (void)hipEventRecord(e1, nullptr);
    st = hip_check(hipEventSynchronize(e1), "probe sync");
  }
  if (st == CLV_OK) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *copy_gbps = 2.0 * (double)n2 * 16.0 * iters / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return st;
}

extern "C" clv_status clv_march_probe(int x_cells, int y_cells, int reads, int writes, int seg, int iters, double *ms_per_launch) {
  using namespace clv;
  if (!ms_per_launch || x_cells < 1 || y_cells < 1 || seg < 1 || iters < 1)
    return set_error(CLV_ERR_ARG, "clv_march_probe: bad argument");
  typedef void (*Kern)(MarchProbe);
  Kern kern = nullptr;
  // the fused kernels' shapes with the geometry arrays read / derived from the mesh
  if (reads == 7 && writes == 1) kern = march_probe_k<7, 1>;
  if (reads == 8 && writes == 6) kern = march_probe_k<8, 6>;
  if (reads == 7 && writes == 5) kern = march_probe_k<7, 5>;
  if (reads == 6 && writes == 5) kern = march_probe_k<6, 5>;
  if (reads == 4 && writes == 1) kern = march_probe_k<4, 1>;
  if (reads == 5 && writes == 6) kern = march_probe_k<5, 6>;
  if (reads == 5 && writes == 5) kern = march_probe_k<5, 5>;
  if (reads == 7 && writes == 0) kern = march_probe_k<7, 0>;
  if (reads == 7 && writes == 6) kern = march_probe_k<7, 6>;
  if (reads == 4 && writes == 0) kern = march_probe_k<4, 0>;
  if (reads == 4 && writes == 6) kern = march_probe_k<4, 6>;
  if (reads == 7 && writes == 4) kern = march_probe_k<7, 4>;
  if (reads == 6 && writes == 4) kern = march_probe_k<6, 4>;
  if (reads == 5 && writes == 4) kern = march_probe_k<5, 4>;
  if (!kern)
    return set_error(CLV_ERR_ARG,
                     "clv_march_probe: (reads, writes) must be (7, 1), (8, 6), (7, 5), (6, 5), (4, 1), (5, 6), (5, 5), (7, 0), (7, 6), "
                     "(4, 0), (4, 6), (7, 4), (6, 4) or (5, 4)");
  MarchProbe p{};
  p.pitch = pitch_for(x_cells);
  p.nx = x_cells;
  p.ny = y_cells;
  p.seg = seg;
  const size_t n = (size_t)p.pitch * (size_t)(y_cells + 5);
  std::vector<double *> bufs((size_t)(reads + writes), nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  clv_status st = CLV_OK;
  // laid out like a chunk's arena: kLeadPad doubles before element (0, 0), so
  // element (2, j) starts a 128-byte line and every wave's loads are line aligned
  const size_t alloc = (n + 32) * sizeof(double);
  for (double *&q : bufs)
    if (st == CLV_OK && (st = hip_check(hipMalloc(&q, alloc), "march probe alloc")) == CLV_OK)
      st = hip_check(hipMemset(q, 0, alloc), "march probe memset");
  for (int a = 0; a < reads; ++a) p.in[a] = bufs[a] + kLeadPad;
  for (int w = 0; w < writes; ++w) p.out[w] = bufs[reads + w] + kLeadPad;
  if (st == CLV_OK) st = hip_check(hipEventCreate(&e0), "march probe event");
  if (st == CLV_OK) st = hip_check(hipEventCreate(&e1), "march probe event");
  const dim3 grid((unsigned)((x_cells + 255) / 256), (unsigned)((y_cells + seg - 1) / seg)), block(256);
  if (st == CLV_OK) {
    hipLaunchKernelGGL(kern, grid, block, 0, nullptr, p);
    (void)hipEventRecord(e0, nullptr);
    for (int k = 0; k < iters; ++k) hipLaunchKernelGGL(kern, grid, block, 0, nullptr, p);
    (void)hipEventRecord(e1, nullptr);
    st = hip_check(hipEventSynchronize(e1), "march probe sync");
  }
  if (st == CLV_OK) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_launch = (double)ms / iters;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (double *q : bufs)
    if (q) (void)hipFree(q);
  return st;
}

namespace clv {

clv_status k_fill_zero(clv_chunk *c) {
  c->d.geom_mesh = 0;  // volume and the areas become zero
  const Dev &d = c->d;
  const size_t n2 = (size_t)(d.ny + 5) * (size_t)d.pitch;
  for (int f = 0; f < CLV_NUM_FIELDS; ++f) {
    // the lead pad before element (0, 0) belongs to no element; start at (0, 0)
    LAUNCH_G("build_field", fill_zero_k, dim3(2048, 1, 1), dim3(256, 1, 1), d.f[f], n2);
  }
  for (int m = 0; m < CLV_NUM_MESH; ++m) {
    const size_t n = clv_mesh_length(d.nx, d.ny, m);
    LAUNCH_G("build_field", fill_zero_k, dim3(1, 1, 1), dim3(256, 1, 1), d.m[m], n);
  }
  return CLV_OK;
}

// a tile whose first cell is (t_left, t_bottom) of the grid, src/serial/initialise_chunk.cpp:32-37
clv_status k_initialise_chunk(clv_chunk *c, const clv_config *cfg, int t_left, int t_bottom) {
  const Dev &d = c->d;
  double dx = (cfg->xmax - cfg->xmin) / (double)(cfg->x_cells);
  double dy = (cfg->ymax - cfg->ymin) / (double)(cfg->y_cells);
  double xmin = cfg->xmin + dx * (double)(t_left - 1);
  double ymin = cfg->ymin + dy * (double)(t_bottom - 1);
  const int n = (d.nx > d.ny ? d.nx : d.ny) + 5;
  LAUNCH_G("initialise_chunk", init_mesh_k, dim3((n + 255) / 256, 1, 1), dim3(256, 1, 1), d, xmin, ymin, dx, dy);
  const dim3 g((unsigned)((d.nx + 4 + 63) / 64), (unsigned)((d.ny + 4 + 3) / 4), 1);
  LAUNCH_G("initialise_chunk", init_geom_k, g, dim3(64, 4, 1), d, dx * dy);
  // volume = dx * dy, xarea = celldy, yarea = celldx: in mesh mode the kernels derive them
  c->d.geom_mesh = c->geom_mode == CLV_GEOMETRY_MESH;
  return CLV_OK;
}

clv_status k_generate_chunk(clv_chunk *c, const clv_config *cfg) {
  const Dev &d = c->d;
  if (cfg->number_of_states < 1 || cfg->number_of_states > CLV_MAX_STATES) return set_error(CLV_ERR_ARG, "generate_chunk: bad number_of_states");
  StateTable tab{};
  tab.n = cfg->number_of_states;
  for (int s = 0; s < tab.n; ++s) tab.s[s] = cfg->states[s];
  const dim3 gc((unsigned)((d.nx + 4 + 63) / 64), (unsigned)((d.ny + 4 + 3) / 4), 1);
  const dim3 gn((unsigned)((d.nx + 5 + 63) / 64), (unsigned)((d.ny + 5 + 3) / 4), 1);
  LAUNCH_G("generate_chunk", generate_cells_k, gc, dim3(64, 4, 1), d, tab);
  LAUNCH_G("generate_chunk", generate_nodes_k, gn, dim3(64, 4, 1), d, tab);
  return CLV_OK;
}

}  // namespace clv

// ---- field_summary with no allocation ------------------------------------------------------
// clv_field_summary_noalloc: the sums of clv_field_summary over a caller's arrays, for
// callers that cannot allocate device memory.  Its reduction partials and results live
// in `scratch`, one fixed array per device for the life of the process: 5 sums x
// RED_MAX_BLOCKS block partials, then the 5 results.  The reduction never runs more
// than RED_MAX_BLOCKS blocks (they stride over the rows), so every grid fits.  In host
// code `scratch` names no device address: the per-device address comes from
// hipGetSymbolAddress, taken once per device with that device current.
namespace clv {
namespace {
constexpr int kScratchResults = 5 * RED_MAX_BLOCKS;                 // offset of the 5 results
constexpr size_t kScratchDoubles = (size_t)kScratchResults + 8;     // 81,984 bytes
}  // namespace
}  // namespace clv

static __device__ __align__(256) double scratch[clv::kScratchDoubles];

namespace clv {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// the address of this device's `scratch` (looked up on the first call per device)
clv_status scratch_on(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return set_error(CLV_ERR_ARG, "clv_field_summary_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    clv_status st = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "clv_field_summary_noalloc: hipGetSymbolAddress(scratch)");
    if (st != CLV_OK) return st;
    if (!p) return set_error(CLV_ERR_HIP, "clv_field_summary_noalloc: no address for the scratch array");
    g_scratch_ptr[device] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[device];
  return CLV_OK;
}
}  // namespace
}  // namespace clv

extern "C" clv_status clv_field_summary_noalloc(const clv_chunk_view *v, void *hip_stream, double out[5]) {
  using namespace clv;
  if (!v || !out) return set_error(CLV_ERR_ARG, "clv_field_summary_noalloc: null view or out");
  if (v->x_cells < 2 || v->y_cells < 2 || v->pitch < (int64_t)v->x_cells + 5)
    return set_error(CLV_ERR_ARG, "clv_field_summary_noalloc: need x_cells, y_cells >= 2 and pitch >= x_cells + 5");
  const int read[] = {CLV_DENSITY0, CLV_ENERGY0, CLV_PRESSURE, CLV_XVEL0, CLV_YVEL0, CLV_VOLUME};
  for (int f : read)
    if (!v->field[f]) return set_error(CLV_ERR_ARG, "clv_field_summary_noalloc: density0, energy0, pressure, xvel0, yvel0 and volume are required");
  int device = 0;
  clv_status st = hip_check(hipGetDevice(&device), "clv_field_summary_noalloc: hipGetDevice");
  if (st != CLV_OK) return st;
  const int nb = red_blocks(v->y_cells, device);
  // the partials of nb blocks and the results must fit the fixed scratch array
  if ((size_t)nb * 5 > (size_t)kScratchResults || nb > RED_MAX_BLOCKS)
    return set_error(CLV_ERR_ARG, "clv_field_summary_noalloc: the reduction needs more partials than the scratch array holds");
  double *ptr = nullptr;
  if ((st = scratch_on(device, &ptr)) != CLV_OK) return st;
  Dev d{};
  for (int f = 0; f < CLV_NUM_FIELDS; ++f) d.f[f] = v->field[f];
  for (int m = 0; m < CLV_NUM_MESH; ++m) d.m[m] = v->mesh[m];
  d.nx = v->x_cells;
  d.ny = v->y_cells;
  d.pitch = v->pitch;
  d.geom_mesh = 0;  // the caller's volume array
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  hipLaunchKernelGGL(field_summary_k, dim3(nb, 1, 1), dim3(RED_THREADS, 1, 1), 0, s, d, ptr);
  hipLaunchKernelGGL(reduce_sum5_k, dim3(1, 1, 1), dim3(RED_THREADS, 1, 1), 0, s, (const double *)ptr, nb, ptr + kScratchResults);
  if ((st = hip_check(hipGetLastError(), "clv_field_summary_noalloc: launch")) != CLV_OK) return st;
  if ((st = hip_check(hipMemcpyAsync(out, ptr + kScratchResults, 5 * sizeof(double), hipMemcpyDeviceToHost, s),
                      "clv_field_summary_noalloc: copy")) != CLV_OK)
    return st;
  return hip_check(hipStreamSynchronize(s), "clv_field_summary_noalloc: sync");
}
