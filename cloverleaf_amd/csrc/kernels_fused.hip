// Fused sweeps for CLV_PLAN_FUSED.
//
// One step of driver/hydro.cpp:57-89 in four single-pass kernels:
//
//   timestep_k  ideal_gas(rho0,e0) + viscosity + calc_dt partial minimum
//               (driver/timestep.cpp:47-78)        writes q (p stays in registers)
//   -- global minimum -> dt on the host --
//   lagrange_k  PdV predictor + ideal_gas + pressure halo + revert, accelerate,
//               PdV corrector, flux_calc (driver/hydro.cpp:63-75)
//                                                   writes u1, v1, rho1, e1, vol_flux_x/y
//   cm_x_k / cm_y_k  one advection sweep: advec_cell (src/serial/advec_cell.cpp),
//               then advec_mom for xvel1 AND yvel1 (the yvel pass reuses the xvel
//               pass's node arrays, src/serial/advec_mom.cpp:81-103, so both
//               velocities share one set); twice per step, x and y in the
//               driver's alternating order   writes rho1, e1, mass_flux_d, u1, v1
//   reset_field: pointer swaps 1 <-> 0.
//
// Every value a kernel writes is computed by the reference expression on the
// same inputs, so the persistent fields are bit-identical to the reference
// plan.  Intermediates the reference keeps in work arrays (pre/post volumes,
// node fluxes and masses, momentum and energy fluxes, the predictor's
// rho1/e1/p) stay in registers; dead stores are not made (revert's copy, the
// predictor's soundspeed, advec_mom's pre_vol, advec_cell's post_vol).
//
// Launch shape.  One wave64 owns a strip of consecutive columns (x-contiguous,
// coalesced) and marches up a segment of rows, so neighbours in y come
// from registers of the previous row and every input element is loaded once.
// Neighbours in x of computed values come from the adjacent lane (wave
// shuffles); the lanes at a strip's edges only supply those values, so strips
// overlap by the stencil width and each lane of the interior writes results.
// Inputs needed at x+1 / x-1 also come from shuffles where the lane ranges
// allow, or from a second (cache-hitting) load.  lagrange_k and cm_x_k store each
// row through LDS so that every store instruction writes whole 128-byte lines
// (BlockRows).  Cell volumes and face areas come from the 1D mesh while the
// chunk's arrays hold what initialise_chunk wrote (geo_*).
//
// In-place hazards: a kernel that updates a field its neighbours' stencils read
// (rho1, e1 in advec_cell; u1, v1 in advec_mom) writes into a work array, and
// the host swaps the two pointers afterwards.
#include "clv_internal.h"
#include "eos.h"
#include "lanes.h"

namespace clv {
namespace {

#define AT(fid, i, j) d.f[fid][(int64_t)(j) * d.pitch + (i)]
#define M(mid, i) d.m[mid][i]
// a 1D mesh value at a wave-uniform index, read through the scalar cache (the mesh
// arrays do not change during a launch)
// a store that bypasses the caches' normal retention (the value is read next by
// another kernel, from HBM): on the block-staged whole-line row stores of lagrange_k
// and cm_x_k it takes 1-2 % off those kernels at 15360^2; on the y walk's direct
// stores it costs 1 %, and non-temporal loads cost lagrange_k 5 % (its strips
// re-read each other's edge columns from L2)
#define ST_STREAM(lhs, v) __builtin_nontemporal_store((v), &(lhs))
#define MU(mid, i) ((const __attribute__((address_space(4))) double *)(d.m[mid]))[i]

constexpr int WAVES = 4;       // waves (strips) per 256-thread block
// waves (strips) per block of the plain y advection pass on a wide chunk (at least
// kWideCmY strips; narrower ones keep WAVES): a block's walks keep in step, and 8 in
// step read 8 adjacent strips of each row together (15360^2: 2-5 % faster than 4 a
// block, 2 % faster than 16; 7680^2: 3-7 %; 3840^2: even; 1920^2: 5-10 % slower)
constexpr int kWavesCmY = 8, kWideCmY = 64;

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Reflective mirror of a cell index (src/serial/update_halo.cpp:47-77: ghost
// 1 - k <- 2 + k, ghost max + 2 + k <- max + 1 - k).  Reading a cell field's
// ghost through its mirror gives what the reference's halo update would have
// put there, whatever the ghost layer holds.
__device__ __forceinline__ int mirror_cell(int i, int n) { return i < 2 ? 3 - i : (i > n + 1 ? 2 * n + 3 - i : i); }

// XCD-aware tile order.  Workgroups are dispatched in linear order (x fastest)
// and dealt round-robin to the 8 XCDs, each with its own L2.  Renumbering so that
// XCD k receives the k-th contiguous eighth of the tiles keeps neighbouring
// strips and row segments -- which re-read each other's edge columns and rows --
// in one L2.  The map is a bijection on [0, n).
constexpr int NUM_XCD = 8;
template <bool Y_FASTEST = false>
__device__ __forceinline__ void xcd_tile(int &bx, int &by) {
  const int n = gridDim.x * gridDim.y;
  const int b = blockIdx.y * gridDim.x + blockIdx.x;
  const int xcd = b % NUM_XCD, k = b / NUM_XCD;
  const int per = n / NUM_XCD, rem = n % NUM_XCD;
  const int t = xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
  if (Y_FASTEST) {
    by = t % gridDim.y;
    bx = t / gridDim.y;
  } else {
    bx = t % gridDim.x;
    by = t / gridDim.x;
  }
}

// ---- block-wide row stores ---------------------------------------------------------------
// A 128-byte line written partly by one wave and partly by another costs HBM time
// (a march that writes 6 arrays with line-misaligned waves runs 14% longer than
// the same march aligned, at 15360^2; misaligned reads cost nothing).  So a
// block's four strips of S owned columns each own the consecutive columns
// [X, X + 4 S), X = 2 + 4 S bx, 4 S a multiple of 16: with element (2, j) on a
// line boundary these are whole lines.  The waves put their owned values of a
// row at LDS slot column - X; after a barrier thread t stores slot t to column
// X + t, every store instruction covering whole lines.  With EXTRA_NODE (lagrange:
// strips cover the cells, the last strip's lane S + 1 owns the boundary node column
// nx + 2) that column can land on slot 4 S.
template <int S, bool EXTRA_NODE = false>
struct BlockRows {
  static_assert((WAVES * S) % 16 == 0, "a block's columns must be whole 128-byte lines");
  static constexpr int kSlots = WAVES * 64;
  int x_first, nx;
  __device__ BlockRows(int bx, int nx_) : x_first(2 + WAVES * S * bx), nx(nx_) {}
  __device__ int pos(int i) const { return i - x_first; }  // owned lanes: [0, 4 S]
  __device__ int column(int t) const { return x_first + t; }
  // slot t holds an owned cell / node (x face) value of the row
  __device__ bool cell(int t) const { return t < WAVES * S && x_first + t <= nx + 1; }
  __device__ bool node(int t) const {
    return (t < WAVES * S && x_first + t <= nx + 2) || (EXTRA_NODE && t == WAVES * S && x_first + t == nx + 2);
  }
};

// ---- timestep: ideal_gas + viscosity + calc_dt ----------------------------------------
// Lanes: column x0 + L, x0 = 1 + kOwnTimestep * strip; lanes 1..62 own cells (p of the
// neighbouring columns comes from lanes 0 and 63).  Its one stored array goes straight
// from the lanes (staged through LDS it ran no faster).

// ---- cell volumes and face areas --------------------------------------------------------
// initialise_chunk (src/serial/initialise_chunk.cpp:78-84) sets volume = dx * dy,
// xarea(i, j) = celldy(j) and yarea(i, j) = celldx(i) over i < nx + 4, j < ny + 4 (the
// elements past those, the last column of xarea, the last row of yarea and the row
// padding, keep build_field's zero), and no kernel of a step writes them.  By default
// the kernels read the three arrays, as the reference asks of any optimisation
// (driver/clover_leaf.cpp:31-36: the volume stays an array).  A chunk in
// CLV_GEOMETRY_MESH mode whose arrays still hold what initialise_chunk wrote
// (d.geom_mesh) has them computed instead, per cell from the 1D spacings:
// volume = celldx(i) * celldy(j), which is dx * dy bit for bit on the reference's
// uniform mesh and the tensor-product cell area on any other.  That streams three
// fewer 2D arrays (3 of timestep's 8 array sweeps, 3 of lagrange's 14, one of each
// advection pass's).  The host picks the instantiation (MESH) per launch.
template <bool MESH, class V>
__device__ __forceinline__ double geo_volume(const V &d, int i, int j) {
  if constexpr (MESH) return i < d.nx + 4 && j < d.ny + 4 ? M(CLV_CELLDX, i) * MU(CLV_CELLDY, j) : 0.0;
  else return AT(CLV_VOLUME, i, j);
}
// the same with celldx(i) already in registers (cdx_i; its value does not matter for
// i >= nx + 4): one load of celldy(j) instead of two loads.  Every caller's row j is
// wave-uniform (lanes run along x), so celldy(j) is a scalar load.
template <bool MESH, class V>
__device__ __forceinline__ double geo_volume_x(const V &d, double cdx_i, int i, int j) {
  if constexpr (MESH) return i < d.nx + 4 && j < d.ny + 4 ? cdx_i * MU(CLV_CELLDY, j) : 0.0;
  else return AT(CLV_VOLUME, i, j);
}
template <bool MESH, class V>
__device__ __forceinline__ double geo_xarea(const V &d, int i, int j) {
  if constexpr (MESH) return i < d.nx + 4 && j < d.ny + 4 ? MU(CLV_CELLDY, j) : 0.0;
  else return AT(CLV_XAREA, i, j);
}
// cdx_i: celldx(i) (used for i < nx + 4 only)
template <bool MESH, class V>
__device__ __forceinline__ double geo_yarea(const V &d, double cdx_i, int i, int j) {
  if constexpr (MESH) return i < d.nx + 4 && j < d.ny + 4 ? cdx_i : 0.0;
  else return AT(CLV_YAREA, i, j);
}

// Every kernel below is a tile function -- strip block bx (WAVES strips of one
// 256-thread block) over rows [j0, j1) of one chunk -- called either once per
// block of a plain launch (<name>_k) or in the loop of a persistent batched launch
// over many chunks (batch_<name>_k).  Both run the same instructions per tile.

// q of one cell, src/serial/viscosity.cpp:39-65: u/v 00, 10 (x + 1), 01 (y + 1), 11 the
// cell's nodes; p_w, p_e, p_m, p_p the pressure of its x and y neighbours; cdx, cdx1
// celldx of its column and the next; cdy, cdy2 celldy of its row and of row + 2.  The
// reference sets q = 0 whenever div >= 0, whatever the limiter (:50), so the limiter and
// its seven divisions are formed only where !(div >= 0): in the at-rest part of the box
// (div exactly 0) whole waves skip them.  The same doubles come out: where the branch
// runs, every expression is the reference's.
__device__ __forceinline__ double viscosity_cell(double u00, double u10, double u01, double u11, double v00, double v10, double v01,
                                                 double v11, double p_w, double p_e, double p_m, double p_p, double cdx, double cdx1,
                                                 double cdy, double cdy2, double rho) {
  double ugrad = (u10 + u11) - (u00 + u01);
  double vgrad = (v01 + v11) - (v00 + v10);
  double div = (cdx * (ugrad) + cdy * (vgrad));
  double q = 0.0;
  if (!(div >= 0.0)) {
    double strain2 = 0.5 * (u01 + u11 - u00 - u10) / cdy + 0.5 * (v10 + v11 - v00 - v01) / cdx;
    double pgradx = (p_e - p_w) / (cdx + cdx1);
    double pgrady = (p_p - p_m) / (cdy + cdy2);
    double pgradx2 = pgradx * pgradx;
    double pgrady2 = pgrady * pgrady;
    double limiter = ((0.5 * (ugrad) / cdx) * pgradx2 + (0.5 * (vgrad) / cdy) * pgrady2 + strain2 * pgradx * pgrady) /
                     fmax(pgradx2 + pgrady2, g_small);
    if (!(limiter > 0.0)) {
      double dirx = 1.0;
      if (pgradx < 0.0) dirx = -1.0;
      pgradx = dirx * fmax(g_small, fabs(pgradx));
      double diry = 1.0;
      if (pgradx < 0.0) diry = -1.0;
      pgrady = diry * fmax(g_small, fabs(pgrady));
      double pgrad = sqrt(pgradx * pgradx + pgrady * pgrady);
      double xgrad = fabs(cdx * pgrad / pgradx);
      double ygrad = fabs(cdy * pgrad / pgrady);
      double grad = fmin(xgrad, ygrad);
      double grad2 = grad * grad;
      q = 2.0 * rho * grad2 * limiter * limiter;
    }
  }
  return q;
}

// The wave's minimum goes to part[strip].  WRITE_PC: also store each owned cell's
// pressure and sound speed, as ideal_gas writes them (the per-entry-point path, whose
// later kernels read the arrays; the fused step recomputes them instead).  STORE_Q:
// store each owned cell's q (the fused step leaves it out: lagrange_tile<..., OWN_Q>
// recomputes q from the same operands).
template <bool MESH, bool WRITE_PC = false, bool STORE_Q = true, class V>
__device__ __forceinline__ void timestep_tile(const V &d, int bx, int j0, int j1, double dtc_safe, double dtu_safe, double dtv_safe,
                                              double dtdiv_safe, double *part, int nstrips) {
  const int lane = threadIdx.x & 63;
  const int strip = bx * WAVES + (threadIdx.x >> 6);
  if (strip >= nstrips) return;  // whole wave
  const int x0 = 1 + kOwnTimestep * strip;
  const int i = x0 + lane;
  const int ic = i < d.nx + 4 ? i : d.nx + 4;  // clamped load column (lanes past the chunk)
  const bool own = lane >= 1 && lane <= kOwnTimestep && i <= d.nx + 1;

  const double cdx = M(CLV_CELLDX, ic), cdx1 = M(CLV_CELLDX, ic + 1 <= d.nx + 4 ? ic + 1 : ic);
  // rho0/e0 of ghost cells through their mirrors (the halo of driver/timestep.cpp:53-60)
  const int im = mirror_cell(ic, d.nx);
  // p at rows r-1, r (window) and r+1
  double p_m = pressure_of(AT(CLV_DENSITY0, im, mirror_cell(j0 - 1, d.ny)), AT(CLV_ENERGY0, im, mirror_cell(j0 - 1, d.ny)));
  double rho_c = AT(CLV_DENSITY0, im, j0), e_c = AT(CLV_ENERGY0, im, j0);
  double p_c, c_c;
  ideal_gas_cell(rho_c, e_c, p_c, c_c);
  // node row r velocities (own column and column + 1 via shuffle)
  double u_r = AT(CLV_XVEL0, ic, j0), v_r = AT(CLV_YVEL0, ic, j0);
  double ya_r = geo_yarea<MESH>(d, cdx, ic, j0);
  double dtmin = g_big;
  for (int r = j0; r < j1; ++r) {
    const int rm = mirror_cell(r + 1, d.ny);
    const double rho_p = AT(CLV_DENSITY0, im, rm), e_p = AT(CLV_ENERGY0, im, rm);
    double p_p, c_p;
    ideal_gas_cell(rho_p, e_p, p_p, c_p);
    const double u_r1 = AT(CLV_XVEL0, ic, r + 1), v_r1 = AT(CLV_YVEL0, ic, r + 1);
    const double ya_r1 = geo_yarea<MESH>(d, cdx, ic, r + 1);
    const double xa = geo_xarea<MESH>(d, ic, r);
    const double vol = geo_volume_x<MESH>(d, cdx, ic, r);
    // neighbours along x
    const double p_w = up1(p_c), p_e = dn1(p_c);
    const double u10 = dn1(u_r), u11 = dn1(u_r1), v10 = dn1(v_r), v11 = dn1(v_r1), xa1 = dn1(xa);
    const double u00 = u_r, u01 = u_r1, v00 = v_r, v01 = v_r1;
    const double cdy = MU(CLV_CELLDY, r), cdy2 = MU(CLV_CELLDY, r + 2);
    const double q = viscosity_cell(u00, u10, u01, u11, v00, v10, v01, v11, p_w, p_e, p_m, p_p, cdx, cdx1, cdy, cdy2, rho_c);
    // calc_dt, src/serial/calc_dt.cpp:52-75
    {
      double dsx = cdx;
      double dsy = cdy;
      double cc = c_c * c_c;
      // 2 q / rho is +-0 when q is +-0 and rho > 0, and cc + (+-0) is cc: skip the division
      if (!(q == 0.0 && rho_c > 0.0)) cc = cc + 2.0 * q / rho_c;
      cc = fmax(sqrt(cc), g_small);
      double dtct = dtc_safe * fmin(dsx, dsy) / cc;
      double dv = 0.0;
      double dv1 = (u00 + u01) * xa;
      double dv2 = (u10 + u11) * xa1;
      dv = dv + dv2 - dv1;
      double dtut = dtu_safe * 2.0 * vol / fmax(fmax(fabs(dv1), fabs(dv2)), g_small * vol);
      dv1 = (v00 + v10) * ya_r;
      dv2 = (v01 + v11) * ya_r1;
      dv = dv + dv2 - dv1;
      double dtvt = dtv_safe * 2.0 * vol / fmax(fmax(fabs(dv1), fabs(dv2)), g_small * vol);
      // dv only meets `dv < -g_small`: a +-0 dv compares the same divided or not
      if (dv != 0.0) dv = dv / (2.0 * vol);
      double dtdivt;
      if (dv < -g_small) {
        dtdivt = dtdiv_safe * (-1.0 / dv);
      } else {
        dtdivt = g_big;
      }
      double mins = fmin(dtct, fmin(dtut, fmin(dtvt, fmin(dtdivt, g_big))));
      if (own) dtmin = fmin(mins, dtmin);
    }
    // p is not stored in the fused step: lagrange_tile recomputes it from the same rho0, e0
    if (STORE_Q && own) AT(CLV_VISCOSITY, i, r) = q;
    if (WRITE_PC && own) {
      AT(CLV_PRESSURE, i, r) = p_c;
      AT(CLV_SOUNDSPEED, i, r) = c_c;
    }
    p_m = p_c;
    p_c = p_p;
    c_c = c_p;
    rho_c = rho_p;
    u_r = u_r1;
    v_r = v_r1;
    ya_r = ya_r1;
  }
  dtmin = wave_min(dtmin);
  if (lane == 0) part[strip] = dtmin;
}

// 5 waves per SIMD fit without spilling (96 VGPRs)
template <bool MESH, bool WRITE_PC = false, bool STORE_Q = true>
__global__ __launch_bounds__(256, 5) void timestep_k(Dev d, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe,
                                                  double *partials, int nstrips, int seg) {
  int bx, by;
  xcd_tile(bx, by);
  const int j0 = 2 + seg * by;
  timestep_tile<MESH, WRITE_PC, STORE_Q>(d, bx, j0, min(j0 + seg, d.ny + 2), dtc_safe, dtu_safe, dtv_safe, dtdiv_safe,
                                partials + (size_t)by * nstrips, nstrips);
}

// Minimum of partials[0, n): block b takes the b-th of gridDim.x contiguous
// parts and writes out[b].
__global__ __launch_bounds__(1024) void reduce_min_big_k(const double *partials, int n, double *out) {
  const int part = (n + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * part, hi = min(n, lo + part);
  double m = g_big;
  for (int b = lo + threadIdx.x; b < hi; b += 1024) m = fmin(m, partials[b]);
  __shared__ double lds[16];
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = lds[0];
    for (int w = 1; w < 16; ++w) r = fmin(r, lds[w]);
    out[blockIdx.x] = r;
  }
}

// ---- lagrange: PdV predictor chain, accelerate, PdV corrector, flux_calc ----------------
// Lanes: column x0 + L, x0 = 1 + kOwnLagrange * strip.  Cells are owned by lanes
// 1..kOwnLagrange; nodes by the same lanes plus the boundary node column nx + 2
// (lane kOwnLagrange + 1).  March: cell/node row r from j0 - 1 (or 2) to j1; at
// row r the predictor pressure p'(r), the node velocities of node row r, then the
// corrector of cell row r - 1 and the fluxes.  The block's results go out through
// LDS (BlockRows): its four strips own the columns [X, X + 240), whole lines.
struct PredIn {
  double xa0, xa1, ya0, ya1, u00, u01, u10, u11, v00, v01, v10, v11, vol, rho0, e0, p, q;
};

// src/serial/PdV.cpp:49-63 (predict) then src/serial/ideal_gas.cpp:41 (pressure only: the
// predictor's soundspeed is overwritten before use; WITH_C: and the sound speed, :42-45).
// rho1_out / e1_out: the predictor's density1 and energy1.
template <bool WITH_C = false>
__device__ __forceinline__ double predicted_pressure(const PredIn &a, double dt, double &c, double &rho1_out, double &e1_out) {
  double left_flux = (a.xa0 * (a.u00 + a.u01 + a.u00 + a.u01)) * 0.25 * dt * 0.5;
  double right_flux = (a.xa1 * (a.u10 + a.u11 + a.u10 + a.u11)) * 0.25 * dt * 0.5;
  double bottom_flux = (a.ya0 * (a.v00 + a.v10 + a.v00 + a.v10)) * 0.25 * dt * 0.5;
  double top_flux = (a.ya1 * (a.v01 + a.v11 + a.v01 + a.v11)) * 0.25 * dt * 0.5;
  double total_flux = right_flux - left_flux + top_flux - bottom_flux;
  double volume_change_s = a.vol / (a.vol + total_flux);
  double recip_volume = 1.0 / a.vol;
  double energy_change = (a.p / a.rho0 + a.q / a.rho0) * total_flux * recip_volume;
  double e1 = a.e0 - energy_change;
  double rho1 = a.rho0 * volume_change_s;
  rho1_out = rho1;
  e1_out = e1;
  if constexpr (WITH_C) {
    double p;
    ideal_gas_cell(rho1, e1, p, c);
    return p;
  } else {
    (void)c;
    return (1.4 - 1.0) * rho1 * e1;
  }
}

// cells [j0, jc1).  WRITE_PC: also store the predictor's pressure and sound speed of
// the owned cells (what ideal_gas(predict) leaves in those arrays; the per-entry-point
// path, whose caller may read them -- the fused step never does).  PRED (with
// WRITE_PC): also the predictor's density1 and energy1 of the owned cells into
// rho_pred / e_pred (what PdV(predict) leaves in density1 / energy1 before revert;
// the adjoint's undo log).  OWN_Q: q of the cells the march needs (rows j0 - 1 .. jc1,
// columns of lanes 1..62) computed here, as timestep_tile computes it, instead of read
// (the fused step, whose timestep pass does not store q); its reflective ghosts (the
// halo of driver/timestep.cpp:66-69) by lane shuffle and row carry.
// Lanes: column x0 + L, x0 = kOwnLagrange * strip; lanes 2..1 + kOwnLagrange own
// cells and nodes (lane 2 + kOwnLagrange the boundary node nx + 2); lanes 0, 1 and
// 62, 63 hold the neighbours the stencils reach (q of lane 1 needs p of lane 0).
template <bool MESH, bool WRITE_PC = false, bool PRED = false, bool OWN_Q = false, class V>
__device__ __forceinline__ void lagrange_tile(const V &d, int bx, int j0, int jc1, double dt, int nstrips, double *rho_pred = nullptr,
                                              double *e_pred = nullptr) {
  static_assert(!PRED || WRITE_PC, "the predictor's density and energy are stored with its pressure");
  const int lane = threadIdx.x & 63;
  const int strip = bx * WAVES + (threadIdx.x >> 6);
  // a strip past the chunk runs the march (every wave meets the block's barriers)
  // with clamped loads and owns nothing
  const bool live = strip < nstrips;
  const int nx = d.nx, ny = d.ny;
  const int x0 = kOwnLagrange * strip;
  const int i = x0 + lane;
  const int ic = i < nx + 4 ? i : nx + 4;
  const bool own_cell = live && lane >= 2 && lane <= kOwnLagrange + 1 && i <= nx + 1;
  const bool own_node =
      live && ((lane >= 2 && lane <= kOwnLagrange + 1 && i <= nx + 2) || (lane == kOwnLagrange + 2 && i == nx + 2));
  __shared__ double stage[2][PRED ? 10 : (WRITE_PC ? 8 : 6)][BlockRows<kOwnLagrange, true>::kSlots];
  const BlockRows<kOwnLagrange, true> out(bx, nx);
  const int pos = out.pos(i);
  const int jn1 = jc1 == ny + 2 ? ny + 3 : jc1;    // owned node rows [j0, jn1)
  const double cdx_i = MESH ? M(CLV_CELLDX, ic < nx + 4 ? ic : nx + 3) : 0.0;
  const double halfdt = 0.5 * dt;                  // src/serial/accelerate.cpp:33

  // per-row loads of column ic: node velocities (rows r, r+1), areas, cell data
  auto load_pred = [&](int r, double u_r, double v_r, double u_r1, double v_r1, PredIn &a) {
    a.xa0 = geo_xarea<MESH>(d, ic, r);
    a.xa1 = dn1(a.xa0);
    a.ya0 = geo_yarea<MESH>(d, cdx_i, ic, r);
    a.ya1 = geo_yarea<MESH>(d, cdx_i, ic, r + 1);
    a.u00 = u_r;
    a.u01 = u_r1;
    a.u10 = dn1(u_r);
    a.u11 = dn1(u_r1);
    a.v00 = v_r;
    a.v01 = v_r1;
    a.v10 = dn1(v_r);
    a.v11 = dn1(v_r1);
    a.vol = geo_volume_x<MESH>(d, cdx_i, ic, r);
    if constexpr (!OWN_Q) {  // else carried by the march below
      // ghost rho0 through the mirror (the halo of driver/timestep.cpp:53-60)
      a.rho0 = AT(CLV_DENSITY0, mirror_cell(ic, nx), mirror_cell(r, ny));
      a.e0 = AT(CLV_ENERGY0, ic, r);
      // ideal_gas of the timestep (driver/timestep.cpp:48), recomputed from the
      // same operands with the same expression instead of re-read: bit-identical
      // on the interior cells, the only cells whose predictor result is used (the
      // ghost rows and columns of p' are mirrored from the interior below)
      a.p = pressure_of(a.rho0, a.e0);
      a.q = AT(CLV_VISCOSITY, ic, r);
    }
  };
  // predictor pressure of column i, with the reflective ghost columns taking the
  // interior value (the pressure halo of src/serial/PdV.cpp:121-125)
  auto mirror_cols = [&](double pp) {
    const double from_right = dn1(pp), from_left = up1(pp);
    return i == 1 ? from_right : (i == nx + 2 ? from_left : pp);
  };

  // window: previous cell row (r - 1) and previous node row (r - 1)
  double pp_prev = 0.0, m_prev = 0.0, q_prev = 0.0, xa_prev = 0.0;
  double u0_prev = 0.0, v0_prev = 0.0, u1_prev = 0.0, v1_prev = 0.0;
  double rho0_prev = 0.0, e0_prev = 0.0, vol_prev = 0.0, q0_prev = 0.0, ya_prev = 0.0;
  // node row r velocities; the march starts one row below the segment (the
  // ghost row 1 for the first segment, whose p' is the mirror of row 2)
  int r = j0 - 1;
  double u_r = AT(CLV_XVEL0, ic, r), v_r = AT(CLV_YVEL0, ic, r);
  // OWN_Q: rho0, e0 and p of cell row r and p of row r - 1, rho0 and e0 read through
  // their mirrors (the halo of driver/timestep.cpp:53-60), p recomputed as the
  // timestep's ideal_gas computes it; celldx of the column and the next (clamped
  // past the chunk, where q is not used)
  const int im = mirror_cell(ic, nx);
  const double cdx_q = OWN_Q ? M(CLV_CELLDX, ic < nx + 3 ? ic : nx + 3) : 0.0;
  const double cdx1_q = OWN_Q ? M(CLV_CELLDX, ic + 1 < nx + 3 ? ic + 1 : nx + 3) : 0.0;
  double rho_c = 0.0, e_c = 0.0, p_c = 0.0, p_m = 0.0;
  if constexpr (OWN_Q) {
    const int rm = mirror_cell(r - 1, ny), rc = mirror_cell(r, ny);
    p_m = pressure_of(AT(CLV_DENSITY0, im, rm), AT(CLV_ENERGY0, im, rm));
    rho_c = AT(CLV_DENSITY0, im, rc);
    e_c = AT(CLV_ENERGY0, im, rc);
    p_c = pressure_of(rho_c, e_c);
  }
  for (; r <= jc1; ++r) {
    const double u_r1 = AT(CLV_XVEL0, ic, r + 1);
    const double v_r1 = AT(CLV_YVEL0, ic, r + 1);
    // cell row r inputs
    PredIn a;
    load_pred(r, u_r, v_r, u_r1, v_r1, a);
    if constexpr (OWN_Q) {
      const int rn = mirror_cell(r + 1, ny);
      const double rho_n = AT(CLV_DENSITY0, im, rn), e_n = AT(CLV_ENERGY0, im, rn);
      const double p_n = pressure_of(rho_n, e_n);
      const double cdy = MU(CLV_CELLDY, r), cdy2 = MU(CLV_CELLDY, r + 2 < ny + 3 ? r + 2 : ny + 3);
      double q = viscosity_cell(u_r, dn1(u_r), u_r1, dn1(u_r1), v_r, dn1(v_r), v_r1, dn1(v_r1), up1(p_c), dn1(p_c), p_m, p_n, cdx_q,
                                cdx1_q, cdy, cdy2, rho_c);
      q = mirror_cols(q);           // ghost columns 1, nx + 2 from 2, nx + 1
      if (r == ny + 2) q = q_prev;  // ghost row ny + 2 from ny + 1 (row 1 from row 2 below)
      a.rho0 = rho_c;
      a.e0 = e_c;
      a.p = p_c;
      a.q = q;
      p_m = p_c;
      p_c = p_n;
      rho_c = rho_n;
      e_c = e_n;
    }
    double pp, cp = 0.0, rp = 0.0, ep = 0.0;
    if (r == 1) {
      pp = 0.0;  // set from row 2 below, before any use
    } else if (r == ny + 2) {
      pp = pp_prev;  // ghost row: p'(ny+2) = p'(ny+1)
    } else {
      pp = mirror_cols(predicted_pressure<WRITE_PC>(a, dt, cp, rp, ep));
    }
    if (r == 2) pp_prev = pp;  // ghost row 1 mirrors row 2
    if (OWN_Q && r == 2) q_prev = a.q;
    const double m = a.rho0 * a.vol;
    if (r >= j0) {
      // accelerate at node (i, r), src/serial/accelerate.cpp:43-53
      const double m_w = up1(m), m_w_prev = up1(m_prev);
      double stepbymass_s = halfdt / ((m_w_prev + m_w + m + m_prev) * 0.25);
      const double pp_w = up1(pp), pp_w_prev = up1(pp_prev);
      const double q_w = up1(a.q), q_w_prev = up1(q_prev);
      const double ya_w = up1(a.ya0);
      double u1 = u_r - stepbymass_s * (a.xa0 * (pp - pp_w) + xa_prev * (pp_prev - pp_w_prev));
      double v1 = v_r - stepbymass_s * (a.ya0 * (pp - pp_prev) + ya_w * (pp_w - pp_w_prev));
      u1 = u1 - stepbymass_s * (a.xa0 * (a.q - q_w) + xa_prev * (q_prev - q_w_prev));
      v1 = v1 - stepbymass_s * (a.ya0 * (a.q - q_prev) + ya_w * (q_w - q_w_prev));
      const double u1_e = dn1(u1), v1_e = dn1(v1), u1_e_prev = dn1(u1_prev), v1_e_prev = dn1(v1_prev);
      const double u0_e = dn1(u_r), v0_e = dn1(v_r), u0_e_prev = dn1(u0_prev), v0_e_prev = dn1(v0_prev);
      const double xa_e_prev = dn1(xa_prev);
      if (r > j0) {
        // corrector for cell (i, r - 1), src/serial/PdV.cpp:72-83
        double left_flux = (xa_prev * (u0_prev + u_r + u1_prev + u1)) * 0.25 * dt;
        double right_flux = (xa_e_prev * (u0_e_prev + u0_e + u1_e_prev + u1_e)) * 0.25 * dt;
        double bottom_flux = (ya_prev * (v0_prev + v0_e_prev + v1_prev + v1_e_prev)) * 0.25 * dt;
        double top_flux = (a.ya0 * (v_r + v0_e + v1 + v1_e)) * 0.25 * dt;
        double total_flux = right_flux - left_flux + top_flux - bottom_flux;
        double volume_change_s = vol_prev / (vol_prev + total_flux);
        double recip_volume = 1.0 / vol_prev;
        double energy_change = (pp_prev / rho0_prev + q0_prev / rho0_prev) * total_flux * recip_volume;
        if (own_cell) {
          stage[r & 1][0][pos] = e0_prev - energy_change;
          stage[r & 1][1][pos] = rho0_prev * volume_change_s;
        }
        // flux_calc, src/serial/flux_calc.cpp:39: x-face (i, r - 1)
        if (own_node) stage[r & 1][2][pos] = 0.25 * dt * xa_prev * (u0_prev + u_r + u1_prev + u1);
      }
      if (r < jn1) {
        if (own_node) {
          stage[r & 1][3][pos] = u1;
          stage[r & 1][4][pos] = v1;
        }
        // flux_calc, src/serial/flux_calc.cpp:40: y-face (i, r)
        if (own_cell) stage[r & 1][5][pos] = 0.25 * dt * a.ya0 * (v_r + v0_e + v1 + v1_e);
      }
      if (WRITE_PC && own_cell && r < jc1) {
        stage[r & 1][WRITE_PC ? 6 : 0][pos] = pp;
        stage[r & 1][WRITE_PC ? 7 : 0][pos] = cp;
        if (PRED) {
          stage[r & 1][PRED ? 8 : 0][pos] = rp;
          stage[r & 1][PRED ? 9 : 0][pos] = ep;
        }
      }
      u1_prev = u1;
      v1_prev = v1;
      // the block's rows (r, j0 uniform over the block)
      __syncthreads();
      const int t = threadIdx.x, xs = out.column(t);
      if (r > j0) {
        const int c = r - 1;
        if (out.cell(t)) {
          ST_STREAM(AT(CLV_ENERGY1, xs, c), stage[r & 1][0][t]);
          ST_STREAM(AT(CLV_DENSITY1, xs, c), stage[r & 1][1][t]);
        }
        if (out.node(t)) ST_STREAM(AT(CLV_VOL_FLUX_X, xs, c), stage[r & 1][2][t]);
      }
      if (r < jn1) {
        if (out.node(t)) {
          ST_STREAM(AT(CLV_XVEL1, xs, r), stage[r & 1][3][t]);
          ST_STREAM(AT(CLV_YVEL1, xs, r), stage[r & 1][4][t]);
        }
        if (out.cell(t)) ST_STREAM(AT(CLV_VOL_FLUX_Y, xs, r), stage[r & 1][5][t]);
      }
      if (WRITE_PC && r < jc1 && out.cell(t)) {
        ST_STREAM(AT(CLV_PRESSURE, xs, r), stage[r & 1][WRITE_PC ? 6 : 0][t]);
        ST_STREAM(AT(CLV_SOUNDSPEED, xs, r), stage[r & 1][WRITE_PC ? 7 : 0][t]);
        if (PRED) {
          ST_STREAM(rho_pred[(int64_t)r * d.pitch + xs], stage[r & 1][PRED ? 8 : 0][t]);
          ST_STREAM(e_pred[(int64_t)r * d.pitch + xs], stage[r & 1][PRED ? 9 : 0][t]);
        }
      }
    }
    pp_prev = pp;
    m_prev = m;
    q_prev = a.q;
    xa_prev = a.xa0;
    ya_prev = a.ya0;
    u0_prev = u_r;
    v0_prev = v_r;
    rho0_prev = a.rho0;
    e0_prev = a.e0;
    vol_prev = a.vol;
    q0_prev = a.q;
    u_r = u_r1;
    v_r = v_r1;
  }
  __syncthreads();  // the stage is reused by the block's next tile (batched form)
}

// 4 waves per SIMD fit without spilling (120-126 VGPRs; unbounded, the OWN_Q form with
// the geometry from the mesh took 156 and 3 waves)
template <bool MESH, bool WRITE_PC = false, bool PRED = false, bool OWN_Q = false>
__global__ __launch_bounds__(256, 4) void lagrange_k(Dev d, double dt, int nstrips, int seg, double *rho_pred, double *e_pred) {
  int bx, by;
  xcd_tile(bx, by);
  const int j0 = 2 + seg * by;
  lagrange_tile<MESH, WRITE_PC, PRED, OWN_Q>(d, bx, j0, min(j0 + seg, d.ny + 2), dt, nstrips, rho_pred, e_pred);
}

// ---- advec_cell -------------------------------------------------------------------------
__device__ __forceinline__ double vl_limiter(double sigma, double sigma3, double sigma4, double uw, double dn, double dw) {
  const double one_by_six = 1.0 / 6.0;
  double diffuw = dn - uw;
  double diffdw = dw - dn;
  double wind = 1.0;
  if (diffdw <= 0.0) wind = -1.0;
  double limiter;
  if (diffuw * diffdw > 0.0) {
    limiter = (1.0 - sigma) * wind * fmin(fmin(fabs(diffuw), fabs(diffdw)), one_by_six * (sigma3 * fabs(diffuw) + sigma4 * fabs(diffdw)));
  } else {
    limiter = 0.0;
  }
  return limiter;
}

// Face flux of src/serial/advec_cell.cpp:71-110 given the cell values around the face.
// *_m2, *_m1, *_0, *_p1 are the cells at face - 2, face - 1, face, face + 1 (the +1
// one already clamped by the caller, :77).
__device__ __forceinline__ void cell_face_flux(double vf, double pre_m1, double pre_0, double rho_m2, double rho_m1, double rho_0,
                                               double rho_p1, double e_m2, double e_m1, double e_0, double e_p1, double vd_face,
                                               double vd_m1, double vd_p1, double &mf, double &ef) {
  double pre_d, rho_uw, rho_d, rho_dw, e_uw, e_d, e_dw, vd_dif;
  if (vf > 0.0) {
    rho_uw = rho_m2; rho_d = rho_m1; rho_dw = rho_0;
    e_uw = e_m2; e_d = e_m1; e_dw = e_0;
    pre_d = pre_m1;
    vd_dif = vd_m1;
  } else {
    rho_uw = rho_p1; rho_d = rho_0; rho_dw = rho_m1;
    e_uw = e_p1; e_d = e_0; e_dw = e_m1;
    pre_d = pre_0;
    vd_dif = vd_p1;
  }
  double sigmat = fabs(vf) / pre_d;
  double sigma3 = (1.0 + sigmat) * (vd_face / vd_dif);
  double sigma4 = 2.0 - sigmat;
  double sigmav = sigmat;
  double limiter = vl_limiter(sigmav, sigma3, sigma4, rho_uw, rho_d, rho_dw);
  mf = vf * (rho_d + limiter);
  double sigmam = fabs(mf) / (rho_d * pre_d);
  limiter = vl_limiter(sigmam, sigma3, sigma4, e_uw, e_d, e_dw);
  ef = mf * (e_d + limiter);
}

// ---- advec_mom --------------------------------------------------------------------------
// Limited momentum flux, src/serial/advec_mom.cpp:114-139 (x) / :186-211 (y).
__device__ __forceinline__ double mom_flux_of(double flux, double mass_pre_donor, double width, double width_dif, double v_uw,
                                              double v_dn, double v_dw) {
  double sigma = fabs(flux) / (mass_pre_donor);
  double vdiffuw = v_dn - v_uw;
  double vdiffdw = v_dw - v_dn;
  double limiter = 0.0;
  if (vdiffuw * vdiffdw > 0.0) {
    double auw = fabs(vdiffuw);
    double adw = fabs(vdiffdw);
    double wind = 1.0;
    if (vdiffdw <= 0.0) wind = -1.0;
    limiter = wind * fmin(fmin(width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / width_dif) / 6.0, auw), adw);
  }
  double advec_vel_s = v_dn + (1.0 - sigma) * limiter;
  return advec_vel_s * flux;
}

// x direction, advec_cell then advec_mom of the same sweep in one pass (cell sweep 1
// with mom_sweep 1, cell sweep 2 with mom_sweep 3), so the advected density and the
// mass fluxes reach advec_mom in registers.  advec_mom reads them through the
// reflective halo the driver updates in between (driver/advection.cpp:42-46); here
// the ghost values are taken from the cells and faces they mirror
// (src/serial/update_halo.cpp: density1 ghost col 1 - k <- 2 + k, nx + 2 + k <- nx + 1 - k,
// rows alike; mass_flux_x col 1 - k <- -(3 + k), nx + 3 + k <- -(nx + 1 - k), row 1 <- 3,
// row ny + 2 <- ny), columns by lane shuffles, rows by computing the mirrored row
// (cell rows are independent in an x sweep).  Lanes: column x0 + L, x0 = kOwnCmX * strip
// - 2; cells, faces and nodes owned on lanes 4..3 + kOwnCmX (4..59).  Cell values hold on
// lanes 2..61, face fluxes on 2..62, and advec_mom's stencil narrows that to 4..60.
// March: cell row r (r from j0 - 1), then node row r from cell rows r - 1 and r.  The
// block's results go out through LDS (BlockRows): its four strips own 224 whole-line
// columns.  RESET: the pass also stores its new density, energy and velocities into
// density0, energy0, xvel0 and yvel0 -- reset_field (src/serial/reset_field.cpp:29-42:
// cells [2, nx + 1] x [2, ny + 1], nodes [2, nx + 2] x [2, ny + 2]) -- over the
// same cells and nodes as its own stores.  STORE_MF: store the mass fluxes (the
// per-entry-point path, whose caller may read them; in the fused step nothing reads
// them after the pass, and the next step's pass rewrites them).
template <int SWEEP, bool MESH, bool RESET = false, bool STORE_MF = true, class V, class O>
__device__ __forceinline__ void cm_x_tile(const V &d, int bx, int j0, int j1, int nstrips, O rho_out, O e_out, O u_out, O v_out) {
  constexpr int MOM_SWEEP = SWEEP == 1 ? 1 : 3;
  const int lane = threadIdx.x & 63;
  const int strip = bx * WAVES + (threadIdx.x >> 6);
  const bool live = strip < nstrips;  // past the chunk: march on clamped loads, own nothing
  const int nx = d.nx, ny = d.ny;
  const int i = kOwnCmX * strip - 2 + lane;
  const int ic = i < 0 ? 0 : (i < nx + 4 ? i : nx + 4);
  const bool own = live && lane >= 4 && lane <= 3 + kOwnCmX;
  const bool own_cell = own && i <= nx + 1, own_face = own && i <= nx + 2;
  __shared__ double stage[2][5][BlockRows<kOwnCmX>::kSlots];
  const BlockRows<kOwnCmX> out(bx, nx);
  const int pos = out.pos(i);
  // halo mirrors of the columns: cells, then x faces (negated)
  const int lane_c = lane + (mirror_cell(i, nx) - i);
  const int fm = i <= 1 ? 4 - i : (i >= nx + 3 ? 2 * nx + 4 - i : i);
  const int lane_f = lane + (fm - i);
  const double sgn_f = fm == i ? 1.0 : -1.0;
  const double vd_face = M(CLV_VERTEXDX, ic);
  const double vd_m1 = up1(vd_face), vd_e = dn1(vd_face);
  const double vd_p1 = i + 1 <= nx + 2 ? vd_e : vd_face;
  const double width = M(CLV_CELLDX, ic);
  const double w_m1 = up1(width), w_p1 = dn1(width);
  // advec_cell of cell row `row` at this lane (src/serial/advec_cell.cpp:46-57, 71-125);
  // returns the new density and energy, the face mass flux and the loads advec_mom's
  // mass reuses
  struct CellRow {
    double rho, e, mf, vol, vfy0, vfy1;
  };
  auto cell_row = [&](int row) {
    const double rho = AT(CLV_DENSITY1, ic, row), e = AT(CLV_ENERGY1, ic, row);
    const double vol = geo_volume_x<MESH>(d, width, ic, row), vfx = AT(CLV_VOL_FLUX_X, ic, row);
    const double vfx_e = dn1(vfx);
    double pre, y0 = 0.0, y1 = 0.0;
    if (SWEEP == 1) {
      y0 = AT(CLV_VOL_FLUX_Y, ic, row);
      y1 = AT(CLV_VOL_FLUX_Y, ic, row + 1);
      pre = vol + (vfx_e - vfx + y1 - y0);  // src/serial/advec_cell.cpp:47
    } else {
      pre = vol + vfx_e - vfx;  // :57
    }
    const double rho_m1 = up1(rho), rho_m2 = up2(rho), e_m1 = up1(e), e_m2 = up2(e);
    const double rho_e = dn1(rho), e_e = dn1(e);
    const double rho_p1 = i + 1 <= nx + 2 ? rho_e : rho, e_p1 = i + 1 <= nx + 2 ? e_e : e;  // :77
    const double pre_m1 = up1(pre);
    double mf, ef;
    cell_face_flux(vfx, pre_m1, pre, rho_m2, rho_m1, rho, rho_p1, e_m2, e_m1, e, e_p1, vd_face, vd_m1, vd_p1, mf, ef);
    const double mf_e = dn1(mf), ef_e = dn1(ef);
    // :120-125
    double pre_mass_s = rho * pre;
    double post_mass_s = pre_mass_s + mf - mf_e;
    double post_ener_s = (e * pre_mass_s + ef - ef_e) / post_mass_s;
    double advec_vol_s = pre + vfx - vfx_e;
    return CellRow{post_mass_s / advec_vol_s, post_ener_s, mf, vol, y0, y1};
  };
  double mp_prev = 0.0, mfx_prev = 0.0;
  for (int r = j0 - 1; r < j1; ++r) {
    // cell row r, or the rows its ghosts mirror (density: mirror_cell; mass flux: 1 <- 3,
    // ny + 2 <- ny)
    const int rc = mirror_cell(r, ny), rf = r == 1 ? 3 : (r == ny + 2 ? ny : r);
    const CellRow cr = cell_row(rc);
    const double mf = rf == rc ? cr.mf : cell_row(rf).mf;
    if (r >= j0 && r <= ny + 1) {
      if (own_cell) {
        stage[r & 1][0][pos] = cr.rho;
        stage[r & 1][1][pos] = cr.e;
      }
      if (STORE_MF && own_face) stage[r & 1][2][pos] = cr.mf;
    }
    // advec_mom's view of the cell row: ghost columns mirrored
    const double rho_h = __shfl(cr.rho, lane_c), mfx = sgn_f * __shfl(mf, lane_f);
    // density1 * post_vol of cell (i, r) (src/serial/advec_mom.cpp:47 / :65, :98-99)
    double post;
    if (MOM_SWEEP == 1) post = rc == r ? cr.vol + cr.vfy1 - cr.vfy0 : geo_volume_x<MESH>(d, width, ic, r) + AT(CLV_VOL_FLUX_Y, ic, r + 1) - AT(CLV_VOL_FLUX_Y, ic, r);
    else post = rc == r ? cr.vol : geo_volume_x<MESH>(d, width, ic, r);
    const double mp = rho_h * post;
    if (r >= j0) {
      // node row r (src/serial/advec_mom.cpp:88-149) on the values in registers:
      // node_flux, node_mass_post / pre, the limited momentum flux of node face i
      // (upwind i+2, donor i+1, downwind i when node_flux < 0, else upwind i-1, donor i,
      // downwind i+1) and the velocity update
      const double node_flux = 0.25 * (mfx_prev + mfx + dn1(mfx_prev) + dn1(mfx));
      const double node_mass_post = 0.25 * (mp_prev + mp + up1(mp_prev) + up1(mp));
      const double node_mass_pre = node_mass_post - up1(node_flux) + node_flux;
      const double pre_e = dn1(node_mass_pre);
      const double u = AT(CLV_XVEL1, ic, r), v = AT(CLV_YVEL1, ic, r);
      const double u_w = up1(u), u_e = dn1(u), u_ee = dn2(u);
      const double v_w = up1(v), v_e = dn1(v), v_ee = dn2(v);
      double mu, mv;
      if (node_flux < 0.0) {
        mu = mom_flux_of(node_flux, pre_e, width, w_p1, u_ee, u_e, u);
        mv = mom_flux_of(node_flux, pre_e, width, w_p1, v_ee, v_e, v);
      } else {
        mu = mom_flux_of(node_flux, node_mass_pre, width, w_m1, u_w, u, u_e);
        mv = mom_flux_of(node_flux, node_mass_pre, width, w_m1, v_w, v, v_e);
      }
      const double mu_w = up1(mu), mv_w = up1(mv);
      if (own_face) {  // nodes [2, nx + 2]
        stage[r & 1][3][pos] = (u * node_mass_pre + mu_w - mu) / node_mass_post;
        stage[r & 1][4][pos] = (v * node_mass_pre + mv_w - mv) / node_mass_post;
      }
      // the block's rows (r, j0 uniform over the block)
      __syncthreads();
      const int t = threadIdx.x, xs = out.column(t);
      const int64_t o = (int64_t)r * d.pitch + xs;
      if (r <= ny + 1) {
        if (out.cell(t)) {
          ST_STREAM(rho_out[o], stage[r & 1][0][t]);
          ST_STREAM(e_out[o], stage[r & 1][1][t]);
          if (RESET) {
            ST_STREAM(AT(CLV_DENSITY0, xs, r), stage[r & 1][0][t]);
            ST_STREAM(AT(CLV_ENERGY0, xs, r), stage[r & 1][1][t]);
          }
        }
        if (STORE_MF && out.node(t)) ST_STREAM(AT(CLV_MASS_FLUX_X, xs, r), stage[r & 1][2][t]);
      }
      if (out.node(t)) {
        ST_STREAM(u_out[o], stage[r & 1][3][t]);
        ST_STREAM(v_out[o], stage[r & 1][4][t]);
        if (RESET) {
          ST_STREAM(AT(CLV_XVEL0, xs, r), stage[r & 1][3][t]);
          ST_STREAM(AT(CLV_YVEL0, xs, r), stage[r & 1][4][t]);
        }
      }
    }
    mp_prev = mp;
    mfx_prev = mfx;
  }
  __syncthreads();  // the stage is reused by the block's next tile (batched form)
}

template <int SWEEP, bool MESH, bool RESET = false, bool STORE_MF = true>
__global__ __launch_bounds__(256) void cm_x_k(Dev d, int nstrips, int seg, double *rho_out, double *e_out, double *u_out, double *v_out) {
  int bx, by;
  xcd_tile(bx, by);
  const int j0 = 2 + seg * by;
  cm_x_tile<SWEEP, MESH, RESET, STORE_MF>(d, bx, j0, min(j0 + seg, d.ny + 3), nstrips, rho_out, e_out, u_out, v_out);
}

// y direction, advec_cell then advec_mom of the same sweep in one pass (cell sweep 1
// with mom_sweep 2, cell sweep 2 with mom_sweep 4).  Lanes: column kOwnCmY * strip + L;
// nodes, cells and faces are owned on lanes 2..1 + kOwnCmY (2..63), and lanes 0 and 1
// supply the two columns left of them.  Stores go straight from the lanes: staged
// through LDS (whole lines, 240-column blocks) this walk ran 5-9 % slower.  Each lane walks its column: the cell part (face r needs cells
// r-2 .. r+1; cell r - 1 is updated from faces r - 1 and r) runs four faces ahead of
// the mom part (node row k: node_flux(k) = 0.25 (mfy(i-1,k) + mfy(i,k) + mfy(i-1,k+1) +
// mfy(i,k+1)), the masses of cells (i-1 .. i, k-1 .. k), the flux of face k - 1 and the
// update of node row k - 1; src/serial/advec_mom.cpp:158-222), and keeps the
// advected density and the mass fluxes of the last rows in registers.  advec_mom reads
// them through the halo update between the two sweeps (driver/advection.cpp:42-46,
// src/serial/update_halo.cpp): density1 rows 1 - k <- 2 + k, ny + 2 + k <- ny + 1 - k;
// mass_flux_y rows 1 - k <- -(3 + k), ny + 3 + k <- -(ny + 1 - k); density1 columns
// 1 <- 2, nx + 2 <- nx + 1; mass_flux_y columns 1 <- 3, nx + 2 <- nx.  Here the ghost
// rows are read from the rows they mirror (kept in the registers) and the ghost
// columns from lane shuffles.
// The rows the mom part reads back are kept per lane in LDS rings (registers would
// hold them as shift registers; the walk is issue-bound at 2-3 waves per SIMD, and the
// rings bring it to 4): mass flux rows k-3 .. k+4 at slot row & 7, advected density
// rows k-3 .. k+3 at kRingMf + row % 7, advec_mom's post_vol rows k .. k+4 at
// kRingMf + kRingRho + row % 5.
constexpr int kRingMf = 8, kRingRho = 7, kRingPost = 5, kRing = kRingMf + kRingRho + kRingPost;
// mass flux row x (k or k + 1) through the mirror: the row it is read from
__device__ __forceinline__ int cm_y_mf_src(int x, int ny, double &sgn) {
  int src = x;
  if (x < 2) src = 4 - x;
  else if (x > ny + 2) src = 2 * ny + 4 - x;
  sgn = src == x ? 1.0 : -1.0;
  return src;
}
struct CmRing {
  double *lane_base;  // this lane's slot 0; slot q at lane_base[64 q]
  __device__ void put(int q, double v) const { lane_base[64 * q] = v; }
  __device__ double get(int q) const { return lane_base[64 * q]; }
  __device__ void put_mf(int row, double v) const { put(row & 7, v); }
  __device__ double mf(int row) const { return get(row & 7); }
  __device__ void put_rho(int row, double v) const { put(kRingMf + row % kRingRho, v); }
  __device__ double rho(int row) const { return get(kRingMf + row % kRingRho); }
  __device__ void put_post(int row, double v) const { put(kRingMf + kRingRho + row % kRingPost, v); }
  __device__ double post(int row) const { return get(kRingMf + kRingRho + row % kRingPost); }
};
// RESET and STORE_MF as for cm_x_tile.
template <int SWEEP, bool MESH, bool LOCKSTEP = true, bool RESET = false, int NW = WAVES, bool STORE_MF = true, class V, class O>
__device__ __forceinline__ void cm_y_tile(const V &d, int bx, int j0, int j1, int nstrips, O rho_out, O e_out, O u_out, O v_out) {
  constexpr int MOM_SWEEP = SWEEP == 1 ? 2 : 4;
  const int lane = threadIdx.x & 63;
  const int strip = bx * NW + (threadIdx.x >> 6);
  const bool live = strip < nstrips;  // past the chunk: walk on clamped loads, own nothing (or leave)
  // walks in step: always in the plain launch; in the batched one when the block holds no
  // walk past its problem's last strip (uniform over the block)
  const bool together = LOCKSTEP || (bx + 1) * NW <= nstrips;
  if (!together && !live) return;
  const int nx = d.nx, ny = d.ny;
  const int i = kOwnCmY * strip + lane;
  const int ic = i < nx + 3 ? i : nx + 3;
  const bool own = live && lane >= 2 && lane <= 1 + kOwnCmY;
  const bool own_cell = own && i <= nx + 1, own_node = own && i <= nx + 2;
  const int lane_c = lane + (i == 1 ? 1 : (i == nx + 2 ? -1 : 0));
  const int lane_f = lane + (i == 1 ? 2 : (i == nx + 2 ? -2 : 0));
  const int Fs = j0 - 2 > 2 ? j0 - 2 : 2, Fe = min(j1 + 1, ny + 2);  // faces the cell part walks
  const double cdx_ic = MESH ? M(CLV_CELLDX, ic) : 0.0;                 // ic <= nx + 3
  __shared__ double ring_lds[NW * kRing * 64];
  const CmRing ring{ring_lds + (threadIdx.x >> 6) * (kRing * 64) + lane};
  for (int q = 0; q < kRing; ++q) ring.put(q, 0.0);
  // pre_vol of cell row `row` (src/serial/advec_cell.cpp:139 / :149) from its bottom and
  // top vol_flux_y, and advec_mom's post_vol (src/serial/advec_mom.cpp:56 / :74)
  auto pre_of = [&](int row, double vfy0, double vfy1, double vol, double vfx0, double vfx1) {
    if (SWEEP == 1) return vol + (vfy1 - vfy0 + vfx1 - vfx0);
    return vol + vfy1 - vfy0;
  };
  auto post_from = [&](double vol, double vfx0, double vfx1) { return MOM_SWEEP == 2 ? vol + vfx1 - vfx0 : vol; };
  auto post_of = [&](int c) {
    return post_from(geo_volume_x<MESH>(d, cdx_ic, ic, c), MOM_SWEEP == 2 ? AT(CLV_VOL_FLUX_X, ic, c) : 0.0, MOM_SWEEP == 2 ? AT(CLV_VOL_FLUX_X, ic + 1, c) : 0.0);
  };
  auto pre_load = [&](int row, double vfy0, double &vfy1, double &post) {
    const double vol = geo_volume_x<MESH>(d, cdx_ic, ic, row);
    vfy1 = AT(CLV_VOL_FLUX_Y, ic, row + 1);
    const double vfx0 = SWEEP == 1 ? AT(CLV_VOL_FLUX_X, ic, row) : 0.0, vfx1 = SWEEP == 1 ? AT(CLV_VOL_FLUX_X, ic + 1, row) : 0.0;
    post = post_from(vol, vfx0, vfx1);
    return pre_of(row, vfy0, vfy1, vol, vfx0, vfx1);
  };
  // the loads of face F's step, issued one step ahead (before the step in hand's
  // stores, which the compiler cannot move loads past): the walk is latency-bound
  struct CellLd {
    double vol, vfy1, vfx0, vfx1, vd, rho_p1, e_p1, vd_p1;
  };
  auto cell_ld = [&](int F) {
    double vol = 0.0, vfy1 = 0.0, vfx0 = 0.0, vfx1 = 0.0, vd = 0.0, rho_p1 = 0.0, e_p1 = 0.0, vd_p1 = 0.0;
    if (F <= Fe) {
      const int rp = F + 1 <= ny + 2 ? F + 1 : ny + 2;  // src/serial/advec_cell.cpp:169
      vol = geo_volume_x<MESH>(d, cdx_ic, ic, F);
      vfy1 = AT(CLV_VOL_FLUX_Y, ic, F + 1);
      if (SWEEP == 1) {
        vfx0 = AT(CLV_VOL_FLUX_X, ic, F);
        vfx1 = AT(CLV_VOL_FLUX_X, ic + 1, F);
      }
      vd = MU(CLV_VERTEXDY, F);
      rho_p1 = AT(CLV_DENSITY1, ic, rp);
      e_p1 = AT(CLV_ENERGY1, ic, rp);
      vd_p1 = MU(CLV_VERTEXDY, rp);
    }
    return CellLd{vol, vfy1, vfx0, vfx1, vd, rho_p1, e_p1, vd_p1};
  };
  // cell part: faces Fs .. Fe (cell F - 1 is updated at face F)
  double rho_m2 = AT(CLV_DENSITY1, ic, Fs - 2), rho_m1 = AT(CLV_DENSITY1, ic, Fs - 1), rho_0 = AT(CLV_DENSITY1, ic, Fs);
  double e_m2 = AT(CLV_ENERGY1, ic, Fs - 2), e_m1 = AT(CLV_ENERGY1, ic, Fs - 1), e_0 = AT(CLV_ENERGY1, ic, Fs);
  double vfy_c = AT(CLV_VOL_FLUX_Y, ic, Fs - 1), post_unused;  // vol_flux_y of the face below the cell in hand
  double pre_m1 = pre_load(Fs - 1, vfy_c, vfy_c, post_unused);
  double vd_m1 = MU(CLV_VERTEXDY, Fs - 1);
  double mf_prev = 0.0, ef_prev = 0.0, vf_prev = 0.0;
  // mom part state: node flux and masses of the previous row, the previous face's
  // momentum fluxes, velocities of rows k-2 .. k
  double nf_prev = 0.0, mo_prev = 0.0, mw_prev = 0.0, pre_prev = 0.0, post_prev = 0.0, mu_prev = 0.0, mv_prev = 0.0;
  double u_m2 = 0.0, v_m2 = 0.0, u_m1 = 0.0, v_m1 = 0.0, u_0 = 0.0, v_0 = 0.0;
  // node row k's flux and masses from the history (src/serial/advec_mom.cpp:160, :170-171);
  // post_vol of row k from the history where the cell part passed that row, else loaded
  auto mom_row = [&](int k, double &nf, double &mo, double &mw) {
    double s0, s1;
    const int x0 = cm_y_mf_src(k, ny, s0), x1 = cm_y_mf_src(k + 1, ny, s1);
    const double m0 = s0 * ring.mf(x0), m1 = s1 * ring.mf(x1);
    nf = 0.25 * (up1(m0) + m0 + up1(m1) + m1);
    const double post = k >= Fs && k <= Fe ? ring.post(k) : post_of(k);
    mo = ring.rho(mirror_cell(k, ny)) * post;
    mw = up1(mo);
  };
  CellLd cur = cell_ld(Fs);
  double u_ahead = 0.0, v_ahead = 0.0;  // velocities of row k + 1 for the next mom step
  for (int F = Fs; F <= j1 + 4; ++F) {
    const CellLd nxt = cell_ld(F + 1);
    // the next mom step (k + 1 = F - 3) reads velocity row F - 2; it runs for
    // F - 2 in [j0, j1 + 1], and outside that a row in it is loaded and not used
    const int rv = min(max(F - 2, j0), j1 + 1);
    const double u_n = AT(CLV_XVEL1, ic, rv), v_n = AT(CLV_YVEL1, ic, rv);
    double mf_h = 0.0, rho_h = 0.0, post_h = 0.0;
    if (F <= Fe) {
      const double vf = vfy_c;
      vfy_c = cur.vfy1;
      post_h = post_from(cur.vol, cur.vfx0, cur.vfx1);
      const double pre_0 = pre_of(F, vf, cur.vfy1, cur.vol, cur.vfx0, cur.vfx1);
      const double vd_face = cur.vd, rho_p1 = cur.rho_p1, e_p1 = cur.e_p1, vd_p1 = cur.vd_p1;
      double mf, ef;
      cell_face_flux(vf, pre_m1, pre_0, rho_m2, rho_m1, rho_0, rho_p1, e_m2, e_m1, e_0, e_p1, vd_face, vd_m1, vd_p1, mf, ef);
      if (STORE_MF && own_cell && F >= j0 && F < j1) AT(CLV_MASS_FLUX_Y, i, F) = mf;
      mf_h = mf;
      if (F > Fs) {
        // cell F - 1 from faces F - 1 and F, src/serial/advec_cell.cpp:211-216
        const double pre = pre_m1, rho = rho_m1, e = e_m1;
        double pre_mass_s = rho * pre;
        double post_mass_s = pre_mass_s + mf_prev - mf;
        double post_ener_s = (e * pre_mass_s + ef_prev - ef) / post_mass_s;
        double advec_vol_s = pre + vf_prev - vf;
        rho_h = post_mass_s / advec_vol_s;
        const int c = F - 1;
        if (own_cell && c >= j0 && c < j1 && c <= ny + 1) {
          rho_out[(int64_t)c * d.pitch + i] = rho_h;
          e_out[(int64_t)c * d.pitch + i] = post_ener_s;
          if (RESET) {
            AT(CLV_DENSITY0, i, c) = rho_h;
            AT(CLV_ENERGY0, i, c) = post_ener_s;
          }
        }
      }
      mf_prev = mf;
      ef_prev = ef;
      vf_prev = vf;
      rho_m2 = rho_m1;
      rho_m1 = rho_0;
      e_m2 = e_m1;
      e_m1 = e_0;
      rho_0 = rho_p1;  // row F + 1 (unclamped while F + 1 <= Fe <= ny + 2)
      e_0 = e_p1;
      pre_m1 = pre_0;
      vd_m1 = vd_face;
    }
    // history: mass flux row F and density row F - 1, ghost columns mirrored
    ring.put_mf(F, __shfl(mf_h, lane_f));
    ring.put_rho(F - 1, __shfl(rho_h, lane_c));
    ring.put_post(F, post_h);
    // mom part at node row k = F - 4
    const int k = F - 4;
    if (k == j0 - 2) {
      mom_row(k, nf_prev, mo_prev, mw_prev);
      u_m1 = AT(CLV_XVEL1, ic, k);
      v_m1 = AT(CLV_YVEL1, ic, k);
      u_0 = AT(CLV_XVEL1, ic, k + 1);
      v_0 = AT(CLV_YVEL1, ic, k + 1);
    } else if (k > j0 - 2) {
      double nf, mo, mw;
      mom_row(k, nf, mo, mw);
      const double nm_post = 0.25 * (mo_prev + mo + mw_prev + mw);  // :170-171
      const double nm_pre = nm_post - nf_prev + nf;                 // :172
      const double u_p1 = u_ahead, v_p1 = v_ahead;  // row k + 1
      if (k >= j0) {
        // momentum flux of face f = k - 1
        const int f = k - 1;
        const double wd = MU(CLV_CELLDY, f);
        double mu, mv;
        if (nf_prev < 0.0) {  // upwind f+2, donor f+1, downwind f, dif = donor (:186-191)
          const double wdif = MU(CLV_CELLDY, f + 1);
          mu = mom_flux_of(nf_prev, nm_pre, wd, wdif, u_p1, u_0, u_m1);
          mv = mom_flux_of(nf_prev, nm_pre, wd, wdif, v_p1, v_0, v_m1);
        } else {  // upwind f-1, donor f, downwind f+1, dif = upwind (:192-196)
          const double wdif = MU(CLV_CELLDY, f - 1);
          mu = mom_flux_of(nf_prev, pre_prev, wd, wdif, u_m2, u_m1, u_0);
          mv = mom_flux_of(nf_prev, pre_prev, wd, wdif, v_m2, v_m1, v_0);
        }
        if (k >= j0 + 1 && own_node) {
          // update node row f from faces f - 1 and f (:221)
          const double u_new = (u_m1 * pre_prev + mu_prev - mu) / post_prev, v_new = (v_m1 * pre_prev + mv_prev - mv) / post_prev;
          u_out[(int64_t)f * d.pitch + i] = u_new;
          v_out[(int64_t)f * d.pitch + i] = v_new;
          if (RESET) {
            AT(CLV_XVEL0, i, f) = u_new;
            AT(CLV_YVEL0, i, f) = v_new;
          }
        }
        mu_prev = mu;
        mv_prev = mv;
      }
      nf_prev = nf;
      mo_prev = mo;
      mw_prev = mw;
      pre_prev = nm_pre;
      post_prev = nm_post;
      u_m2 = u_m1;
      u_m1 = u_0;
      u_0 = u_p1;
      v_m2 = v_m1;
      v_m1 = v_0;
      v_0 = v_p1;
    }
    // The block's walks stay on the same rows (3-6 % faster at 15360^2 than
    // free-running waves; F, j1 are uniform over the block).  The batched form runs a
    // block with walks past its problem's last strip free, and those walks exit at
    // once: over many small problems that is 7 % faster for the whole batch than
    // keeping every block in step; keeping the full blocks in step gains 0.6 % more.
    if (together) __syncthreads();
    cur = nxt;
    u_ahead = u_n;
    v_ahead = v_n;
  }
}

// 4 waves per SIMD (the bound's second argument is the minimum waves per SIMD) with
// either block size: sweep 1 takes 128 VGPRs and spills 12-36 bytes a lane; bounded at
// 2, its 8-walk form took 130-136 VGPRs, 3 waves and one block a CU, 2 % slower
template <int SWEEP, bool MESH, bool RESET = false, int NW = WAVES, bool STORE_MF = true>
__global__ __launch_bounds__(64 * NW, 4) void cm_y_k(Dev d, int nstrips, int seg, double *rho_out, double *e_out, double *u_out, double *v_out) {
  int bx, by;
  xcd_tile(bx, by);
  const int j0 = 2 + seg * by;
  cm_y_tile<SWEEP, MESH, true, RESET, NW, STORE_MF>(d, bx, j0, min(j0 + seg, d.ny + 3), nstrips, rho_out, e_out, u_out, v_out);
}

// ---- batched form: persistent launches over the tiles of many chunks ----------------------
// A launch walks a list of BatchItems (problem, strip block, row segment).  Its
// blocks are dealt round-robin to the 8 XCDs (block b on XCD b % 8), so the
// blocks of XCD x take the x-th contiguous eighth of the list, keeping
// neighbouring tiles of one problem -- which re-read each other's edge rows and
// columns -- in one L2.  With fewer blocks than XCDs: a plain block stride.
struct ItemRange {
  int first, end, stride;
};
__device__ __forceinline__ ItemRange batch_items(int n) {
  const int G = gridDim.x, b = blockIdx.x;
  if (G < NUM_XCD) return {b, n, G};
  const int x = b % NUM_XCD, k = b / NUM_XCD;
  const int nbx = (G - x + NUM_XCD - 1) / NUM_XCD;  // blocks on XCD x
  const int lo = (int)((int64_t)n * x / NUM_XCD), hi = (int)((int64_t)n * (x + 1) / NUM_XCD);
  return {lo + k, hi, nbx};
}

// strips of each tile family
__device__ __forceinline__ int strips_of(int kind, int nx) {
  switch (kind) {
    case BK_TIMESTEP: return (nx + kOwnTimestep - 1) / kOwnTimestep;
    case BK_LAGRANGE: return (nx + kOwnLagrange - 1) / kOwnLagrange;
    case BK_CM_X1: case BK_CM_X2: return (nx + kOwnCmX) / kOwnCmX;  // nx + 1 face columns
    default: return (nx + kOwnCmY) / kOwnCmY;                     // cm y: nx + 1 node columns
  }
}

// A view loaded from memory holds generic pointers, which would make every
// access a flat one with 64-bit vector address arithmetic.  They all point to
// global memory: the batched kernels use this copy typed so (global loads and
// stores with scalar bases, as a view passed by value gets).
typedef __attribute__((address_space(1))) double gdouble;
struct GDev {
  gdouble *f[CLV_NUM_FIELDS];
  gdouble *m[CLV_NUM_MESH];
  int nx, ny;
  int64_t pitch;
  int geom_mesh;
};
__device__ __forceinline__ GDev global_view(const Dev &s) {
  GDev d;
  for (int k = 0; k < CLV_NUM_FIELDS; ++k) d.f[k] = (gdouble *)s.f[k];
  for (int k = 0; k < CLV_NUM_MESH; ++k) d.m[k] = (gdouble *)s.m[k];
  d.nx = s.nx;
  d.ny = s.ny;
  d.pitch = s.pitch;
  d.geom_mesh = s.geom_mesh;
  return d;
}

// One 8-byte load (a scalar load: the list is read-only for the launch).
__device__ __forceinline__ BatchItem load_item(const BatchItem *__restrict__ items, int w) {
  const uint64_t v = reinterpret_cast<const uint64_t *>(items)[w];
  return BatchItem{(uint16_t)v, (uint16_t)(v >> 16), (uint16_t)(v >> 32), (uint16_t)(v >> 48)};
}

template <bool MESH>
__global__ __launch_bounds__(256, 5) void batch_timestep_k(const Dev *__restrict__ tab, const BatchConst *__restrict__ pc,
                                                        const BatchState *__restrict__ ps, const BatchItem *__restrict__ items,
                                                        int nitems) {
  const ItemRange rg = batch_items(nitems);
  for (int w = rg.first; w < rg.end; w += rg.stride) {
    const BatchItem it = load_item(items, w);
    if (ps[it.p].complete) continue;
    const GDev d = global_view(tab[it.p]);
    const BatchConst c = pc[it.p];
    const int ns = strips_of(BK_TIMESTEP, d.nx);
    const int j0 = 2 + (int)it.k * it.seg;
    timestep_tile<MESH, false, false>(d, it.bx, j0, min(j0 + (int)it.seg, d.ny + 2), c.dtc_safe, c.dtu_safe, c.dtv_safe, c.dtdiv_safe,
                  c.partials + (size_t)it.k * ns, ns);
  }
}

template <bool MESH>
__global__ __launch_bounds__(256, 4) void batch_lagrange_k(const Dev *__restrict__ tab, const BatchState *__restrict__ ps,
                                                      const BatchItem *__restrict__ items, int nitems) {
  const ItemRange rg = batch_items(nitems);
  for (int w = rg.first; w < rg.end; w += rg.stride) {
    const BatchItem it = load_item(items, w);
    if (!ps[it.p].run) continue;
    const GDev d = global_view(tab[it.p]);
    const double dt = ps[it.p].dt;
    const int j0 = 2 + (int)it.k * it.seg;
    lagrange_tile<MESH, false, false, true>(d, it.bx, j0, min(j0 + (int)it.seg, d.ny + 2), dt, strips_of(BK_LAGRANGE, d.nx));
  }
}

// advection sweeps write the updated pair into work arrays 1/2 (cells) or 3/4
// (nodes) of the stage's view; the next stage's view has them renamed.
template <int KIND, bool MESH>
__global__ __launch_bounds__(256) void batch_advec_k(const Dev *__restrict__ tab, const BatchState *__restrict__ ps,
                                                   const BatchItem *__restrict__ items, int nitems) {
  const ItemRange rg = batch_items(nitems);
  for (int w = rg.first; w < rg.end; w += rg.stride) {
    const BatchItem it = load_item(items, w);
    if (!ps[it.p].run) continue;
    const GDev d = global_view(tab[it.p]);
    const int ns = strips_of(KIND, d.nx);
    const int j0 = 2 + (int)it.k * it.seg, jn1 = min(j0 + (int)it.seg, d.ny + 3);
    if constexpr (KIND == BK_CM_X1)
      cm_x_tile<1, MESH, false, false>(d, it.bx, j0, jn1, ns, d.f[CLV_WORK1], d.f[CLV_WORK2], d.f[CLV_WORK3], d.f[CLV_WORK4]);
    if constexpr (KIND == BK_CM_X2)
      cm_x_tile<2, MESH, false, false>(d, it.bx, j0, jn1, ns, d.f[CLV_WORK1], d.f[CLV_WORK2], d.f[CLV_WORK3], d.f[CLV_WORK4]);
    if constexpr (KIND == BK_CM_Y1)
      cm_y_tile<1, MESH, false, false, WAVES, false>(d, it.bx, j0, jn1, ns, d.f[CLV_WORK1], d.f[CLV_WORK2], d.f[CLV_WORK3], d.f[CLV_WORK4]);
    if constexpr (KIND == BK_CM_Y2)
      cm_y_tile<2, MESH, false, false, WAVES, false>(d, it.bx, j0, jn1, ns, d.f[CLV_WORK1], d.f[CLV_WORK2], d.f[CLV_WORK3], d.f[CLV_WORK4]);
  }
}

// Per problem: the step's dt from its timestep minima, with the host logic of
// clv_timestep / clv_hydro_step (driver/timestep.cpp:78-90, driver/hydro.cpp:108).
__global__ __launch_bounds__(256) void batch_dt_k(const BatchConst *pc, BatchState *ps) {
  const int p = blockIdx.x;
  BatchState s = ps[p];
  if (s.complete) {
    if (threadIdx.x == 0) ps[p].run = 0;
    return;
  }
  const BatchConst c = pc[p];
  double m = g_big;
  for (int b = threadIdx.x; b < c.nparts; b += 256) m = fmin(m, c.partials[b]);
  __shared__ double lds[4];
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double dtlp = fmin(fmin(lds[0], lds[1]), fmin(lds[2], lds[3]));
  double dt = g_big;
  if (dtlp <= dt) dt = dtlp;
  const double a = dt, b = s.dtold * c.dtrise;  // std::min(std::min(dt, dtold * dtrise), dtmax)
  const double mm = (b < a) ? b : a;
  dt = (c.dtmax < mm) ? c.dtmax : mm;
  s.step += 1;
  s.dt = dt;
  s.dtold = dt;
  if (dt < c.dtmin) {  // the host path stops with CLV_ERR_SMALL_DT before PdV
    s.failed_step = s.step;
    s.complete = 1;
    s.run = 0;
  } else {
    s.run = 1;
    s.ran += 1;
    if (s.step <= c.hist_len) c.dt_hist[s.step - 1] = dt;
    s.time += dt;
    s.complete = (s.time + g_small > c.end_time || s.step >= c.end_step) ? 1 : 0;
  }
  ps[p] = s;
}

}  // namespace

template <bool MESH>
static hipError_t launch_tiles(int kind, hipStream_t s, int grid, const Dev *tab, const BatchConst *pc, const BatchState *ps,
                               const BatchItem *items, int nitems) {
  const dim3 g(grid, 1, 1), blk(256, 1, 1);
  switch (kind) {
    case BK_TIMESTEP: hipLaunchKernelGGL(batch_timestep_k<MESH>, g, blk, 0, s, tab, pc, ps, items, nitems); break;
    case BK_LAGRANGE: hipLaunchKernelGGL(batch_lagrange_k<MESH>, g, blk, 0, s, tab, ps, items, nitems); break;
    case BK_CM_X1: hipLaunchKernelGGL((batch_advec_k<BK_CM_X1, MESH>), g, blk, 0, s, tab, ps, items, nitems); break;
    case BK_CM_X2: hipLaunchKernelGGL((batch_advec_k<BK_CM_X2, MESH>), g, blk, 0, s, tab, ps, items, nitems); break;
    case BK_CM_Y1: hipLaunchKernelGGL((batch_advec_k<BK_CM_Y1, MESH>), g, blk, 0, s, tab, ps, items, nitems); break;
    case BK_CM_Y2: hipLaunchKernelGGL((batch_advec_k<BK_CM_Y2, MESH>), g, blk, 0, s, tab, ps, items, nitems); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t batch_launch_tiles(int kind, hipStream_t s, int grid, const Dev *tab, const BatchConst *pc, const BatchState *ps,
                              const BatchItem *items, int nitems, int mesh) {
  return mesh ? launch_tiles<true>(kind, s, grid, tab, pc, ps, items, nitems) : launch_tiles<false>(kind, s, grid, tab, pc, ps, items, nitems);
}

hipError_t batch_launch_dt(hipStream_t s, int nprob, const BatchConst *pc, BatchState *ps) {
  hipLaunchKernelGGL(batch_dt_k, dim3(nprob, 1, 1), dim3(256, 1, 1), 0, s, pc, ps);
  return hipGetLastError();
}

// ---- host side ----------------------------------------------------------------------------
#define LAUNCH_FT(name, kern, grid, threads, ...)                                    \
  do {                                                                               \
    before_launch(c, name);                                                          \
    hipLaunchKernelGGL(kern, grid, dim3(threads, 1, 1), 0, c->stream, __VA_ARGS__);  \
    clv_status st_ = after_launch(c, name);                                          \
    if (st_ != CLV_OK) return st_;                                                   \
  } while (0)
#define LAUNCH_F(name, kern, grid, ...) LAUNCH_FT(name, kern, grid, 256, __VA_ARGS__)

static int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Rows per wave segment of each plain launch, per kernel family (timestep,
// lagrange, x advection pass, y advection pass).  Longer segments re-read fewer
// window rows at segment edges; shorter ones give more, shorter workgroups (a
// smaller tail, a tighter working set per XCD L2).  At 15360^2 (ms per launch over
// segments of 8 ... 256 rows) the row marches are fastest at 8-16 rows and the
// per-lane y walk at 128 (64 and 256 within 1 %).  On a smaller grid a family's
// segment is halved until it has 4 groups of 4 waves per CU (or 8 rows).  The timestep and
// lagrange marches halve until 16 per CU (3840^2: 8 rows, not 16, a 1.5 % faster
// step; 15360^2 keeps 16, 2-3 % faster there than 8); the y walk, which spends 6
// extra steps at each segment's start, stops at 3.5 (1920^2: 16 rows, not 8, 10 %
// faster; 3840^2: 64, not 32, 5 % faster).
enum { SF_TIMESTEP, SF_LAGRANGE, SF_CM_X, SF_CM_Y, SF_COUNT };
constexpr int kSegLarge[SF_COUNT] = {16, 16, 8, 128};
constexpr int kBlocksPer2Cu[SF_COUNT] = {32, 32, 8, 7};
constexpr int kSegMin = 8;

static void choose_segments(clv_chunk *c) {
  if (c->seg[0]) return;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1) cus = 256;
  const int nx = c->d.nx, ny = c->d.ny;
  const int strips[SF_COUNT] = {ceil_div(nx, kOwnTimestep), ceil_div(nx, kOwnLagrange), ceil_div(nx + 1, kOwnCmX), ceil_div(nx + 1, kOwnCmY)};
  const int rows[SF_COUNT] = {ny, ny, ny + 1, ny + 1};
  for (int f = 0; f < SF_COUNT; ++f) {
    int seg = kSegLarge[f];
    while (seg > kSegMin && 2 * ceil_div(strips[f], WAVES) * ceil_div(rows[f], seg) < kBlocksPer2Cu[f] * cus) seg /= 2;
    c->seg[f] = seg;
  }
}

// strips of the timestep kernel (kOwnTimestep owned columns each) times row segments (of
// at least kSegMin rows), and the second level of the minimum
constexpr int kMinLevel2 = 256;
size_t fused_partials_needed(int nx, int ny) { return (size_t)ceil_div(nx, kOwnTimestep) * (size_t)ceil_div(ny, kSegMin) + kMinLevel2; }

static clv_status halo(clv_chunk *c, std::initializer_list<int> ids, int depth) {
  int fields[CLV_NUM_HALO_FIELDS] = {0};
  for (int id : ids) fields[id] = 1;
  return k_update_halo(c, fields, depth);
}

static void swap_fields(clv_chunk *c, int a, int b) {
  double *t = c->d.f[a];
  c->d.f[a] = c->d.f[b];
  c->d.f[b] = t;
}

static clv_status require_owning(clv_chunk *c) {
  // the fused plan renames buffers (double buffering, reset_field by swap), which
  // only a chunk that owns its arena may do
  if (!c->owns) return set_error(CLV_ERR_ARG, "CLV_PLAN_FUSED needs a chunk created by clv_chunk_create (not clv_chunk_wrap)");
  if (c->d.nx < 4 || c->d.ny < 4) return set_error(CLV_ERR_ARG, "CLV_PLAN_FUSED needs x_cells, y_cells >= 4");
  choose_segments(c);
  return CLV_OK;
}

// timestep_k over the chunk and the two-level minimum of its partials into
// c->result_dev, then copied to c->result_host (asynchronously).  write_pc: store p and
// c too; store_q: store q (the fused step does not: its lagrange pass recomputes q).
static clv_status timestep_launch(clv_chunk *c, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe, bool write_pc,
                                  bool store_q) {
  const Dev &d = c->d;
  const int seg = c->seg[SF_TIMESTEP];
  const int ns = ceil_div(d.nx, kOwnTimestep), nseg = ceil_div(d.ny, seg);
  const int n1 = ns * nseg;
  if ((size_t)n1 + kMinLevel2 > c->partials_len) return set_error(CLV_ERR_ARG, "fused timestep: partials buffer too small");
  const dim3 g(ceil_div(ns, WAVES), nseg, 1);
  if (write_pc) {
    const auto kern = d.geom_mesh ? timestep_k<true, true> : timestep_k<false, true>;
    LAUNCH_F("f_timestep_pc", kern, g, d, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, c->partials, ns, seg);
  } else if (store_q) {
    const auto kern = d.geom_mesh ? timestep_k<true> : timestep_k<false>;
    LAUNCH_F("f_timestep_q", kern, g, d, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, c->partials, ns, seg);
  } else {
    const auto kern = d.geom_mesh ? timestep_k<true, false, false> : timestep_k<false, false, false>;
    LAUNCH_F("f_timestep", kern, g, d, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, c->partials, ns, seg);
  }
  before_launch(c, "calc_dt_final");
  if (n1 <= 16384) {
    hipLaunchKernelGGL(reduce_min_big_k, dim3(1, 1, 1), dim3(1024, 1, 1), 0, c->stream, (const double *)c->partials, n1, c->result_dev);
  } else {  // two levels: kMinLevel2 block minima, then one
    double *level2 = c->partials + n1;
    hipLaunchKernelGGL(reduce_min_big_k, dim3(kMinLevel2, 1, 1), dim3(1024, 1, 1), 0, c->stream, (const double *)c->partials, n1, level2);
    hipLaunchKernelGGL(reduce_min_big_k, dim3(1, 1, 1), dim3(1024, 1, 1), 0, c->stream, (const double *)level2, kMinLevel2, c->result_dev);
  }
  clv_status st = after_launch(c, "calc_dt_final");
  if (st != CLV_OK) return st;
  return hip_check(hipMemcpyAsync(c->result_host, c->result_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream), "fused dt copy");
}

clv_status fused_timestep_kernels(clv_chunk *c, const clv_config *cfg, double *dt_min_val) {
  clv_status st = require_owning(c);
  if (st != CLV_OK) return st;
  // q is not stored: the lagrange pass recomputes it with its ghosts (driver/timestep.cpp:62-69)
  if ((st = timestep_launch(c, cfg->dtc_safe, cfg->dtu_safe, cfg->dtv_safe, cfg->dtdiv_safe, false, false)) != CLV_OK) return st;
  if ((st = hip_check(hipStreamSynchronize(c->stream), "fused dt sync")) != CLV_OK) return st;
  *dt_min_val = c->result_host[0];
  return CLV_OK;
}

// driver/timestep.cpp:47-69 for the per-entry-point path, the volume and face areas
// and density0, energy0, xvel0, yvel0 halos already in place: pressure, soundspeed
// and viscosity written as ideal_gas and viscosity write them, then the pressure
// halo of depth p_depth and viscosity's of depth 1, then the dt minimum.  p_depth 0:
// pressure and soundspeed are not stored (nor the pressure halo run), for a caller
// whose next pass rewrites them.  Any chunk (a wrapped one too): nothing is renamed.
clv_status fused_timestep_entry(clv_chunk *c, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe, int p_depth,
                                double *dt_min_val) {
  if (c->d.nx < 4 || c->d.ny < 4) return set_error(CLV_ERR_ARG, "clv_timestep_kernels needs x_cells, y_cells >= 4");
  choose_segments(c);
  clv_status st = timestep_launch(c, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, p_depth > 0, true);
  if (st != CLV_OK) return st;
  if (p_depth > 0 && (st = halo(c, {CLV_PRESSURE}, p_depth)) != CLV_OK) return st;
  if ((st = halo(c, {CLV_VISCOSITY}, 1)) != CLV_OK) return st;
  if ((st = hip_check(hipStreamSynchronize(c->stream), "fused dt sync")) != CLV_OK) return st;
  *dt_min_val = c->result_host[0];
  return CLV_OK;
}

// src/serial/PdV.cpp:92-133 (predict), accelerate, PdV (correct) and flux_calc for the
// per-entry-point path, the timestep's halos in place: every array as those entry
// points leave it -- the predictor's pressure (with its depth-1 reflective halo) and
// sound speed included -- in one pass.  Any chunk (a wrapped one too): nothing is
// renamed.
clv_status fused_lagrange_entry(clv_chunk *c, double dt, double *rho_pred, double *e_pred) {
  if (c->d.nx < 4 || c->d.ny < 4) return set_error(CLV_ERR_ARG, "clv_lagrange_kernels needs x_cells, y_cells >= 4");
  if ((rho_pred == nullptr) != (e_pred == nullptr)) return set_error(CLV_ERR_ARG, "fused lagrange: both predictor outputs or neither");
  choose_segments(c);
  const Dev &d = c->d;
  const int seg = c->seg[SF_LAGRANGE];
  const int ns = ceil_div(d.nx, kOwnLagrange), nseg = ceil_div(d.ny, seg);
  const dim3 g(ceil_div(ns, WAVES), nseg, 1);
  if (rho_pred) {
    const auto kern = d.geom_mesh ? lagrange_k<true, true, true> : lagrange_k<false, true, true>;
    LAUNCH_F("f_lagrange_pred", kern, g, d, dt, ns, seg, rho_pred, e_pred);
  } else {
    const auto kern = d.geom_mesh ? lagrange_k<true, true> : lagrange_k<false, true>;
    LAUNCH_F("f_lagrange_pc", kern, g, d, dt, ns, seg, rho_pred, e_pred);
  }
  clv_status st = k_flux_calc_edges(c, dt);
  if (st != CLV_OK) return st;
  return halo(c, {CLV_PRESSURE}, 1);
}

clv_status fused_lagrange(clv_chunk *c, double dt) {
  clv_status st = require_owning(c);
  if (st != CLV_OK) return st;
  const Dev &d = c->d;
  const int seg = c->seg[SF_LAGRANGE];
  const int ns = ceil_div(d.nx, kOwnLagrange), nseg = ceil_div(d.ny, seg);
  const auto kern = d.geom_mesh ? lagrange_k<true, false, false, true> : lagrange_k<false, false, false, true>;
  LAUNCH_F("f_lagrange", kern, dim3(ceil_div(ns, WAVES), nseg, 1), d, dt, ns, seg, nullptr, nullptr);
  return CLV_OK;
}

// advec_cell and advec_mom of one sweep in one launch (cm_x_k, cm_y_k); the halo
// update the driver runs between the two (driver/advection.cpp:42-46) is split: the
// velocities' before (the cell sweep does not change them), the rest after (the mom
// sweep does not change those), and the kernels read the ghosts they need by mirror.
// reset: the pass also does reset_field's stores (the kernels' RESET form).
using CmKernel = void (*)(Dev, int, int, double *, double *, double *, double *);
template <bool MESH, bool RESET, bool STORE_MF>
static CmKernel cm_kernel(int sweep, int dir, bool wide) {
  if (dir == CLV_XDIR) return sweep == 1 ? cm_x_k<1, MESH, RESET, STORE_MF> : cm_x_k<2, MESH, RESET, STORE_MF>;
  if (wide) return sweep == 1 ? cm_y_k<1, MESH, RESET, kWavesCmY, STORE_MF> : cm_y_k<2, MESH, RESET, kWavesCmY, STORE_MF>;
  return sweep == 1 ? cm_y_k<1, MESH, RESET, WAVES, STORE_MF> : cm_y_k<2, MESH, RESET, WAVES, STORE_MF>;
}
template <bool MESH>
static CmKernel cm_kernel(int sweep, int dir, bool wide, bool reset, bool store_mf) {
  if (reset) return cm_kernel<MESH, true, true>(sweep, dir, wide);
  return store_mf ? cm_kernel<MESH, false, true>(sweep, dir, wide) : cm_kernel<MESH, false, false>(sweep, dir, wide);
}

// store_mf false: the fused step's form, the mass fluxes not stored (reset implies them stored)
static clv_status cm_sweep(clv_chunk *c, int sweep, int dir, bool reset, bool store_mf) {
  const Dev &d = c->d;
  double *rho_out = d.f[CLV_WORK1], *e_out = d.f[CLV_WORK2], *u_out = d.f[CLV_WORK3], *v_out = d.f[CLV_WORK4];
  const int own = dir == CLV_XDIR ? kOwnCmX : kOwnCmY, seg = c->seg[dir == CLV_XDIR ? SF_CM_X : SF_CM_Y];
  const int ns = ceil_div(d.nx + 1, own);
  const bool wide = dir == CLV_YDIR && ns >= kWideCmY;
  const int waves = wide ? kWavesCmY : WAVES;
  const CmKernel kern = d.geom_mesh ? cm_kernel<true>(sweep, dir, wide, reset, store_mf) : cm_kernel<false>(sweep, dir, wide, reset, store_mf);
  static const char *const names[2][2][3] = {{{"f_cm_x1", "f_cm_x1_mf", "f_cm_x1_reset"}, {"f_cm_x2", "f_cm_x2_mf", "f_cm_x2_reset"}},
                                             {{"f_cm_y1", "f_cm_y1_mf", "f_cm_y1_reset"}, {"f_cm_y2", "f_cm_y2_mf", "f_cm_y2_reset"}}};
  const char *name = names[dir == CLV_XDIR ? 0 : 1][sweep == 1 ? 0 : 1][reset ? 2 : (store_mf ? 1 : 0)];
  LAUNCH_FT(name, kern, dim3(ceil_div(ns, waves), ceil_div(d.ny + 1, seg), 1), 64 * waves, d, ns, seg, rho_out, e_out,
            u_out, v_out);
  swap_fields(c, CLV_DENSITY1, CLV_WORK1);
  swap_fields(c, CLV_ENERGY1, CLV_WORK2);
  swap_fields(c, CLV_XVEL1, CLV_WORK3);
  swap_fields(c, CLV_YVEL1, CLV_WORK4);
  return CLV_OK;
}

// One sweep of driver/advection.cpp (:47-73 or :75-103) for the per-entry-point path,
// the velocities' depth-2 halo (and for sweep 1 energy1, density1 and the volume
// fluxes') in place: advec_cell, the halo of density1, energy1 and both mass fluxes,
// advec_mom for xvel1 and yvel1, in one pass.  The new fields are exchanged with work
// arrays (the caller re-points its handles); the velocities' ghost frames are carried
// over first, so that they hold the halo the reference's advec_mom leaves untouched.
// reset: reset_field (src/serial/reset_field.cpp) follows, its stores done by the pass.
clv_status fused_advection_entry(clv_chunk *c, int sweep, int dir, bool reset) {
  if (c->d.nx < 4 || c->d.ny < 4) return set_error(CLV_ERR_ARG, "clv_advection_pass needs x_cells, y_cells >= 4");
  choose_segments(c);
  clv_status st = k_copy_frame(c, CLV_XVEL1, CLV_WORK3);
  if (st == CLV_OK) st = k_copy_frame(c, CLV_YVEL1, CLV_WORK4);
  if (st == CLV_OK) st = cm_sweep(c, sweep, dir, reset, true);
  if (st != CLV_OK) return st;
  return halo(c, {CLV_DENSITY1, CLV_ENERGY1, CLV_MASS_FLUX_X, CLV_MASS_FLUX_Y}, 2);
}

// driver/advection.cpp:29-105 followed by reset_field (src/serial/reset_field.cpp)
clv_status fused_advection(clv_chunk *c, int advect_x) {
  clv_status st = require_owning(c);
  if (st != CLV_OK) return st;
  const int dir1 = advect_x ? CLV_XDIR : CLV_YDIR, dir2 = advect_x ? CLV_YDIR : CLV_XDIR;
  if ((st = halo(c, {CLV_ENERGY1, CLV_DENSITY1, CLV_VOL_FLUX_X, CLV_VOL_FLUX_Y, CLV_XVEL1, CLV_YVEL1}, 2)) != CLV_OK) return st;
  // the passes leave the mass fluxes unstored: nothing reads them before the next
  // step's passes rewrite them (driver/advection.cpp:47-103)
  if ((st = cm_sweep(c, 1, dir1, false, false)) != CLV_OK) return st;
  // the first pass's halo after it and the second pass's velocity halo before it
  if ((st = halo(c, {CLV_DENSITY1, CLV_ENERGY1, CLV_XVEL1, CLV_YVEL1}, 2)) != CLV_OK) return st;
  if ((st = cm_sweep(c, 2, dir2, false, false)) != CLV_OK) return st;
  if ((st = halo(c, {CLV_DENSITY1, CLV_ENERGY1}, 2)) != CLV_OK) return st;
  // reset_field: the end-of-step fields become the start-of-step fields
  swap_fields(c, CLV_DENSITY0, CLV_DENSITY1);
  swap_fields(c, CLV_ENERGY0, CLV_ENERGY1);
  swap_fields(c, CLV_XVEL0, CLV_XVEL1);
  swap_fields(c, CLV_YVEL0, CLV_YVEL1);
  return CLV_OK;
}

}  // namespace clv
