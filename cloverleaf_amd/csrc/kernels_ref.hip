// gfx950 kernels for the per-step hydro loops, one launch per reference loop.
//
// Each kernel keeps the reference loop's iteration range and floating-point
// expression order (built with -ffp-contract=off, no fast-math), so the fields
// it writes are bit-identical to the serial model's.  Launch shape: 2D blocks
// of 64 x 4 threads, one wave64 per row segment so that every wave reads and
// writes whole 128-byte lines of x-contiguous rows; one thread per cell/node.
#include "clv_internal.h"
#include "geom.h"

namespace clv {
namespace {

#define AT(fid, i, j) d.f[fid][(int64_t)(j) * d.pitch + (i)]
#define M(mid, i) d.m[mid][i]
// every output element is stored once and read next by another launch, from HBM:
// streaming (non-temporal) stores, 2.7 % off the drop-in's step at 15360^2
#define ST_STREAM(lhs, v) __builtin_nontemporal_store((v), &(lhs))

constexpr int BX = 64, BY = 4;

inline dim3 grid2(int x0, int x1, int y0, int y1) {
  return dim3((unsigned)((x1 - x0 + BX - 1) / BX), (unsigned)((y1 - y0 + BY - 1) / BY), 1);
}

#define RANGE_GUARD(x0, x1, y0, y1)                   \
  const int i = (x0) + blockIdx.x * BX + threadIdx.x; \
  const int j = (y0) + blockIdx.y * BY + threadIdx.y; \
  if (i >= (x1) || j >= (y1)) return;

// ---- ideal_gas: src/serial/ideal_gas.cpp:38-47 ------------------------------------------
__global__ __launch_bounds__(256) void ideal_gas_k(Dev d, int rho_id, int e_id) {
  RANGE_GUARD(2, d.nx + 2, 2, d.ny + 2)
  const double rho = AT(rho_id, i, j);
  double v = 1.0 / rho;
  double p = (1.4 - 1.0) * rho * AT(e_id, i, j);
  double pressurebyenergy = (1.4 - 1.0) * rho;
  double pressurebyvolume = -rho * p;
  double sound_speed_squared = v * v * (p * pressurebyenergy - pressurebyvolume);
  ST_STREAM(AT(CLV_PRESSURE, i, j), p);
  ST_STREAM(AT(CLV_SOUNDSPEED, i, j), sqrt(sound_speed_squared));
}

// ---- viscosity: src/serial/viscosity.cpp:37-67 (diry tests pgradx, :57) ------------------
__global__ __launch_bounds__(256) void viscosity_k(Dev d) {
  RANGE_GUARD(2, d.nx + 2, 2, d.ny + 2)
  const double u00 = AT(CLV_XVEL0, i, j), u10 = AT(CLV_XVEL0, i + 1, j), u01 = AT(CLV_XVEL0, i, j + 1), u11 = AT(CLV_XVEL0, i + 1, j + 1);
  const double v00 = AT(CLV_YVEL0, i, j), v10 = AT(CLV_YVEL0, i + 1, j), v01 = AT(CLV_YVEL0, i, j + 1), v11 = AT(CLV_YVEL0, i + 1, j + 1);
  const double cdx = M(CLV_CELLDX, i), cdy = M(CLV_CELLDY, j);
  double ugrad = (u10 + u11) - (u00 + u01);
  double vgrad = (v01 + v11) - (v00 + v10);
  double div = (cdx * (ugrad) + cdy * (vgrad));
  // q = 0 whenever div >= 0, whatever the limiter (src/serial/viscosity.cpp:50): the
  // limiter, its seven divisions and the pressure and density loads are needed only
  // where !(div >= 0), and there every expression is the reference's
  double q = 0.0;
  if (!(div >= 0.0)) {
    double strain2 = 0.5 * (u01 + u11 - u00 - u10) / cdy + 0.5 * (v10 + v11 - v00 - v01) / cdx;
    double pgradx = (AT(CLV_PRESSURE, i + 1, j) - AT(CLV_PRESSURE, i - 1, j)) / (cdx + M(CLV_CELLDX, i + 1));
    double pgrady = (AT(CLV_PRESSURE, i, j + 1) - AT(CLV_PRESSURE, i, j - 1)) / (cdy + M(CLV_CELLDY, j + 2));
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
      q = 2.0 * AT(CLV_DENSITY0, i, j) * grad2 * limiter * limiter;
    }
  }
  ST_STREAM(AT(CLV_VISCOSITY, i, j), q);
}

// ---- PdV: src/serial/PdV.cpp:46-86 -------------------------------------------------------
template <bool PREDICT>
__global__ __launch_bounds__(256) void pdv_k(Dev d, double dt) {
  RANGE_GUARD(2, d.nx + 2, 2, d.ny + 2)
  const int U1 = PREDICT ? CLV_XVEL0 : CLV_XVEL1;
  const int V1 = PREDICT ? CLV_YVEL0 : CLV_YVEL1;
  double left_flux, right_flux, bottom_flux, top_flux;
  if (PREDICT) {
    left_flux = (geom_xarea(d, i, j) * (AT(CLV_XVEL0, i, j) + AT(CLV_XVEL0, i, j + 1) + AT(U1, i, j) + AT(U1, i, j + 1))) * 0.25 * dt * 0.5;
    right_flux = (geom_xarea(d, i + 1, j) * (AT(CLV_XVEL0, i + 1, j) + AT(CLV_XVEL0, i + 1, j + 1) + AT(U1, i + 1, j) + AT(U1, i + 1, j + 1))) *
                 0.25 * dt * 0.5;
    bottom_flux = (geom_yarea(d, i, j) * (AT(CLV_YVEL0, i, j) + AT(CLV_YVEL0, i + 1, j) + AT(V1, i, j) + AT(V1, i + 1, j))) * 0.25 * dt * 0.5;
    top_flux = (geom_yarea(d, i, j + 1) * (AT(CLV_YVEL0, i, j + 1) + AT(CLV_YVEL0, i + 1, j + 1) + AT(V1, i, j + 1) + AT(V1, i + 1, j + 1))) *
               0.25 * dt * 0.5;
  } else {
    left_flux = (geom_xarea(d, i, j) * (AT(CLV_XVEL0, i, j) + AT(CLV_XVEL0, i, j + 1) + AT(U1, i, j) + AT(U1, i, j + 1))) * 0.25 * dt;
    right_flux = (geom_xarea(d, i + 1, j) * (AT(CLV_XVEL0, i + 1, j) + AT(CLV_XVEL0, i + 1, j + 1) + AT(U1, i + 1, j) + AT(U1, i + 1, j + 1))) *
                 0.25 * dt;
    bottom_flux = (geom_yarea(d, i, j) * (AT(CLV_YVEL0, i, j) + AT(CLV_YVEL0, i + 1, j) + AT(V1, i, j) + AT(V1, i + 1, j))) * 0.25 * dt;
    top_flux = (geom_yarea(d, i, j + 1) * (AT(CLV_YVEL0, i, j + 1) + AT(CLV_YVEL0, i + 1, j + 1) + AT(V1, i, j + 1) + AT(V1, i + 1, j + 1))) *
               0.25 * dt;
  }
  double total_flux = right_flux - left_flux + top_flux - bottom_flux;
  const double vol = geom_volume(d, i, j);
  double volume_change_s = vol / (vol + total_flux);
  double recip_volume = 1.0 / vol;
  const double rho0 = AT(CLV_DENSITY0, i, j);
  double energy_change = (AT(CLV_PRESSURE, i, j) / rho0 + AT(CLV_VISCOSITY, i, j) / rho0) * total_flux * recip_volume;
  ST_STREAM(AT(CLV_ENERGY1, i, j), AT(CLV_ENERGY0, i, j) - energy_change);
  ST_STREAM(AT(CLV_DENSITY1, i, j), rho0 * volume_change_s);
}

// ---- revert: src/serial/revert.cpp:35-40 -------------------------------------------------
__global__ __launch_bounds__(256) void revert_k(Dev d) {
  RANGE_GUARD(2, d.nx + 2, 2, d.ny + 2)
  ST_STREAM(AT(CLV_DENSITY1, i, j), AT(CLV_DENSITY0, i, j));
  ST_STREAM(AT(CLV_ENERGY1, i, j), AT(CLV_ENERGY0, i, j));
}

// ---- accelerate: src/serial/accelerate.cpp:41-55 -----------------------------------------
__global__ __launch_bounds__(256) void accelerate_k(Dev d, double halfdt) {
  RANGE_GUARD(2, d.nx + 3, 2, d.ny + 3)
  double stepbymass_s = halfdt / ((AT(CLV_DENSITY0, i - 1, j - 1) * geom_volume(d, i - 1, j - 1) + AT(CLV_DENSITY0, i - 1, j) * geom_volume(d, i - 1, j) +
                                   AT(CLV_DENSITY0, i, j) * geom_volume(d, i, j) + AT(CLV_DENSITY0, i, j - 1) * geom_volume(d, i, j - 1)) *
                                  0.25);
  const double xa0 = geom_xarea(d, i, j), xa1 = geom_xarea(d, i, j - 1);
  const double ya0 = geom_yarea(d, i, j), ya1 = geom_yarea(d, i - 1, j);
  const double p00 = AT(CLV_PRESSURE, i, j), pm0 = AT(CLV_PRESSURE, i - 1, j), p0m = AT(CLV_PRESSURE, i, j - 1), pmm = AT(CLV_PRESSURE, i - 1, j - 1);
  const double q00 = AT(CLV_VISCOSITY, i, j), qm0 = AT(CLV_VISCOSITY, i - 1, j), q0m = AT(CLV_VISCOSITY, i, j - 1), qmm = AT(CLV_VISCOSITY, i - 1, j - 1);
  double u = AT(CLV_XVEL0, i, j) - stepbymass_s * (xa0 * (p00 - pm0) + xa1 * (p0m - pmm));
  double v = AT(CLV_YVEL0, i, j) - stepbymass_s * (ya0 * (p00 - p0m) + ya1 * (pm0 - pmm));
  u = u - stepbymass_s * (xa0 * (q00 - qm0) + xa1 * (q0m - qmm));
  v = v - stepbymass_s * (ya0 * (q00 - q0m) + ya1 * (qm0 - qmm));
  ST_STREAM(AT(CLV_XVEL1, i, j), u);
  ST_STREAM(AT(CLV_YVEL1, i, j), v);
}

// ---- flux_calc: src/serial/flux_calc.cpp:37-42 -------------------------------------------
__global__ __launch_bounds__(256) void flux_calc_k(Dev d, double dt) {
  RANGE_GUARD(2, d.nx + 3, 2, d.ny + 3)
  AT(CLV_VOL_FLUX_X, i, j) =
      0.25 * dt * geom_xarea(d, i, j) * (AT(CLV_XVEL0, i, j) + AT(CLV_XVEL0, i, j + 1) + AT(CLV_XVEL1, i, j) + AT(CLV_XVEL1, i, j + 1));
  AT(CLV_VOL_FLUX_Y, i, j) =
      0.25 * dt * geom_yarea(d, i, j) * (AT(CLV_YVEL0, i, j) + AT(CLV_YVEL0, i + 1, j) + AT(CLV_YVEL1, i, j) + AT(CLV_YVEL1, i + 1, j));
}

// The one extra flux of each direction that flux_calc's loops compute beyond what the
// advection reads (src/serial/flux_calc.cpp:33-35): vol_flux_x of the ghost row ny + 2
// and vol_flux_y of the ghost column nx + 2, from the ghost velocities as they stand.
// The fused Lagrangian pass leaves them out; the per-entry-point form adds them here.
__global__ __launch_bounds__(256) void flux_calc_edges_k(Dev d, double dt) {
  const int t = blockIdx.x * (BX * BY) + threadIdx.y * BX + threadIdx.x;  // BX x BY blocks (LAUNCH)
  if (t < d.nx + 1) {
    const int i = 2 + t, j = d.ny + 2;
    ST_STREAM(AT(CLV_VOL_FLUX_X, i, j),
              0.25 * dt * geom_xarea(d, i, j) * (AT(CLV_XVEL0, i, j) + AT(CLV_XVEL0, i, j + 1) + AT(CLV_XVEL1, i, j) + AT(CLV_XVEL1, i, j + 1)));
  }
  if (t < d.ny + 1) {
    const int i = d.nx + 2, j = 2 + t;
    ST_STREAM(AT(CLV_VOL_FLUX_Y, i, j),
              0.25 * dt * geom_yarea(d, i, j) * (AT(CLV_YVEL0, i, j) + AT(CLV_YVEL0, i + 1, j) + AT(CLV_YVEL1, i, j) + AT(CLV_YVEL1, i + 1, j)));
  }
}

// ---- reset_field: src/serial/reset_field.cpp:35-50 (both loops in one launch) ------------
__global__ __launch_bounds__(256) void reset_field_k(Dev d) {
  RANGE_GUARD(2, d.nx + 3, 2, d.ny + 3)
  if (i < d.nx + 2 && j < d.ny + 2) {
    ST_STREAM(AT(CLV_DENSITY0, i, j), AT(CLV_DENSITY1, i, j));
    ST_STREAM(AT(CLV_ENERGY0, i, j), AT(CLV_ENERGY1, i, j));
  }
  ST_STREAM(AT(CLV_XVEL0, i, j), AT(CLV_XVEL1, i, j));
  ST_STREAM(AT(CLV_YVEL0, i, j), AT(CLV_YVEL1, i, j));
}

// ---- advec_cell: src/serial/advec_cell.cpp:37-219 ----------------------------------------
// pre_vol = work1, post_vol = work2, ener_flux = work7 (src/serial/advec_cell.cpp:228-231)
template <int DIR, int SWEEP>
__global__ __launch_bounds__(256) void advec_cell_vol_k(Dev d) {
  RANGE_GUARD(0, d.nx + 4, 0, d.ny + 4)
  const double vol = geom_volume(d, i, j);
  double pre, post;
  if (DIR == 1) {
    const double fx = AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
    if (SWEEP == 1) {
      pre = vol + (fx + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j));
      post = pre - fx;
    } else {
      pre = vol + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
      post = vol;
    }
  } else {
    const double fy = AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
    if (SWEEP == 1) {
      pre = vol + (fy + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j));
      post = pre - fy;
    } else {
      pre = vol + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
      post = vol;
    }
  }
  ST_STREAM(AT(CLV_WORK1, i, j), pre);
  ST_STREAM(AT(CLV_WORK2, i, j), post);
}

// One van Leer limited donor value (src/serial/advec_cell.cpp:86-96 and :99-109).
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

template <int DIR>
__global__ __launch_bounds__(256) void advec_cell_flux_k(Dev d) {
  // x: j in [2, ny+2), i in [2, nx+4); y: j in [2, ny+4), i in [2, nx+2)
  RANGE_GUARD(2, DIR == 1 ? d.nx + 4 : d.nx + 2, 2, DIR == 1 ? d.ny + 2 : d.ny + 4)
  const int vf_id = DIR == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y;
  const int mf_id = DIR == 1 ? CLV_MASS_FLUX_X : CLV_MASS_FLUX_Y;
  const double vf = AT(vf_id, i, j);
  const int s = DIR == 1 ? i : j;
  const int smax = DIR == 1 ? d.nx + 2 : d.ny + 2;
  int upwind, donor, downwind, dif;
  if (vf > 0.0) {
    upwind = s - 2;
    donor = s - 1;
    downwind = s;
    dif = donor;
  } else {
    upwind = min(s + 1, smax);
    donor = s;
    downwind = s - 1;
    dif = upwind;
  }
#define AL(fid, t) (DIR == 1 ? AT(fid, t, j) : AT(fid, i, t))
  const double vd = DIR == 1 ? M(CLV_VERTEXDX, s) / M(CLV_VERTEXDX, dif) : M(CLV_VERTEXDY, s) / M(CLV_VERTEXDY, dif);
  const double pre_d = AL(CLV_WORK1, donor);
  double sigmat = fabs(vf) / pre_d;
  double sigma3 = (1.0 + sigmat) * vd;
  double sigma4 = 2.0 - sigmat;
  double sigmav = sigmat;
  const double rho_d = AL(CLV_DENSITY1, donor);
  double limiter = vl_limiter(sigmav, sigma3, sigma4, AL(CLV_DENSITY1, upwind), rho_d, AL(CLV_DENSITY1, downwind));
  double mf = vf * (rho_d + limiter);
  double sigmam = fabs(mf) / (rho_d * pre_d);
  const double e_d = AL(CLV_ENERGY1, donor);
  limiter = vl_limiter(sigmam, sigma3, sigma4, AL(CLV_ENERGY1, upwind), e_d, AL(CLV_ENERGY1, downwind));
  ST_STREAM(AT(mf_id, i, j), mf);
  ST_STREAM(AT(CLV_WORK7, i, j), mf * (e_d + limiter));
#undef AL
}

template <int DIR>
__global__ __launch_bounds__(256) void advec_cell_update_k(Dev d) {
  RANGE_GUARD(2, d.nx + 2, 2, d.ny + 2)
  const int vf_id = DIR == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y;
  const int mf_id = DIR == 1 ? CLV_MASS_FLUX_X : CLV_MASS_FLUX_Y;
  const int i1 = DIR == 1 ? i + 1 : i, j1 = DIR == 1 ? j : j + 1;
  const double pre = AT(CLV_WORK1, i, j);
  double pre_mass_s = AT(CLV_DENSITY1, i, j) * pre;
  double post_mass_s = pre_mass_s + AT(mf_id, i, j) - AT(mf_id, i1, j1);
  double post_ener_s = (AT(CLV_ENERGY1, i, j) * pre_mass_s + AT(CLV_WORK7, i, j) - AT(CLV_WORK7, i1, j1)) / post_mass_s;
  double advec_vol_s = pre + AT(vf_id, i, j) - AT(vf_id, i1, j1);
  ST_STREAM(AT(CLV_DENSITY1, i, j), post_mass_s / advec_vol_s);
  ST_STREAM(AT(CLV_ENERGY1, i, j), post_ener_s);
}

// ---- advec_mom: src/serial/advec_mom.cpp:37-224 ------------------------------------------
// node_flux = work1, node_mass_post = work2, node_mass_pre = work3, mom_flux = work4,
// pre_vol = work5, post_vol = work6 (src/serial/advec_mom.cpp:234-242)
template <int MOM_SWEEP>
__global__ __launch_bounds__(256) void advec_mom_vol_k(Dev d) {
  RANGE_GUARD(0, d.nx + 4, 0, d.ny + 4)
  double post, pre;
  if (MOM_SWEEP == 1) {
    post = geom_volume(d, i, j) + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
    pre = post + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
  } else if (MOM_SWEEP == 2) {
    post = geom_volume(d, i, j) + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
    pre = post + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
  } else if (MOM_SWEEP == 3) {
    post = geom_volume(d, i, j);
    pre = post + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
  } else {
    post = geom_volume(d, i, j);
    pre = post + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
  }
  ST_STREAM(AT(CLV_WORK6, i, j), post);
  ST_STREAM(AT(CLV_WORK5, i, j), pre);
}

template <int DIR>
__global__ __launch_bounds__(256) void advec_mom_node_flux_k(Dev d) {
  if (DIR == 1) {
    RANGE_GUARD(0, d.nx + 4, 2, d.ny + 3)
    AT(CLV_WORK1, i, j) = 0.25 * (AT(CLV_MASS_FLUX_X, i, j - 1) + AT(CLV_MASS_FLUX_X, i, j) + AT(CLV_MASS_FLUX_X, i + 1, j - 1) +
                                  AT(CLV_MASS_FLUX_X, i + 1, j));
  } else {
    RANGE_GUARD(2, d.nx + 3, 0, d.ny + 4)
    AT(CLV_WORK1, i, j) = 0.25 * (AT(CLV_MASS_FLUX_Y, i - 1, j) + AT(CLV_MASS_FLUX_Y, i, j) + AT(CLV_MASS_FLUX_Y, i - 1, j + 1) +
                                  AT(CLV_MASS_FLUX_Y, i, j + 1));
  }
}

template <int DIR>
__global__ __launch_bounds__(256) void advec_mom_node_mass_k(Dev d) {
  // x: j in [2, ny+3), i in [1, nx+4); y: j in [1, ny+4), i in [2, nx+3)
  RANGE_GUARD(DIR == 1 ? 1 : 2, DIR == 1 ? d.nx + 4 : d.nx + 3, DIR == 1 ? 2 : 1, DIR == 1 ? d.ny + 3 : d.ny + 4)
  double post = 0.25 * (AT(CLV_DENSITY1, i, j - 1) * AT(CLV_WORK6, i, j - 1) + AT(CLV_DENSITY1, i, j) * AT(CLV_WORK6, i, j) +
                        AT(CLV_DENSITY1, i - 1, j - 1) * AT(CLV_WORK6, i - 1, j - 1) + AT(CLV_DENSITY1, i - 1, j) * AT(CLV_WORK6, i - 1, j));
  ST_STREAM(AT(CLV_WORK2, i, j), post);
  if (DIR == 1) {
    ST_STREAM(AT(CLV_WORK3, i, j), post - AT(CLV_WORK1, i - 1, j) + AT(CLV_WORK1, i, j));
  } else {
    ST_STREAM(AT(CLV_WORK3, i, j), post - AT(CLV_WORK1, i, j - 1) + AT(CLV_WORK1, i, j));
  }
}

// Limited momentum flux (src/serial/advec_mom.cpp:114-139 and :186-211).
__device__ __forceinline__ double mom_flux_value(double flux, double mass_pre_donor, double width, double width_dif, double v_uw,
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

template <int DIR>
__global__ __launch_bounds__(256) void advec_mom_flux_k(Dev d, int vel_id) {
  // x: j in [2, ny+3), i in [1, nx+3); y: j in [1, ny+3), i in [2, nx+3)
  RANGE_GUARD(DIR == 1 ? 1 : 2, d.nx + 3, DIR == 1 ? 2 : 1, d.ny + 3)
  const double nf = AT(CLV_WORK1, i, j);
  const int s = DIR == 1 ? i : j;
  int upwind, donor, downwind, dif;
  if (nf < 0.0) {
    upwind = s + 2;
    donor = s + 1;
    downwind = s;
    dif = donor;
  } else {
    upwind = s - 1;
    donor = s;
    downwind = s + 1;
    dif = upwind;
  }
  double r;
  if (DIR == 1) {
    r = mom_flux_value(nf, AT(CLV_WORK3, donor, j), M(CLV_CELLDX, s), M(CLV_CELLDX, dif), AT(vel_id, upwind, j), AT(vel_id, donor, j),
                       AT(vel_id, downwind, j));
  } else {
    r = mom_flux_value(nf, AT(CLV_WORK3, i, donor), M(CLV_CELLDY, s), M(CLV_CELLDY, dif), AT(vel_id, i, upwind), AT(vel_id, i, donor),
                       AT(vel_id, i, downwind));
  }
  ST_STREAM(AT(CLV_WORK4, i, j), r);
}

template <int DIR>
__global__ __launch_bounds__(256) void advec_mom_update_k(Dev d, int vel_id) {
  RANGE_GUARD(2, d.nx + 3, 2, d.ny + 3)
  const double mf_m = DIR == 1 ? AT(CLV_WORK4, i - 1, j) : AT(CLV_WORK4, i, j - 1);
  ST_STREAM(AT(vel_id, i, j), (AT(vel_id, i, j) * AT(CLV_WORK3, i, j) + mf_m - AT(CLV_WORK4, i, j)) / AT(CLV_WORK2, i, j));
}

}  // namespace

// ---- host wrappers ------------------------------------------------------------------------
#define LAUNCH(name, kern, grid, ...)                                             \
  do {                                                                            \
    before_launch(c, name);                                                       \
    hipLaunchKernelGGL(kern, grid, dim3(BX, BY, 1), 0, c->stream, __VA_ARGS__);   \
    clv_status st_ = after_launch(c, name);                                       \
    if (st_ != CLV_OK) return st_;                                                \
  } while (0)

clv_status k_ideal_gas(clv_chunk *c, int predict) {
  const Dev &d = c->d;
  LAUNCH("ideal_gas", ideal_gas_k, grid2(2, d.nx + 2, 2, d.ny + 2), d, predict ? CLV_DENSITY1 : CLV_DENSITY0,
         predict ? CLV_ENERGY1 : CLV_ENERGY0);
  return CLV_OK;
}

clv_status k_viscosity(clv_chunk *c) {
  const Dev &d = c->d;
  LAUNCH("viscosity", viscosity_k, grid2(2, d.nx + 2, 2, d.ny + 2), d);
  return CLV_OK;
}

clv_status k_PdV(clv_chunk *c, int predict, double dt) {
  const Dev &d = c->d;
  if (predict) LAUNCH("PdV_predict", pdv_k<true>, grid2(2, d.nx + 2, 2, d.ny + 2), d, dt);
  else LAUNCH("PdV_correct", pdv_k<false>, grid2(2, d.nx + 2, 2, d.ny + 2), d, dt);
  return CLV_OK;
}

clv_status k_revert(clv_chunk *c) {
  const Dev &d = c->d;
  LAUNCH("revert", revert_k, grid2(2, d.nx + 2, 2, d.ny + 2), d);
  return CLV_OK;
}

clv_status k_accelerate(clv_chunk *c, double dt) {
  const Dev &d = c->d;
  const double halfdt = 0.5 * dt;  // src/serial/accelerate.cpp:33
  LAUNCH("accelerate", accelerate_k, grid2(2, d.nx + 3, 2, d.ny + 3), d, halfdt);
  return CLV_OK;
}

clv_status k_flux_calc(clv_chunk *c, double dt) {
  const Dev &d = c->d;
  LAUNCH("flux_calc", flux_calc_k, grid2(2, d.nx + 3, 2, d.ny + 3), d, dt);
  return CLV_OK;
}

clv_status k_flux_calc_edges(clv_chunk *c, double dt) {
  const Dev &d = c->d;
  const int n = (d.nx > d.ny ? d.nx : d.ny) + 1;
  LAUNCH("flux_calc_edges", flux_calc_edges_k, dim3((unsigned)((n + BX * BY - 1) / (BX * BY)), 1, 1), d, dt);
  return CLV_OK;
}

clv_status k_reset_field(clv_chunk *c) {
  const Dev &d = c->d;
  LAUNCH("reset_field", reset_field_k, grid2(2, d.nx + 3, 2, d.ny + 3), d);
  return CLV_OK;
}

clv_status k_advec_cell(clv_chunk *c, int sweep_number, int dir) {
  const Dev &d = c->d;
  const dim3 gv = grid2(0, d.nx + 4, 0, d.ny + 4);
  if (dir == 1) {
    if (sweep_number == 1) LAUNCH("advec_cell_vol", (advec_cell_vol_k<1, 1>), gv, d);
    else LAUNCH("advec_cell_vol", (advec_cell_vol_k<1, 2>), gv, d);
    LAUNCH("advec_cell_flux_x", advec_cell_flux_k<1>, grid2(2, d.nx + 4, 2, d.ny + 2), d);
    LAUNCH("advec_cell_update_x", advec_cell_update_k<1>, grid2(2, d.nx + 2, 2, d.ny + 2), d);
  } else {
    if (sweep_number == 1) LAUNCH("advec_cell_vol", (advec_cell_vol_k<2, 1>), gv, d);
    else LAUNCH("advec_cell_vol", (advec_cell_vol_k<2, 2>), gv, d);
    LAUNCH("advec_cell_flux_y", advec_cell_flux_k<2>, grid2(2, d.nx + 2, 2, d.ny + 4), d);
    LAUNCH("advec_cell_update_y", advec_cell_update_k<2>, grid2(2, d.nx + 2, 2, d.ny + 2), d);
  }
  return CLV_OK;
}

clv_status k_advec_mom(clv_chunk *c, int which_vel, int direction, int sweep_number) {
  const Dev &d = c->d;
  const int vel = which_vel == 1 ? CLV_XVEL1 : CLV_YVEL1;
  const int mom_sweep = direction + 2 * (sweep_number - 1);
  const dim3 gv = grid2(0, d.nx + 4, 0, d.ny + 4);
  switch (mom_sweep) {
    case 1: LAUNCH("advec_mom_vol", advec_mom_vol_k<1>, gv, d); break;
    case 2: LAUNCH("advec_mom_vol", advec_mom_vol_k<2>, gv, d); break;
    case 3: LAUNCH("advec_mom_vol", advec_mom_vol_k<3>, gv, d); break;
    default: LAUNCH("advec_mom_vol", advec_mom_vol_k<4>, gv, d); break;
  }
  if (direction == 1) {
    if (which_vel == 1) {
      LAUNCH("advec_mom_node_flux_x", advec_mom_node_flux_k<1>, grid2(0, d.nx + 4, 2, d.ny + 3), d);
      LAUNCH("advec_mom_node_mass_x", advec_mom_node_mass_k<1>, grid2(1, d.nx + 4, 2, d.ny + 3), d);
    }
    LAUNCH("advec_mom_flux_x", advec_mom_flux_k<1>, grid2(1, d.nx + 3, 2, d.ny + 3), d, vel);
    LAUNCH("advec_mom_update_x", advec_mom_update_k<1>, grid2(2, d.nx + 3, 2, d.ny + 3), d, vel);
  } else {
    if (which_vel == 1) {
      LAUNCH("advec_mom_node_flux_y", advec_mom_node_flux_k<2>, grid2(2, d.nx + 3, 0, d.ny + 4), d);
      LAUNCH("advec_mom_node_mass_y", advec_mom_node_mass_k<2>, grid2(2, d.nx + 3, 1, d.ny + 4), d);
    }
    LAUNCH("advec_mom_flux_y", advec_mom_flux_k<2>, grid2(2, d.nx + 3, 1, d.ny + 3), d, vel);
    LAUNCH("advec_mom_update_y", advec_mom_update_k<2>, grid2(2, d.nx + 3, 2, d.ny + 3), d, vel);
  }
  return CLV_OK;
}

}  // namespace clv
