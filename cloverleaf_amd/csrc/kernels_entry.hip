// Fused kernels behind the plugin model's own entry points (CLV_ENTRY_FUSED).
//
// The reference's model API hands the driver one entry point per phase, and several
// of those run a chain of loops over the whole grid:
//   advec_cell_driver   3 loops: pre/post volume, face fluxes, update
//                       (src/serial/advec_cell.cpp:37-219)
//   advec_mom_driver    5 (xvel) or 3 (yvel) loops: volumes, node flux, node masses,
//                       momentum flux, update (src/serial/advec_mom.cpp:37-224)
//   PdV(predict)        PdV, ideal_gas(predict), update_halo(pressure, 1), revert
//                       (src/serial/PdV.cpp:92-133)
// Here each is one launch, one thread per output element, with the values the
// reference passes between its loops through work arrays recomputed from the same
// inputs by the same expressions (so the same doubles) and kept in registers.
//
// In-place updates.  advec_cell rewrites density1/energy1 and advec_mom the velocity
// while neighbouring threads still read the old values, so the new values go to a
// spare work array and the host exchanges the two arrays' pointers (chunk.cpp).
// Afterwards every array the reference reads later holds the reference's values:
// density1, energy1, the velocity and the mass flux over the reference's ranges.
// Work arrays hold other values (they are scratch between entry points), and the
// ghost layers of an exchanged array hold the spare's old contents until the halo
// update that the driver runs before any read of them (driver/advection.cpp:42-46,
// 57-60; reset_field copies interior nodes only).
#include "clv_internal.h"
#include "geom.h"

namespace clv {
namespace {

#define AT(fid, i, j) d.f[fid][(int64_t)(j) * d.pitch + (i)]
#define M(mid, i) d.m[mid][i]

constexpr int BX = 64, BY = 4;

inline dim3 grid2(int x0, int x1, int y0, int y1) {
  return dim3((unsigned)((x1 - x0 + BX - 1) / BX), (unsigned)((y1 - y0 + BY - 1) / BY), 1);
}

// ---- PdV(predict) chain: src/serial/PdV.cpp:46-66 + ideal_gas.cpp:40-46 + the depth-1
// pressure halo + revert.cpp:35-40 ------------------------------------------------------------
// For a tile whose four faces are external (update_halo's reflective pass on every
// face): each boundary cell writes its own pressure to the ghosts that mirror it --
// ghost (1, j) <- (2, j), (nx + 2, j) <- (nx + 1, j), likewise in y, and a corner ghost
// <- the diagonal interior cell, which is where the reference's bottom/top then
// left/right passes leave it (src/serial/update_halo.cpp:84-164).
__global__ __launch_bounds__(256) void pdv_predict_chain_k(Dev d, double dt) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  if (i >= d.nx + 2 || j >= d.ny + 2) return;
  const double u00 = AT(CLV_XVEL0, i, j), u01 = AT(CLV_XVEL0, i, j + 1), u10 = AT(CLV_XVEL0, i + 1, j), u11 = AT(CLV_XVEL0, i + 1, j + 1);
  const double v00 = AT(CLV_YVEL0, i, j), v01 = AT(CLV_YVEL0, i, j + 1), v10 = AT(CLV_YVEL0, i + 1, j), v11 = AT(CLV_YVEL0, i + 1, j + 1);
  // PdV predictor (the reference's xvel1 argument is xvel0 here, :49-57)
  double left_flux = (geom_xarea(d, i, j) * (u00 + u01 + u00 + u01)) * 0.25 * dt * 0.5;
  double right_flux = (geom_xarea(d, i + 1, j) * (u10 + u11 + u10 + u11)) * 0.25 * dt * 0.5;
  double bottom_flux = (geom_yarea(d, i, j) * (v00 + v10 + v00 + v10)) * 0.25 * dt * 0.5;
  double top_flux = (geom_yarea(d, i, j + 1) * (v01 + v11 + v01 + v11)) * 0.25 * dt * 0.5;
  double total_flux = right_flux - left_flux + top_flux - bottom_flux;
  const double vol = geom_volume(d, i, j);
  double volume_change_s = vol / (vol + total_flux);
  double recip_volume = 1.0 / vol;
  const double rho0 = AT(CLV_DENSITY0, i, j), e0 = AT(CLV_ENERGY0, i, j);
  double energy_change = (AT(CLV_PRESSURE, i, j) / rho0 + AT(CLV_VISCOSITY, i, j) / rho0) * total_flux * recip_volume;
  const double e1 = e0 - energy_change;
  const double rho1 = rho0 * volume_change_s;
  // ideal_gas(predict)
  double v = 1.0 / rho1;
  double p = (1.4 - 1.0) * rho1 * e1;
  double pressurebyenergy = (1.4 - 1.0) * rho1;
  double pressurebyvolume = -rho1 * p;
  double sound_speed_squared = v * v * (p * pressurebyenergy - pressurebyvolume);
  AT(CLV_PRESSURE, i, j) = p;
  AT(CLV_SOUNDSPEED, i, j) = sqrt(sound_speed_squared);
  // update_halo(pressure, 1)
  const int gi = i == 2 ? 1 : (i == d.nx + 1 ? d.nx + 2 : 0);
  const int gj = j == 2 ? 1 : (j == d.ny + 1 ? d.ny + 2 : 0);
  if (gi) AT(CLV_PRESSURE, gi, j) = p;
  if (gj) AT(CLV_PRESSURE, i, gj) = p;
  if (gi && gj) AT(CLV_PRESSURE, gi, gj) = p;
  // revert
  AT(CLV_DENSITY1, i, j) = rho0;
  AT(CLV_ENERGY1, i, j) = e0;
}

// ---- advec_cell: one launch per call (src/serial/advec_cell.cpp:37-219) ------------------------
// pre_vol of cell (i, j), :46-57 (x) and :139-150 (y)
template <int DIR, int SWEEP>
__device__ __forceinline__ double cell_pre_vol(const Dev &d, int i, int j) {
  const double vol = geom_volume(d, i, j);
  if (DIR == 1) {
    if (SWEEP == 1) return vol + (AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j) + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j));
    return vol + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
  }
  if (SWEEP == 1) return vol + (AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j) + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j));
  return vol + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
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

// Mass and energy flux through face s of the line (column i for DIR 2, row j for
// DIR 1), :71-125 (x) and :163-216 (y).  The upwind index is clamped at the last
// ghost (:77, :169).
template <int DIR, int SWEEP>
__device__ __forceinline__ void cell_face_flux(const Dev &d, int i, int j, int s, double &mf, double &ef) {
  const int fi = DIR == 1 ? s : i, fj = DIR == 1 ? j : s;
  const double vf = DIR == 1 ? AT(CLV_VOL_FLUX_X, fi, fj) : AT(CLV_VOL_FLUX_Y, fi, fj);
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
  const double pre_d = DIR == 1 ? cell_pre_vol<DIR, SWEEP>(d, donor, j) : cell_pre_vol<DIR, SWEEP>(d, i, donor);
  double sigmat = fabs(vf) / pre_d;
  double sigma3 = (1.0 + sigmat) * vd;
  double sigma4 = 2.0 - sigmat;
  double sigmav = sigmat;
  const double rho_d = AL(CLV_DENSITY1, donor);
  double limiter = vl_limiter(sigmav, sigma3, sigma4, AL(CLV_DENSITY1, upwind), rho_d, AL(CLV_DENSITY1, downwind));
  mf = vf * (rho_d + limiter);
  double sigmam = fabs(mf) / (rho_d * pre_d);
  const double e_d = AL(CLV_ENERGY1, donor);
  limiter = vl_limiter(sigmam, sigma3, sigma4, AL(CLV_ENERGY1, upwind), e_d, AL(CLV_ENERGY1, downwind));
  ef = mf * (e_d + limiter);
#undef AL
}

// Threads over the reference's face range (x: faces [2, nx + 4) of rows [2, ny + 2);
// y: faces [2, ny + 4) of columns [2, nx + 2)): each writes its face's mass flux, and
// the thread of an interior cell s also forms face s + 1 and updates the cell
// (:128-136 / :205-216) into rho_out / e_out.
template <int DIR, int SWEEP>
__global__ __launch_bounds__(256) void advec_cell_entry_k(Dev d, double *__restrict__ rho_out, double *__restrict__ e_out) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  if (i >= (DIR == 1 ? d.nx + 4 : d.nx + 2) || j >= (DIR == 1 ? d.ny + 2 : d.ny + 4)) return;
  const int s = DIR == 1 ? i : j;
  double mf0, ef0;
  cell_face_flux<DIR, SWEEP>(d, i, j, s, mf0, ef0);
  if (DIR == 1) AT(CLV_MASS_FLUX_X, i, j) = mf0;
  else AT(CLV_MASS_FLUX_Y, i, j) = mf0;
  if (s > (DIR == 1 ? d.nx + 1 : d.ny + 1)) return;
  double mf1, ef1;
  cell_face_flux<DIR, SWEEP>(d, i, j, s + 1, mf1, ef1);
  const int i1 = DIR == 1 ? i + 1 : i, j1 = DIR == 1 ? j : j + 1;
  const double pre = cell_pre_vol<DIR, SWEEP>(d, i, j);
  double pre_mass_s = AT(CLV_DENSITY1, i, j) * pre;
  double post_mass_s = pre_mass_s + mf0 - mf1;
  double post_ener_s = (AT(CLV_ENERGY1, i, j) * pre_mass_s + ef0 - ef1) / post_mass_s;
  const int vf_id = DIR == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y;
  double advec_vol_s = pre + AT(vf_id, i, j) - AT(vf_id, i1, j1);
  const int64_t o = (int64_t)j * d.pitch + i;
  rho_out[o] = post_mass_s / advec_vol_s;
  e_out[o] = post_ener_s;
}

// ---- advec_mom: one launch per call (src/serial/advec_mom.cpp:37-224) -------------------------
// post_vol of cell (i, j) for the momentum sweep (:41-78)
template <int MOM_SWEEP>
__device__ __forceinline__ double mom_post_vol(const Dev &d, int i, int j) {
  if (MOM_SWEEP == 1) return geom_volume(d, i, j) + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
  if (MOM_SWEEP == 2) return geom_volume(d, i, j) + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
  return geom_volume(d, i, j);
}

// node flux of node (i, j), :81-89 (x) / :153-161 (y)
template <int DIR>
__device__ __forceinline__ double mom_node_flux(const Dev &d, int i, int j) {
  if (DIR == 1)
    return 0.25 * (AT(CLV_MASS_FLUX_X, i, j - 1) + AT(CLV_MASS_FLUX_X, i, j) + AT(CLV_MASS_FLUX_X, i + 1, j - 1) + AT(CLV_MASS_FLUX_X, i + 1, j));
  return 0.25 * (AT(CLV_MASS_FLUX_Y, i - 1, j) + AT(CLV_MASS_FLUX_Y, i, j) + AT(CLV_MASS_FLUX_Y, i - 1, j + 1) + AT(CLV_MASS_FLUX_Y, i, j + 1));
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

// One thread per node of [2, nx + 3) x [2, ny + 3).  Along the sweep (t = node index
// along DIR, o = the other index) it needs node fluxes at t - 2 .. t + 1, node masses
// at t - 1 .. t + 1 and the velocity at t - 2 .. t + 2, all inside the ranges the
// reference's loops fill (node flux [0, n + 4), node masses [1, n + 4)).  The yvel
// call recomputes the node flux and masses the reference reuses from the xvel call
// (:81-103 run only for xvel): their inputs are unchanged between the two calls.
template <int DIR, int MOM_SWEEP>
__global__ __launch_bounds__(256) void advec_mom_entry_k(Dev d, int vel_id, double *__restrict__ vel_out) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  if (i >= d.nx + 3 || j >= d.ny + 3) return;
  // point (k) along the sweep: node (k, j) for DIR 1, (i, k) for DIR 2
  auto node_flux = [&](int k) { return DIR == 1 ? mom_node_flux<1>(d, k, j) : mom_node_flux<2>(d, i, k); };
  // node_mass_post (:91-95 / :163-167)
  auto node_mass_post = [&](int k) {
    const int ni = DIR == 1 ? k : i, nj = DIR == 1 ? j : k;
    return 0.25 * (AT(CLV_DENSITY1, ni, nj - 1) * mom_post_vol<MOM_SWEEP>(d, ni, nj - 1) +
                   AT(CLV_DENSITY1, ni, nj) * mom_post_vol<MOM_SWEEP>(d, ni, nj) +
                   AT(CLV_DENSITY1, ni - 1, nj - 1) * mom_post_vol<MOM_SWEEP>(d, ni - 1, nj - 1) +
                   AT(CLV_DENSITY1, ni - 1, nj) * mom_post_vol<MOM_SWEEP>(d, ni - 1, nj));
  };
  const int t = DIR == 1 ? i : j;
  double nf[4], post[3], pre[3], vel[5];
  for (int q = 0; q < 4; ++q) nf[q] = node_flux(t - 2 + q);  // t - 2 .. t + 1
  for (int q = 0; q < 3; ++q) {                               // t - 1 .. t + 1
    post[q] = node_mass_post(t - 1 + q);
    pre[q] = post[q] - nf[q] + nf[q + 1];                     // :99-103 / :171-175
  }
  for (int q = 0; q < 5; ++q) vel[q] = DIR == 1 ? AT(vel_id, t - 2 + q, j) : AT(vel_id, i, t - 2 + q);
  const int W = DIR == 1 ? CLV_CELLDX : CLV_CELLDY;
  // momentum flux through point a (= t - 1 or t): node flux nf[a - t + 2]
  auto mom_flux = [&](int a) {
    const double f = nf[a - t + 2];
    if (f < 0.0) {  // upwind a + 2, donor a + 1, downwind a, dif donor
      return mom_flux_value(f, pre[a + 1 - t + 1], M(W, a), M(W, a + 1), vel[a + 2 - t + 2], vel[a + 1 - t + 2], vel[a - t + 2]);
    }
    // upwind a - 1, donor a, downwind a + 1, dif upwind
    return mom_flux_value(f, pre[a - t + 1], M(W, a), M(W, a - 1), vel[a - 1 - t + 2], vel[a - t + 2], vel[a + 1 - t + 2]);
  };
  const double mf_m = mom_flux(t - 1), mf_0 = mom_flux(t);
  vel_out[(int64_t)j * d.pitch + i] = (vel[2] * pre[1] + mf_m - mf_0) / post[1];  // :141-148 / :213-220
}

}  // namespace

#define LAUNCH_E(name, kern, grid, ...)                                           \
  do {                                                                            \
    before_launch(c, name);                                                       \
    hipLaunchKernelGGL(kern, grid, dim3(BX, BY, 1), 0, c->stream, __VA_ARGS__);   \
    clv_status st_ = after_launch(c, name);                                       \
    if (st_ != CLV_OK) return st_;                                                \
  } while (0)

clv_status e_PdV_predict_chain(clv_chunk *c, double dt) {
  const Dev &d = c->d;
  LAUNCH_E("PdV_predict_chain", pdv_predict_chain_k, grid2(2, d.nx + 2, 2, d.ny + 2), d, dt);
  return CLV_OK;
}

clv_status e_advec_cell(clv_chunk *c, int sweep_number, int dir, double *rho_out, double *e_out) {
  const Dev &d = c->d;
  if (dir == 1) {
    const dim3 g = grid2(2, d.nx + 4, 2, d.ny + 2);
    if (sweep_number == 1) LAUNCH_E("advec_cell_entry_x", (advec_cell_entry_k<1, 1>), g, d, rho_out, e_out);
    else LAUNCH_E("advec_cell_entry_x", (advec_cell_entry_k<1, 2>), g, d, rho_out, e_out);
  } else {
    const dim3 g = grid2(2, d.nx + 2, 2, d.ny + 4);
    if (sweep_number == 1) LAUNCH_E("advec_cell_entry_y", (advec_cell_entry_k<2, 1>), g, d, rho_out, e_out);
    else LAUNCH_E("advec_cell_entry_y", (advec_cell_entry_k<2, 2>), g, d, rho_out, e_out);
  }
  return CLV_OK;
}

clv_status e_advec_mom(clv_chunk *c, int which_vel, int direction, int sweep_number, double *vel_out) {
  const Dev &d = c->d;
  const int vel = which_vel == 1 ? CLV_XVEL1 : CLV_YVEL1;
  const int mom_sweep = direction + 2 * (sweep_number - 1);
  const dim3 g = grid2(2, d.nx + 3, 2, d.ny + 3);
  const char *name = direction == 1 ? "advec_mom_entry_x" : "advec_mom_entry_y";
  switch (mom_sweep) {
    case 1: LAUNCH_E(name, (advec_mom_entry_k<1, 1>), g, d, vel, vel_out); break;  // x, sweep 1
    case 2: LAUNCH_E(name, (advec_mom_entry_k<2, 2>), g, d, vel, vel_out); break;  // y, sweep 1
    case 3: LAUNCH_E(name, (advec_mom_entry_k<1, 3>), g, d, vel, vel_out); break;  // x, sweep 2
    default: LAUNCH_E(name, (advec_mom_entry_k<2, 4>), g, d, vel, vel_out); break; // y, sweep 2
  }
  return CLV_OK;
}

}  // namespace clv
