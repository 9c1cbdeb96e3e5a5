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
// Here each is one launch, with the values the reference passes between its loops
// through work arrays recomputed from the same inputs by the same expressions (so the
// same doubles) and kept in registers: x sweeps take one row per wave with the
// neighbouring columns from lane shifts, y sweeps march each lane up a segment of its
// column with the rows the stencil reads in registers.
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
#include "lanes.h"

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
  ST_STREAM(AT(CLV_PRESSURE, i, j), p);
  ST_STREAM(AT(CLV_SOUNDSPEED, i, j), sqrt(sound_speed_squared));
  // update_halo(pressure, 1)
  const int gi = i == 2 ? 1 : (i == d.nx + 1 ? d.nx + 2 : 0);
  const int gj = j == 2 ? 1 : (j == d.ny + 1 ? d.ny + 2 : 0);
  if (gi) ST_STREAM(AT(CLV_PRESSURE, gi, j), p);
  if (gj) ST_STREAM(AT(CLV_PRESSURE, i, gj), p);
  if (gi && gj) ST_STREAM(AT(CLV_PRESSURE, gi, gj), p);
  // revert
  ST_STREAM(AT(CLV_DENSITY1, i, j), rho0);
  ST_STREAM(AT(CLV_ENERGY1, i, j), e0);
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

// Mass and energy flux of one face from the values the reference's loop reads
// (:71-115 / :163-204): vf the face's volume flux, vd_ratio = vertexd(face) /
// vertexd(dif), pre_d the donor's pre_vol, and density / energy at upwind, donor and
// downwind.
__device__ __forceinline__ void face_flux(double vf, double vd_ratio, double pre_d, double rho_u, double rho_d, double rho_w, double e_u,
                                          double e_d, double e_w, double &mf, double &ef) {
  double sigmat = fabs(vf) / pre_d;
  double sigma3 = (1.0 + sigmat) * vd_ratio;
  double sigma4 = 2.0 - sigmat;
  double sigmav = sigmat;
  double limiter = vl_limiter(sigmav, sigma3, sigma4, rho_u, rho_d, rho_w);
  mf = vf * (rho_d + limiter);
  double sigmam = fabs(mf) / (rho_d * pre_d);
  limiter = vl_limiter(sigmam, sigma3, sigma4, e_u, e_d, e_w);
  ef = mf * (e_d + limiter);
}

// x sweep.  One wave per row j and strip s: lane L holds column i = kOwnEX * s + L;
// faces and cells are owned by lanes 2 .. 1 + kOwnEX, whose stencils (cells i - 2 ..
// i + 1, pre_vol of i - 1 .. i, the flux of face i + 1) come from neighbouring lanes.
// Faces [2, nx + 3] get their mass flux, cells [2, nx + 1] the update (:120-125) into
// rho_out / e_out.  Loads past the arrays are clamped; only unowned lanes see them.
constexpr int kOwnEX = 60;
template <int SWEEP>
__global__ __launch_bounds__(256) void advec_cell_x_entry_k(Dev d, double *__restrict__ rho_out, double *__restrict__ e_out) {
  const int lane = threadIdx.x & 63;
  const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (j >= d.ny + 2) return;  // the whole wave
  const int nx = d.nx;
  const int i = kOwnEX * blockIdx.x + lane;
  const int ic = min(i, nx + 3), iv = min(i, nx + 4);  // cell / face and vertex columns
  const double rho = AT(CLV_DENSITY1, ic, j), e = AT(CLV_ENERGY1, ic, j);
  const double vfx = AT(CLV_VOL_FLUX_X, iv, j), vfx_e = dn1(vfx);
  double pre;
  if (SWEEP == 1) pre = geom_volume(d, ic, j) + (vfx_e - vfx + AT(CLV_VOL_FLUX_Y, ic, j + 1) - AT(CLV_VOL_FLUX_Y, ic, j));  // :47
  else pre = geom_volume(d, ic, j) + vfx_e - vfx;                                                                        // :57
  const double vd = M(CLV_VERTEXDX, iv);
  const double rho_m1 = up1(rho), rho_m2 = up2(rho), rho_p1 = dn1(rho);
  const double e_m1 = up1(e), e_m2 = up2(e), e_p1 = dn1(e);
  const double pre_m1 = up1(pre), vd_m1 = up1(vd), vd_p1 = dn1(vd);
  double mf, ef;
  if (vfx > 0.0) {  // upwind i - 2, donor i - 1, downwind i, dif = donor
    face_flux(vfx, vd / vd_m1, pre_m1, rho_m2, rho_m1, rho, e_m2, e_m1, e, mf, ef);
  } else {  // upwind min(i + 1, nx + 2) (:77), donor i, downwind i - 1, dif = upwind
    const bool east = i + 1 <= nx + 2, self = i == nx + 2;
    const double ru = east ? rho_p1 : (self ? rho : rho_m1), eu = east ? e_p1 : (self ? e : e_m1);
    const double vdu = east ? vd_p1 : (self ? vd : vd_m1);
    face_flux(vfx, vd / vdu, pre, ru, rho, rho_m1, eu, e, e_m1, mf, ef);
  }
  const double mf_e = dn1(mf), ef_e = dn1(ef);
  const bool own = lane >= 2 && lane <= 1 + kOwnEX;
  if (!own) return;
  if (i <= nx + 3) ST_STREAM(AT(CLV_MASS_FLUX_X, i, j), mf);
  if (i > nx + 1) return;
  double pre_mass_s = rho * pre;
  double post_mass_s = pre_mass_s + mf - mf_e;
  double post_ener_s = (e * pre_mass_s + ef - ef_e) / post_mass_s;
  double advec_vol_s = pre + vfx - vfx_e;
  const int64_t o = (int64_t)j * d.pitch + i;
  ST_STREAM(rho_out[o], post_mass_s / advec_vol_s);
  ST_STREAM(e_out[o], post_ener_s);
}

// y sweep.  One lane per cell column i in [2, nx + 1], marching up the faces
// [c0, c1) of its segment with the cells and pre_vols the faces read in registers
// (face f reads cells f - 2 .. min(f + 1, ny + 2) and pre_vol f - 1 .. f, :163-204).
// Faces [2, ny + 3] get their mass flux, cells [2, ny + 1] the update (:205-216).
template <int SWEEP>
__global__ __launch_bounds__(256) void advec_cell_y_entry_k(Dev d, int seg, double *__restrict__ rho_out, double *__restrict__ e_out) {
  const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int c0 = 2 + (blockIdx.y * 4 + (threadIdx.x >> 6)) * seg;
  const int ny = d.ny;
  if (i > d.nx + 1 || c0 > ny + 3) return;
  const int c1 = min(c0 + seg, ny + 4);
  auto rho_at = [&](int r) { return AT(CLV_DENSITY1, i, r); };
  auto e_at = [&](int r) { return AT(CLV_ENERGY1, i, r); };
  // window at face f: rows f - 2, f - 1, f, min(f + 1, ny + 2); pre_vol f - 1, f
  double r_m2 = rho_at(c0 - 2), r_m1 = rho_at(c0 - 1), r_0 = rho_at(c0), r_p1 = rho_at(min(c0 + 1, ny + 2));
  double e_m2 = e_at(c0 - 2), e_m1 = e_at(c0 - 1), e_0 = e_at(c0), e_p1 = e_at(min(c0 + 1, ny + 2));
  double p_m1 = cell_pre_vol<2, SWEEP>(d, i, c0 - 1), p_0 = cell_pre_vol<2, SWEEP>(d, i, c0);
  auto flux = [&](int f, double vf, double &mf, double &ef) {
    if (vf > 0.0) {  // upwind f - 2, donor f - 1, downwind f, dif = donor
      face_flux(vf, M(CLV_VERTEXDY, f) / M(CLV_VERTEXDY, f - 1), p_m1, r_m2, r_m1, r_0, e_m2, e_m1, e_0, mf, ef);
    } else {  // upwind min(f + 1, ny + 2), donor f, downwind f - 1, dif = upwind
      face_flux(vf, M(CLV_VERTEXDY, f) / M(CLV_VERTEXDY, min(f + 1, ny + 2)), p_0, r_p1, r_0, r_m1, e_p1, e_0, e_m1, mf, ef);
    }
  };
  double vf0 = AT(CLV_VOL_FLUX_Y, i, c0), mf0, ef0;
  flux(c0, vf0, mf0, ef0);
  for (int c = c0; c < c1; ++c) {
    ST_STREAM(AT(CLV_MASS_FLUX_Y, i, c), mf0);
    if (c + 1 > ny + 3 || (c > ny + 1 && c + 1 >= c1)) break;
    // the window at face c + 1
    const double rho_c = r_0, e_c = e_0, pre_c = p_0;
    r_m2 = r_m1; r_m1 = r_0; r_0 = r_p1; r_p1 = rho_at(min(c + 2, ny + 2));
    e_m2 = e_m1; e_m1 = e_0; e_0 = e_p1; e_p1 = e_at(min(c + 2, ny + 2));
    p_m1 = p_0; p_0 = cell_pre_vol<2, SWEEP>(d, i, c + 1);
    if (c + 1 == ny + 3) r_0 = rho_at(ny + 3), e_0 = e_at(ny + 3);  // the donor of face ny + 3 (vf <= 0) is cell ny + 3
    const double vf1 = AT(CLV_VOL_FLUX_Y, i, c + 1);
    double mf1, ef1;
    flux(c + 1, vf1, mf1, ef1);
    if (c <= ny + 1) {
      double pre_mass_s = rho_c * pre_c;
      double post_mass_s = pre_mass_s + mf0 - mf1;
      double post_ener_s = (e_c * pre_mass_s + ef0 - ef1) / post_mass_s;
      double advec_vol_s = pre_c + vf0 - vf1;
      const int64_t o = (int64_t)c * d.pitch + i;
      ST_STREAM(rho_out[o], post_mass_s / advec_vol_s);
      ST_STREAM(e_out[o], post_ener_s);
    }
    mf0 = mf1;
    ef0 = ef1;
    vf0 = vf1;
  }
}

// ---- advec_mom: one launch per call (src/serial/advec_mom.cpp:37-224) -------------------------
// post_vol of cell (i, j) for the momentum sweep (:41-78)
template <int MOM_SWEEP>
__device__ __forceinline__ double mom_post_vol(const Dev &d, int i, int j) {
  if (MOM_SWEEP == 1) return geom_volume(d, i, j) + AT(CLV_VOL_FLUX_Y, i, j + 1) - AT(CLV_VOL_FLUX_Y, i, j);
  if (MOM_SWEEP == 2) return geom_volume(d, i, j) + AT(CLV_VOL_FLUX_X, i + 1, j) - AT(CLV_VOL_FLUX_X, i, j);
  return geom_volume(d, i, j);
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

// x sweep (mom_sweep 1 or 3).  One wave per node row j and strip s: lane L holds
// column i = kOwnEX * s + L, nodes owned by lanes 2 .. 1 + kOwnEX.  Per lane: the node
// flux of node i (:85-89), the density * post_vol of cells (i, j - 1) and (i, j), the
// node masses of node i (:91-103), the momentum flux through the face between nodes i
// and i + 1 (:106-139), and for owned nodes [2, nx + 2] the update (:141-148) into
// vel_out.  The yvel call recomputes the node flux and masses the reference reuses
// from the xvel call (:81-103 run only for xvel): their inputs do not change between
// the two calls.
template <int MOM_SWEEP>
__global__ __launch_bounds__(256) void advec_mom_x_entry_k(Dev d, int vel_id, double *__restrict__ vel_out) {
  const int lane = threadIdx.x & 63;
  const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (j >= d.ny + 3) return;  // the whole wave
  const int nx = d.nx;
  const int i = kOwnEX * blockIdx.x + lane;
  const int ic = min(i, nx + 3), in = min(i, nx + 4);  // cell / node and face columns
  const double t0 = AT(CLV_DENSITY1, ic, j - 1) * mom_post_vol<MOM_SWEEP>(d, ic, j - 1);
  const double t1 = AT(CLV_DENSITY1, ic, j) * mom_post_vol<MOM_SWEEP>(d, ic, j);
  const double node_mass_post = 0.25 * (t0 + t1 + up1(t0) + up1(t1));
  const double mf0 = AT(CLV_MASS_FLUX_X, in, j - 1), mf1 = AT(CLV_MASS_FLUX_X, in, j);
  const double node_flux = 0.25 * (mf0 + mf1 + dn1(mf0) + dn1(mf1));
  const double node_mass_pre = node_mass_post - up1(node_flux) + node_flux;
  const double vel = AT(vel_id, in, j);
  const double width = M(CLV_CELLDX, ic);
  // every lane shift before the branch: a shift reads nothing from lanes that are off
  const double pre_e = dn1(node_mass_pre), w_e = dn1(width), w_w = up1(width);
  const double vel_e = dn1(vel), vel_ee = dn2(vel), vel_w = up1(vel);
  double mom_flux;
  if (node_flux < 0.0) {  // upwind i + 2, donor i + 1, downwind i, dif = donor
    mom_flux = mom_flux_value(node_flux, pre_e, width, w_e, vel_ee, vel_e, vel);
  } else {  // upwind i - 1, donor i, downwind i + 1, dif = upwind
    mom_flux = mom_flux_value(node_flux, node_mass_pre, width, w_w, vel_w, vel, vel_e);
  }
  const double mom_flux_w = up1(mom_flux);
  if (lane >= 2 && lane <= 1 + kOwnEX && i <= nx + 2)
    ST_STREAM(vel_out[(int64_t)j * d.pitch + i], (vel * node_mass_pre + mom_flux_w - mom_flux) / node_mass_post);
}

// y sweep (mom_sweep 2 or 4).  Lane L holds node column i = 1 + kOwnEY * s + L (nodes
// owned by lanes 1 .. 63; lane 0 supplies column i - 1) and marches up the node rows
// [r0, r1) of its segment with the node fluxes, masses and velocities of the rows the
// momentum flux reads in registers (:153-222).
constexpr int kOwnEY = 63;
template <int MOM_SWEEP>
__global__ __launch_bounds__(256) void advec_mom_y_entry_k(Dev d, int vel_id, int seg, double *__restrict__ vel_out) {
  const int lane = threadIdx.x & 63;
  const int r0 = 2 + (blockIdx.y * 4 + (threadIdx.x >> 6)) * seg;
  const int ny = d.ny, nx = d.nx;
  if (r0 > ny + 2) return;  // the whole wave
  const int r1 = min(r0 + seg, ny + 3);
  const int i = 1 + kOwnEY * blockIdx.x + lane;
  const int ic = min(i, nx + 3), in = min(i, nx + 4);
  const bool own = lane >= 1 && i <= nx + 2;
  // density * post_vol of cell (i, q) and its west neighbour's; mass flux of y face (i, q)
  auto t_of = [&](int q) { return AT(CLV_DENSITY1, ic, q) * mom_post_vol<MOM_SWEEP>(d, ic, q); };
  auto nf_of = [&](double mfa, double mfb) { return 0.25 * (up1(mfa) + mfa + up1(mfb) + mfb); };  // :157-161
  auto nmp_of = [&](double ta, double tb) { return 0.25 * (ta + tb + up1(ta) + up1(tb)); };        // :163-167
  auto vel_at = [&](int q) { return AT(vel_id, in, q); };
  // prologue: rows r0 - 2 .. r0 + 1
  double mf_m2 = AT(CLV_MASS_FLUX_Y, in, r0 - 2), mf_m1 = AT(CLV_MASS_FLUX_Y, in, r0 - 1), mf_0 = AT(CLV_MASS_FLUX_Y, in, r0),
         mf_p1 = AT(CLV_MASS_FLUX_Y, in, r0 + 1);
  double t_m2 = t_of(r0 - 2), t_m1 = t_of(r0 - 1), t_0 = t_of(r0);
  const double nf_m2 = nf_of(mf_m2, mf_m1);
  double nf_m1 = nf_of(mf_m1, mf_0), nf_0 = nf_of(mf_0, mf_p1);
  const double nmp_m1 = nmp_of(t_m2, t_m1);
  double nmp_0 = nmp_of(t_m1, t_0);
  const double pre_m1 = nmp_m1 - nf_m2 + nf_m1;  // :171-175
  double pre_0 = nmp_0 - nf_m1 + nf_0;
  double v_m2 = vel_at(r0 - 2), v_m1 = vel_at(r0 - 1), v_0 = vel_at(r0), v_p1 = vel_at(r0 + 1);
  // momentum flux through the face between node rows a and a + 1
  auto mom_flux = [&](int a, double nf, double pre_a, double pre_a1, double va_m1, double va, double va_p1, double va_p2) {
    if (nf < 0.0)  // upwind a + 2, donor a + 1, downwind a, dif = donor
      return mom_flux_value(nf, pre_a1, M(CLV_CELLDY, a), M(CLV_CELLDY, a + 1), va_p2, va_p1, va);
    // upwind a - 1, donor a, downwind a + 1, dif = upwind
    return mom_flux_value(nf, pre_a, M(CLV_CELLDY, a), M(CLV_CELLDY, a - 1), va_m1, va, va_p1);
  };
  double flux_m1 = mom_flux(r0 - 1, nf_m1, pre_m1, pre_0, v_m2, v_m1, v_0, v_p1);
  for (int r = r0; r < r1; ++r) {
    // row r + 1's node flux and masses, row r + 2's velocity
    const double mf_p2 = AT(CLV_MASS_FLUX_Y, in, r + 2);
    const double t_p1 = t_of(r + 1);
    const double v_p2 = vel_at(r + 2);
    const double nf_p1 = nf_of(mf_p1, mf_p2);
    const double nmp_p1 = nmp_of(t_0, t_p1);
    const double pre_p1 = nmp_p1 - nf_0 + nf_p1;
    const double flux_0 = mom_flux(r, nf_0, pre_0, pre_p1, v_m1, v_0, v_p1, v_p2);
    if (own) vel_out[(int64_t)r * d.pitch + i] = (v_0 * pre_0 + flux_m1 - flux_0) / nmp_0;  // :213-220
    mf_p1 = mf_p2;
    t_0 = t_p1;
    nf_0 = nf_p1;
    nmp_0 = nmp_p1;
    pre_0 = pre_p1;
    v_m1 = v_0;
    v_0 = v_p1;
    v_p1 = v_p2;
    flux_m1 = flux_0;
  }
}

}  // namespace

// 256-thread blocks of four waves
#define LAUNCH_W(name, kern, grid, ...)                                           \
  do {                                                                            \
    before_launch(c, name);                                                       \
    hipLaunchKernelGGL(kern, grid, dim3(256, 1, 1), 0, c->stream, __VA_ARGS__);   \
    clv_status st_ = after_launch(c, name);                                       \
    if (st_ != CLV_OK) return st_;                                                \
  } while (0)

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

// rows of a marching segment of the y kernels
constexpr int kSegEY = 16;

clv_status e_advec_cell(clv_chunk *c, int sweep_number, int dir, double *rho_out, double *e_out) {
  const Dev &d = c->d;
  if (dir == 1) {
    const dim3 g((unsigned)((d.nx + 2 + kOwnEX - 1) / kOwnEX), (unsigned)((d.ny + 3) / 4), 1);  // faces [2, nx + 3], rows [2, ny + 1]
    if (sweep_number == 1) LAUNCH_W("advec_cell_entry_x", advec_cell_x_entry_k<1>, g, d, rho_out, e_out);
    else LAUNCH_W("advec_cell_entry_x", advec_cell_x_entry_k<2>, g, d, rho_out, e_out);
  } else {
    const int segs = (d.ny + 2 + kSegEY - 1) / kSegEY;  // faces [2, ny + 3]
    const dim3 g((unsigned)((d.nx + 63) / 64), (unsigned)((segs + 3) / 4), 1);
    if (sweep_number == 1) LAUNCH_W("advec_cell_entry_y", advec_cell_y_entry_k<1>, g, d, kSegEY, rho_out, e_out);
    else LAUNCH_W("advec_cell_entry_y", advec_cell_y_entry_k<2>, g, d, kSegEY, rho_out, e_out);
  }
  return CLV_OK;
}

clv_status e_advec_mom(clv_chunk *c, int which_vel, int direction, int sweep_number, double *vel_out) {
  const Dev &d = c->d;
  const int vel = which_vel == 1 ? CLV_XVEL1 : CLV_YVEL1;
  const int mom_sweep = direction + 2 * (sweep_number - 1);
  if (direction == 1) {
    const dim3 g((unsigned)((d.nx + 1 + kOwnEX - 1) / kOwnEX), (unsigned)((d.ny + 1 + 3) / 4), 1);  // nodes [2, nx + 2] x [2, ny + 2]
    if (mom_sweep == 1) LAUNCH_W("advec_mom_entry_x", advec_mom_x_entry_k<1>, g, d, vel, vel_out);
    else LAUNCH_W("advec_mom_entry_x", advec_mom_x_entry_k<3>, g, d, vel, vel_out);
  } else {
    const int segs = (d.ny + 1 + kSegEY - 1) / kSegEY;  // node rows [2, ny + 2]
    const dim3 g((unsigned)((d.nx + 1 + kOwnEY - 1) / kOwnEY), (unsigned)((segs + 3) / 4), 1);
    if (mom_sweep == 2) LAUNCH_W("advec_mom_entry_y", advec_mom_y_entry_k<2>, g, d, vel, kSegEY, vel_out);
    else LAUNCH_W("advec_mom_entry_y", advec_mom_y_entry_k<4>, g, d, vel, kSegEY, vel_out);
  }
  return CLV_OK;
}

}  // namespace clv
