// Reverse-mode (adjoint) kernels of the fixed-dt hydro step, one per reference
// kernel, for the gradient of a scalar of the final state with respect to the
// initial fields.
//
// Conventions.  p is the primal chunk holding the forward kernel's INPUT state;
// a is a second chunk of the same size holding adjoints.  On entry a holds the
// adjoints of the kernel's outputs (and whatever has accumulated on its
// inputs); on exit the outputs' adjoints are consumed (zeroed where the kernel
// overwrote them) and the inputs' adjoints are accumulated.  Geometry (volume,
// areas, mesh) is constant; soundspeed, which only calc_dt reads, carries no
// adjoint; work arrays of the adjoint chunk are scratch.
//
// Every kernel is a gather: a thread owns one input element and sums, over the
// few output elements that read it, d(output)/d(input) * adjoint(output) --
// deterministic, no atomics.  The local derivative d(output)/d(input) comes from
// evaluating that output's reference expression on dual numbers with the input
// seeded (several inputs of one element at once with an N-component dual), so
// every branch (upwind choice, van Leer limiter, viscosity switch, fmin/fmax,
// fabs) is differentiated on its active piece exactly as the forward took it.
// Ties follow the oracle's tangent-linear model: fmin/fmax return their first
// argument, fabs(0) has derivative +1.
#include "clv_internal.h"

namespace clv {
namespace {

// ---- N-component dual numbers ------------------------------------------------------------
template <int N>
struct Dn {
  double v;
  double d[N];
};
template <int N>
__device__ __forceinline__ Dn<N> cst(double v) {
  Dn<N> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = 0.0;
  return r;
}
// value v with derivative 1 in each component k whose bit is set in mask
template <int N>
__device__ __forceinline__ Dn<N> var(double v, unsigned mask) {
  Dn<N> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (mask >> k) & 1u ? 1.0 : 0.0;
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator+(Dn<N> a, Dn<N> b) {
  Dn<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator-(Dn<N> a, Dn<N> b) {
  Dn<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator*(Dn<N> a, Dn<N> b) {
  Dn<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator/(Dn<N> a, Dn<N> b) {
  Dn<N> r;
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator-(Dn<N> a) {
  Dn<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator+(Dn<N> a, double b) { return a + cst<N>(b); }
template <int N>
__device__ __forceinline__ Dn<N> operator+(double a, Dn<N> b) { return cst<N>(a) + b; }
template <int N>
__device__ __forceinline__ Dn<N> operator-(Dn<N> a, double b) { return a - cst<N>(b); }
template <int N>
__device__ __forceinline__ Dn<N> operator-(double a, Dn<N> b) { return cst<N>(a) - b; }
template <int N>
__device__ __forceinline__ Dn<N> operator*(Dn<N> a, double b) {
  Dn<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator*(double a, Dn<N> b) {
  Dn<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dn<N> operator/(Dn<N> a, double b) { return a / cst<N>(b); }
template <int N>
__device__ __forceinline__ Dn<N> operator/(double a, Dn<N> b) { return cst<N>(a) / b; }
template <int N>
__device__ __forceinline__ bool operator<(Dn<N> a, double b) { return a.v < b; }
template <int N>
__device__ __forceinline__ bool operator>(Dn<N> a, double b) { return a.v > b; }
template <int N>
__device__ __forceinline__ bool operator<=(Dn<N> a, double b) { return a.v <= b; }
template <int N>
__device__ __forceinline__ bool operator>=(Dn<N> a, double b) { return a.v >= b; }
template <int N>
__device__ __forceinline__ Dn<N> fabs(Dn<N> a) { return a.v < 0.0 ? -a : a; }
template <int N>
__device__ __forceinline__ Dn<N> fmin(Dn<N> a, Dn<N> b) { return b.v < a.v ? b : a; }
template <int N>
__device__ __forceinline__ Dn<N> fmax(Dn<N> a, Dn<N> b) { return b.v > a.v ? b : a; }
template <int N>
__device__ __forceinline__ Dn<N> fmax(double a, Dn<N> b) { return fmax(cst<N>(a), b); }
template <int N>
__device__ __forceinline__ Dn<N> fmax(Dn<N> a, double b) { return fmax(a, cst<N>(b)); }
template <int N>
__device__ __forceinline__ Dn<N> sqrt(Dn<N> a) {
  Dn<N> r;
  r.v = ::sqrt(a.v);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
  return r;
}

#define PA(fid, i, j) p.f[fid][(int64_t)(j) * p.pitch + (i)]
#define AA(fid, i, j) a.f[fid][(int64_t)(j) * a.pitch + (i)]
#define PM(mid, i) p.m[mid][i]

// ---- element functions: the reference expressions (oracle order) on dual numbers ----------
// viscosity of one cell, src/serial/viscosity.cpp:43-66
template <int N>
__device__ Dn<N> el_viscosity(Dn<N> u00, Dn<N> u10, Dn<N> u01, Dn<N> u11, Dn<N> v00, Dn<N> v10, Dn<N> v01, Dn<N> v11, Dn<N> pl,
                              Dn<N> pr, Dn<N> pb, Dn<N> pt, Dn<N> rho, double dxi, double dxi1, double dyj, double dyj2) {
  Dn<N> ugrad = (u10 + u11) - (u00 + u01);
  Dn<N> vgrad = (v01 + v11) - (v00 + v10);
  Dn<N> div = (dxi * (ugrad) + dyj * (vgrad));
  Dn<N> strain2 = 0.5 * (u01 + u11 - u00 - u10) / dyj + 0.5 * (v10 + v11 - v00 - v01) / dxi;
  Dn<N> pgradx = (pr - pl) / (dxi + dxi1);
  Dn<N> pgrady = (pt - pb) / (dyj + dyj2);
  Dn<N> pgradx2 = pgradx * pgradx;
  Dn<N> pgrady2 = pgrady * pgrady;
  Dn<N> limiter = ((0.5 * (ugrad) / dxi) * pgradx2 + (0.5 * (vgrad) / dyj) * pgrady2 + strain2 * pgradx * pgrady) /
                  fmax(pgradx2 + pgrady2, g_small);
  if ((limiter > 0.0) || (div >= 0.0)) return cst<N>(0.0);
  double dirx = 1.0;
  if (pgradx < 0.0) dirx = -1.0;
  pgradx = dirx * fmax(g_small, fabs(pgradx));
  double diry = 1.0;
  if (pgradx < 0.0) diry = -1.0;  // the reference tests pgradx here (:57)
  pgrady = diry * fmax(g_small, fabs(pgrady));
  Dn<N> pgrad = sqrt(pgradx * pgradx + pgrady * pgrady);
  Dn<N> xgrad = fabs(dxi * pgrad / pgradx);
  Dn<N> ygrad = fabs(dyj * pgrad / pgrady);
  Dn<N> grad = fmin(xgrad, ygrad);
  Dn<N> grad2 = grad * grad;
  return 2.0 * rho * grad2 * limiter * limiter;
}

// PdV of one cell, src/serial/PdV.cpp:49-63 (predict) / :72-83 (correct)
template <int N>
__device__ void el_pdv(bool predict, double dt, double xa0, double xa1, double ya0, double ya1, Dn<N> u00, Dn<N> u01, Dn<N> u10,
                       Dn<N> u11, Dn<N> v00, Dn<N> v10, Dn<N> v01, Dn<N> v11, Dn<N> w00, Dn<N> w01, Dn<N> w10, Dn<N> w11,
                       Dn<N> z00, Dn<N> z10, Dn<N> z01, Dn<N> z11, double vol, Dn<N> rho0, Dn<N> e0, Dn<N> pr, Dn<N> q,
                       Dn<N> &rho1, Dn<N> &e1) {
  Dn<N> left_flux, right_flux, bottom_flux, top_flux;
  if (predict) {
    left_flux = (xa0 * (u00 + u01 + u00 + u01)) * 0.25 * dt * 0.5;
    right_flux = (xa1 * (u10 + u11 + u10 + u11)) * 0.25 * dt * 0.5;
    bottom_flux = (ya0 * (v00 + v10 + v00 + v10)) * 0.25 * dt * 0.5;
    top_flux = (ya1 * (v01 + v11 + v01 + v11)) * 0.25 * dt * 0.5;
  } else {
    left_flux = (xa0 * (u00 + u01 + w00 + w01)) * 0.25 * dt;
    right_flux = (xa1 * (u10 + u11 + w10 + w11)) * 0.25 * dt;
    bottom_flux = (ya0 * (v00 + v10 + z00 + z10)) * 0.25 * dt;
    top_flux = (ya1 * (v01 + v11 + z01 + z11)) * 0.25 * dt;
  }
  Dn<N> total_flux = right_flux - left_flux + top_flux - bottom_flux;
  Dn<N> volume_change_s = vol / (vol + total_flux);
  double recip_volume = 1.0 / vol;
  Dn<N> energy_change = (pr / rho0 + q / rho0) * total_flux * recip_volume;
  e1 = e0 - energy_change;
  rho1 = rho0 * volume_change_s;
}

// accelerate of one node, src/serial/accelerate.cpp:43-53; cells (i-1,j-1)=mm,
// (i-1,j)=m0, (i,j)=00, (i,j-1)=0m
template <int N>
__device__ void el_accel(double halfdt, Dn<N> rmm, Dn<N> rm0, Dn<N> r00, Dn<N> r0m, double vmm, double vm0, double v00, double v0m,
                         double xa00, double xa0m, double ya00, double yam0, Dn<N> p00, Dn<N> pm0, Dn<N> p0m, Dn<N> pmm, Dn<N> q00,
                         Dn<N> qm0, Dn<N> q0m, Dn<N> qmm, Dn<N> u0, Dn<N> w0, Dn<N> &u1, Dn<N> &w1) {
  Dn<N> stepbymass_s = halfdt / ((rmm * vmm + rm0 * vm0 + r00 * v00 + r0m * v0m) * 0.25);
  u1 = u0 - stepbymass_s * (xa00 * (p00 - pm0) + xa0m * (p0m - pmm));
  w1 = w0 - stepbymass_s * (ya00 * (p00 - p0m) + yam0 * (pm0 - pmm));
  u1 = u1 - stepbymass_s * (xa00 * (q00 - qm0) + xa0m * (q0m - qmm));
  w1 = w1 - stepbymass_s * (ya00 * (q00 - q0m) + yam0 * (qm0 - qmm));
}

// van Leer limiter of advec_cell, src/serial/advec_cell.cpp:88-104
template <int N>
__device__ Dn<N> el_vl(Dn<N> sigma, Dn<N> sigma3, Dn<N> sigma4, Dn<N> uw, Dn<N> dn, Dn<N> dw) {
  const double one_by_six = 1.0 / 6.0;
  Dn<N> diffuw = dn - uw;
  Dn<N> diffdw = dw - dn;
  double wind = 1.0;
  if (diffdw <= 0.0) wind = -1.0;
  if (diffuw.v * diffdw.v > 0.0)
    return (1.0 - sigma) * wind * fmin(fmin(fabs(diffuw), fabs(diffdw)), one_by_six * (sigma3 * fabs(diffuw) + sigma4 * fabs(diffdw)));
  return cst<N>(0.0);
}

// mass and energy flux of one face, src/serial/advec_cell.cpp:84-110 / :176-202
template <int N>
__device__ void el_face(Dn<N> vf, Dn<N> pre_dn, Dn<N> r_uw, Dn<N> r_dn, Dn<N> r_dw, Dn<N> e_uw, Dn<N> e_dn, Dn<N> e_dw, double vd_face,
                        double vd_dif, Dn<N> &mf, Dn<N> &ef) {
  Dn<N> sigmat = fabs(vf) / pre_dn;
  Dn<N> sigma3 = (1.0 + sigmat) * (vd_face / vd_dif);
  Dn<N> sigma4 = 2.0 - sigmat;
  Dn<N> sigmav = sigmat;
  Dn<N> limiter = el_vl(sigmav, sigma3, sigma4, r_uw, r_dn, r_dw);
  mf = vf * (r_dn + limiter);
  Dn<N> sigmam = fabs(mf) / (r_dn * pre_dn);
  limiter = el_vl(sigmam, sigma3, sigma4, e_uw, e_dn, e_dw);
  ef = mf * (e_dn + limiter);
}

// cell update of advec_cell, src/serial/advec_cell.cpp:120-125 / :211-216
template <int N>
__device__ void el_cell_update(Dn<N> rho, Dn<N> e, Dn<N> pre, Dn<N> mf_l, Dn<N> mf_r, Dn<N> ef_l, Dn<N> ef_r, Dn<N> vf_l, Dn<N> vf_r,
                               Dn<N> &rho_new, Dn<N> &e_new) {
  Dn<N> pre_mass_s = rho * pre;
  Dn<N> post_mass_s = pre_mass_s + mf_l - mf_r;
  Dn<N> post_ener_s = (e * pre_mass_s + ef_l - ef_r) / post_mass_s;
  Dn<N> advec_vol_s = pre + vf_l - vf_r;
  rho_new = post_mass_s / advec_vol_s;
  e_new = post_ener_s;
}

// momentum flux of one node face, src/serial/advec_mom.cpp:114-139 / :186-211
template <int N>
__device__ Dn<N> el_momflux(Dn<N> flux, Dn<N> mass_pre_donor, double width, double width_dif, Dn<N> v_uw, Dn<N> v_dn, Dn<N> v_dw) {
  Dn<N> sigma = fabs(flux) / (mass_pre_donor);
  Dn<N> vdiffuw = v_dn - v_uw;
  Dn<N> vdiffdw = v_dw - v_dn;
  Dn<N> limiter = cst<N>(0.0);
  if (vdiffuw.v * vdiffdw.v > 0.0) {
    Dn<N> auw = fabs(vdiffuw);
    Dn<N> adw = fabs(vdiffdw);
    double wind = 1.0;
    if (vdiffdw <= 0.0) wind = -1.0;
    limiter = wind * fmin(fmin(width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / width_dif) / 6.0, auw), adw);
  }
  Dn<N> advec_vel_s = v_dn + (1.0 - sigma) * limiter;
  return advec_vel_s * flux;
}

// velocity update of advec_mom, src/serial/advec_mom.cpp:142-150 / :214-222
template <int N>
__device__ Dn<N> el_vel_update(Dn<N> vel, Dn<N> mpre, Dn<N> mf_l, Dn<N> mf_r, Dn<N> mpost) {
  return (vel * mpre + mf_l - mf_r) / mpost;
}

__device__ __forceinline__ bool in(int i, int lo, int hi) { return i >= lo && i <= hi; }

// ---- ideal_gas, src/serial/ideal_gas.cpp:38-46: p = 0.4 rho e on the interior ---------------
__global__ void adj_ideal_gas_k(Dev p, Dev a, int rho_f, int e_f) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > p.nx + 1 || j > p.ny + 1) return;
  const double pb = AA(CLV_PRESSURE, i, j);
  AA(rho_f, i, j) += (1.4 - 1.0) * PA(e_f, i, j) * pb;
  AA(e_f, i, j) += (1.4 - 1.0) * PA(rho_f, i, j) * pb;
  AA(CLV_PRESSURE, i, j) = 0.0;
}

// ---- update_halo: ghost = sign * source, transposed ----------------------------------------
// The forward runs bottom, top, left, right per field (corners from ghosts the
// bottom/top passes wrote); the transpose runs left/right then bottom/top, each
// ghost adding sign * its adjoint to its source and clearing its own.
struct AdjHaloRule {
  signed char bsrc, tdst, tsrc, lsrc, rdst, rsrc, xext, yext, sy, sx;
};
__constant__ AdjHaloRule kAdjRules[CLV_NUM_HALO_FIELDS] = {
    {2, 2, 1, 2, 2, 1, 0, 0, 1, 1},  {2, 2, 1, 2, 2, 1, 0, 0, 1, 1},  {2, 2, 1, 2, 2, 1, 0, 0, 1, 1},
    {2, 2, 1, 2, 2, 1, 0, 0, 1, 1},  {2, 2, 1, 2, 2, 1, 0, 0, 1, 1},  {2, 2, 1, 2, 2, 1, 0, 0, 1, 1},
    {0, 2, 1, 2, 2, 1, 0, 0, 1, 1},  {3, 3, 1, 3, 3, 1, 1, 1, 1, -1}, {3, 3, 1, 3, 3, 1, 1, 1, 1, -1},
    {3, 3, 1, 3, 3, 1, 1, 1, -1, 1}, {3, 3, 1, 3, 3, 1, 1, 1, -1, 1}, {3, 2, 0, 3, 3, 1, 1, 0, 1, -1},
    {3, 3, 1, 3, 2, 0, 0, 1, -1, 1}, {3, 2, 0, 3, 3, 1, 1, 0, 1, -1}, {3, 3, 1, 3, 2, 0, 0, 1, -1, 1},
};
struct AdjHaloList {
  int n;
  int ids[CLV_NUM_HALO_FIELDS];
};

// one row per thread: left and right ghosts of rows [2 - depth, ny + yext + depth + 1]
__global__ void adj_halo_lr_k(Dev a, AdjHaloList list, int depth) {
  const int fid = list.ids[blockIdx.y];
  const AdjHaloRule r = kAdjRules[fid];
  const int k = 2 - depth + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k > a.ny + r.yext + depth + 1) return;
  double *f = a.f[fid];
  const int64_t row = (int64_t)k * a.pitch;
  for (int jj = depth - 1; jj >= 0; --jj) {  // right then left, each in reverse order
    const int dst = a.nx + r.rdst + jj, src = a.nx + r.rsrc - jj;
    f[row + src] += r.sx * f[row + dst];
    f[row + dst] = 0.0;
  }
  for (int jj = depth - 1; jj >= 0; --jj) {
    const int dst = 1 - jj, src = r.lsrc + jj;
    f[row + src] += r.sx * f[row + dst];
    f[row + dst] = 0.0;
  }
}

// one column per thread: bottom and top ghosts of columns [2 - depth, nx + xext + depth + 1]
__global__ void adj_halo_bt_k(Dev a, AdjHaloList list, int depth) {
  const int fid = list.ids[blockIdx.y];
  const AdjHaloRule r = kAdjRules[fid];
  const int i = 2 - depth + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > a.nx + r.xext + depth + 1) return;
  double *f = a.f[fid];
  const int64_t P = a.pitch;
  for (int kk = depth - 1; kk >= 0; --kk) {
    const int dst = a.ny + r.tdst + kk, src = a.ny + r.tsrc - kk;
    f[(int64_t)src * P + i] += r.sy * f[(int64_t)dst * P + i];
    f[(int64_t)dst * P + i] = 0.0;
  }
  for (int kk = depth - 1; kk >= 0; --kk) {
    const int dst = 1 - kk, src = r.bsrc + kk;
    f[(int64_t)src * P + i] += r.sy * f[(int64_t)dst * P + i];
    f[(int64_t)dst * P + i] = 0.0;
  }
}

// ---- viscosity --------------------------------------------------------------------------
__device__ Dn<2> visc_at(const Dev &p, int ci, int cj, int ni, int nj, int which, int pi, int pj) {
  // which: 0 = seed u0,v0 of node (ni,nj); 1 = seed p(pi,pj) (component 0) and rho0 (component 1, own cell)
  auto U = [&](int i, int j) { return which == 0 && i == ni && j == nj ? var<2>(PA(CLV_XVEL0, i, j), 1u) : cst<2>(PA(CLV_XVEL0, i, j)); };
  auto V = [&](int i, int j) { return which == 0 && i == ni && j == nj ? var<2>(PA(CLV_YVEL0, i, j), 2u) : cst<2>(PA(CLV_YVEL0, i, j)); };
  auto Pp = [&](int i, int j) { return which == 1 && i == pi && j == pj ? var<2>(PA(CLV_PRESSURE, i, j), 1u) : cst<2>(PA(CLV_PRESSURE, i, j)); };
  const Dn<2> rho = which == 1 && ci == pi && cj == pj ? var<2>(PA(CLV_DENSITY0, ci, cj), 2u) : cst<2>(PA(CLV_DENSITY0, ci, cj));
  return el_viscosity<2>(U(ci, cj), U(ci + 1, cj), U(ci, cj + 1), U(ci + 1, cj + 1), V(ci, cj), V(ci + 1, cj), V(ci, cj + 1),
                         V(ci + 1, cj + 1), Pp(ci - 1, cj), Pp(ci + 1, cj), Pp(ci, cj - 1), Pp(ci, cj + 1), rho, PM(CLV_CELLDX, ci),
                         PM(CLV_CELLDX, ci + 1), PM(CLV_CELLDY, cj), PM(CLV_CELLDY, cj + 2));
}

// site (i, j) in [1, nx+2] x [1, ny+2]: node (i, j) velocities, cell (i, j) pressure and density
__global__ void adj_viscosity_k(Dev p, Dev a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i > nx + 2 || j > ny + 2) return;
  // node (i, j): cells (i-1..i, j-1..j)
  if (i >= 2 && j >= 2) {
    double ub = 0.0, vb = 0.0;
    for (int cj = j - 1; cj <= j; ++cj)
      for (int ci = i - 1; ci <= i; ++ci) {
        if (!in(ci, 2, nx + 1) || !in(cj, 2, ny + 1)) continue;
        const double qb = AA(CLV_VISCOSITY, ci, cj);
        if (qb == 0.0) continue;
        const Dn<2> q = visc_at(p, ci, cj, i, j, 0, 0, 0);
        ub += q.d[0] * qb;
        vb += q.d[1] * qb;
      }
    AA(CLV_XVEL0, i, j) += ub;
    AA(CLV_YVEL0, i, j) += vb;
  }
  // cell (i, j): its pressure is read by cells (i+-1, j), (i, j+-1); its density by itself
  double pb = 0.0, rb = 0.0;
  const int nci[5] = {i + 1, i - 1, i, i, i}, ncj[5] = {j, j, j + 1, j - 1, j};
  for (int k = 0; k < 5; ++k) {
    const int ci = nci[k], cj = ncj[k];
    if (!in(ci, 2, nx + 1) || !in(cj, 2, ny + 1)) continue;
    const double qb = AA(CLV_VISCOSITY, ci, cj);
    if (qb == 0.0) continue;
    const Dn<2> q = visc_at(p, ci, cj, 0, 0, 1, i, j);
    if (k < 4) pb += q.d[0] * qb;
    else rb += q.d[1] * qb;
  }
  AA(CLV_PRESSURE, i, j) += pb;
  AA(CLV_DENSITY0, i, j) += rb;
}

// ---- PdV ----------------------------------------------------------------------------------
// cell (ci, cj) with seeds: mode 0 = node (ni, nj) [u0, v0, u1, v1]; mode 1 = own cell [rho0, e0, p, q]
__device__ void pdv_at(const Dev &p, bool predict, double dt, int ci, int cj, int mode, int ni, int nj, Dn<4> &r1, Dn<4> &e1) {
  auto N0 = [&](int fid, int i, int j, unsigned bit) {
    return mode == 0 && i == ni && j == nj ? var<4>(PA(fid, i, j), bit) : cst<4>(PA(fid, i, j));
  };
  const Dn<4> u00 = N0(CLV_XVEL0, ci, cj, 1u), u01 = N0(CLV_XVEL0, ci, cj + 1, 1u), u10 = N0(CLV_XVEL0, ci + 1, cj, 1u),
              u11 = N0(CLV_XVEL0, ci + 1, cj + 1, 1u);
  const Dn<4> v00 = N0(CLV_YVEL0, ci, cj, 2u), v10 = N0(CLV_YVEL0, ci + 1, cj, 2u), v01 = N0(CLV_YVEL0, ci, cj + 1, 2u),
              v11 = N0(CLV_YVEL0, ci + 1, cj + 1, 2u);
  Dn<4> w00 = cst<4>(0.0), w01 = w00, w10 = w00, w11 = w00, z00 = w00, z10 = w00, z01 = w00, z11 = w00;
  if (!predict) {
    w00 = N0(CLV_XVEL1, ci, cj, 4u), w01 = N0(CLV_XVEL1, ci, cj + 1, 4u), w10 = N0(CLV_XVEL1, ci + 1, cj, 4u),
    w11 = N0(CLV_XVEL1, ci + 1, cj + 1, 4u);
    z00 = N0(CLV_YVEL1, ci, cj, 8u), z10 = N0(CLV_YVEL1, ci + 1, cj, 8u), z01 = N0(CLV_YVEL1, ci, cj + 1, 8u),
    z11 = N0(CLV_YVEL1, ci + 1, cj + 1, 8u);
  }
  const bool own = mode == 1;
  const Dn<4> rho0 = own ? var<4>(PA(CLV_DENSITY0, ci, cj), 1u) : cst<4>(PA(CLV_DENSITY0, ci, cj));
  const Dn<4> e0 = own ? var<4>(PA(CLV_ENERGY0, ci, cj), 2u) : cst<4>(PA(CLV_ENERGY0, ci, cj));
  const Dn<4> pr = own ? var<4>(PA(CLV_PRESSURE, ci, cj), 4u) : cst<4>(PA(CLV_PRESSURE, ci, cj));
  const Dn<4> q = own ? var<4>(PA(CLV_VISCOSITY, ci, cj), 8u) : cst<4>(PA(CLV_VISCOSITY, ci, cj));
  el_pdv<4>(predict, dt, PA(CLV_XAREA, ci, cj), PA(CLV_XAREA, ci + 1, cj), PA(CLV_YAREA, ci, cj), PA(CLV_YAREA, ci, cj + 1), u00, u01,
            u10, u11, v00, v10, v01, v11, w00, w01, w10, w11, z00, z10, z01, z11, PA(CLV_VOLUME, ci, cj), rho0, e0, pr, q, r1, e1);
}

// site (i, j) in [2, nx+2] x [2, ny+2]: node velocities (and, inside, the cell's own inputs)
__global__ void adj_pdv_k(Dev p, Dev a, int predict, double dt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i > nx + 2 || j > ny + 2) return;
  double g[4] = {0.0, 0.0, 0.0, 0.0};
  for (int cj = j - 1; cj <= j; ++cj)
    for (int ci = i - 1; ci <= i; ++ci) {
      if (!in(ci, 2, nx + 1) || !in(cj, 2, ny + 1)) continue;
      const double rb = AA(CLV_DENSITY1, ci, cj), eb = AA(CLV_ENERGY1, ci, cj);
      Dn<4> r1, e1;
      pdv_at(p, predict != 0, dt, ci, cj, 0, i, j, r1, e1);
      for (int k = 0; k < 4; ++k) g[k] += r1.d[k] * rb + e1.d[k] * eb;
    }
  AA(CLV_XVEL0, i, j) += g[0];
  AA(CLV_YVEL0, i, j) += g[1];
  if (!predict) {
    AA(CLV_XVEL1, i, j) += g[2];
    AA(CLV_YVEL1, i, j) += g[3];
  }
  if (i <= nx + 1 && j <= ny + 1) {
    const double rb = AA(CLV_DENSITY1, i, j), eb = AA(CLV_ENERGY1, i, j);
    Dn<4> r1, e1;
    pdv_at(p, predict != 0, dt, i, j, 1, 0, 0, r1, e1);
    AA(CLV_DENSITY0, i, j) += r1.d[0] * rb + e1.d[0] * eb;
    AA(CLV_ENERGY0, i, j) += r1.d[1] * rb + e1.d[1] * eb;
    AA(CLV_PRESSURE, i, j) += r1.d[2] * rb + e1.d[2] * eb;
    AA(CLV_VISCOSITY, i, j) += r1.d[3] * rb + e1.d[3] * eb;
  }
}

// ---- accelerate ---------------------------------------------------------------------------
// node (ni, nj) with the inputs of cell (ci, cj) seeded [rho0, p, q]
__device__ void accel_at(const Dev &p, double halfdt, int ni, int nj, int ci, int cj, Dn<3> &u1, Dn<3> &v1) {
  auto C3 = [&](int fid, int i, int j, unsigned bit) { return i == ci && j == cj ? var<3>(PA(fid, i, j), bit) : cst<3>(PA(fid, i, j)); };
  const int i = ni, j = nj;
  el_accel<3>(halfdt, C3(CLV_DENSITY0, i - 1, j - 1, 1u), C3(CLV_DENSITY0, i - 1, j, 1u), C3(CLV_DENSITY0, i, j, 1u),
              C3(CLV_DENSITY0, i, j - 1, 1u), PA(CLV_VOLUME, i - 1, j - 1), PA(CLV_VOLUME, i - 1, j), PA(CLV_VOLUME, i, j),
              PA(CLV_VOLUME, i, j - 1), PA(CLV_XAREA, i, j), PA(CLV_XAREA, i, j - 1), PA(CLV_YAREA, i, j), PA(CLV_YAREA, i - 1, j),
              C3(CLV_PRESSURE, i, j, 2u), C3(CLV_PRESSURE, i - 1, j, 2u), C3(CLV_PRESSURE, i, j - 1, 2u), C3(CLV_PRESSURE, i - 1, j - 1, 2u),
              C3(CLV_VISCOSITY, i, j, 4u), C3(CLV_VISCOSITY, i - 1, j, 4u), C3(CLV_VISCOSITY, i, j - 1, 4u),
              C3(CLV_VISCOSITY, i - 1, j - 1, 4u), cst<3>(PA(CLV_XVEL0, i, j)), cst<3>(PA(CLV_YVEL0, i, j)), u1, v1);
}

// site (i, j) in [1, nx+2] x [1, ny+2]: cell (i, j) inputs from nodes (i..i+1, j..j+1); node (i, j) u0, v0
__global__ void adj_accelerate_k(Dev p, Dev a, double dt) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i > nx + 2 || j > ny + 2) return;
  const double halfdt = 0.5 * dt;
  double g[3] = {0.0, 0.0, 0.0};
  for (int nj = j; nj <= j + 1; ++nj)
    for (int ni = i; ni <= i + 1; ++ni) {
      if (!in(ni, 2, nx + 2) || !in(nj, 2, ny + 2)) continue;
      const double ub = AA(CLV_XVEL1, ni, nj), vb = AA(CLV_YVEL1, ni, nj);
      Dn<3> u1, v1;
      accel_at(p, halfdt, ni, nj, i, j, u1, v1);
      for (int k = 0; k < 3; ++k) g[k] += u1.d[k] * ub + v1.d[k] * vb;
    }
  AA(CLV_DENSITY0, i, j) += g[0];
  AA(CLV_PRESSURE, i, j) += g[1];
  AA(CLV_VISCOSITY, i, j) += g[2];
  if (i >= 2 && j >= 2) {  // d u1 / d u0 = 1
    AA(CLV_XVEL0, i, j) += AA(CLV_XVEL1, i, j);
    AA(CLV_YVEL0, i, j) += AA(CLV_YVEL1, i, j);
  }
}

// ---- flux_calc (linear), src/serial/flux_calc.cpp:34-41 ------------------------------------
// site (i, j) in [2, nx+3] x [2, ny+3]: node velocities from faces (i, j-1..j) and (i-1..i, j)
__global__ void adj_flux_calc_k(Dev p, Dev a, double dt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i > nx + 3 || j > ny + 3) return;
  double gu = 0.0, gv = 0.0;
  if (i <= nx + 2) {  // x faces (i, j) and (i, j - 1) read u at (i, j)
    if (j <= ny + 2) gu += 0.25 * dt * PA(CLV_XAREA, i, j) * AA(CLV_VOL_FLUX_X, i, j);
    if (j - 1 >= 2) gu += 0.25 * dt * PA(CLV_XAREA, i, j - 1) * AA(CLV_VOL_FLUX_X, i, j - 1);
  }
  if (j <= ny + 2) {  // y faces (i, j) and (i - 1, j) read v at (i, j)
    if (i <= nx + 2) gv += 0.25 * dt * PA(CLV_YAREA, i, j) * AA(CLV_VOL_FLUX_Y, i, j);
    if (i - 1 >= 2) gv += 0.25 * dt * PA(CLV_YAREA, i - 1, j) * AA(CLV_VOL_FLUX_Y, i - 1, j);
  }
  AA(CLV_XVEL0, i, j) += gu;
  AA(CLV_XVEL1, i, j) += gu;
  AA(CLV_YVEL0, i, j) += gv;
  AA(CLV_YVEL1, i, j) += gv;
}

// ---- zero / accumulate helpers ------------------------------------------------------------
__global__ void adj_zero_k(double *f, int64_t pitch, int i0, int i1, int j0, int j1) {
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x, j = j0 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > i1 || j > j1) return;
  f[(int64_t)j * pitch + i] = 0.0;
}
// dst += src over a range, then src = 0 (revert, reset_field)
__global__ void adj_move_k(double *dst, double *src, int64_t pitch, int i0, int i1, int j0, int j1) {
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x, j = j0 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > i1 || j > j1) return;
  const int64_t k = (int64_t)j * pitch + i;
  dst[k] += src[k];
  src[k] = 0.0;
}

// ---- advec_cell ---------------------------------------------------------------------------
// Index helpers along the sweep: s = position along the sweep direction, t = the
// other; for dir x (s, t) = (i, j), for dir y (s, t) = (j, i).
struct Sweep {
  int dir;
  __device__ int64_t at(const Dev &d, int s, int t) const { return dir == 1 ? (int64_t)t * d.pitch + s : (int64_t)s * d.pitch + t; }
};
#define PS(fid, s, t) p.f[fid][sw.at(p, s, t)]
#define AS(fid, s, t) a.f[fid][sw.at(a, s, t)]

// pre_vol of advec_cell, src/serial/advec_cell.cpp:46-58 / :138-150
__device__ __forceinline__ double cell_pre_vol(const Dev &p, int sweep, int dir, int i, int j) {
  const double vol = PA(CLV_VOLUME, i, j);
  if (dir == 1) {
    if (sweep == 1)
      return vol + (PA(CLV_VOL_FLUX_X, i + 1, j) - PA(CLV_VOL_FLUX_X, i, j) + PA(CLV_VOL_FLUX_Y, i, j + 1) - PA(CLV_VOL_FLUX_Y, i, j));
    return vol + PA(CLV_VOL_FLUX_X, i + 1, j) - PA(CLV_VOL_FLUX_X, i, j);
  }
  if (sweep == 1)
    return vol + (PA(CLV_VOL_FLUX_Y, i, j + 1) - PA(CLV_VOL_FLUX_Y, i, j) + PA(CLV_VOL_FLUX_X, i + 1, j) - PA(CLV_VOL_FLUX_X, i, j));
  return vol + PA(CLV_VOL_FLUX_Y, i, j + 1) - PA(CLV_VOL_FLUX_Y, i, j);
}
__device__ __forceinline__ double pre_at(const Dev &p, int sweep, const Sweep &sw, int s, int t) {
  return sw.dir == 1 ? cell_pre_vol(p, sweep, 1, s, t) : cell_pre_vol(p, sweep, 2, t, s);
}

// face f (along the sweep) of line t: roles and the flux with the inputs of cell c seeded
// [rho (bit 1), e (bit 2), pre (bit 4, donor only), vf (bit 8, when seed_vf)]
__device__ void face_at(const Dev &p, int sweep, const Sweep &sw, int f, int t, int c, bool seed_vf, Dn<4> &mf, Dn<4> &ef) {
  const int vf_id = sw.dir == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y;
  const int smax = sw.dir == 1 ? p.nx : p.ny;
  const double vfv = PS(vf_id, f, t);
  int uw, dn, dw, dif;
  if (vfv > 0.0) {
    uw = f - 2, dn = f - 1, dw = f, dif = dn;
  } else {
    uw = min(f + 1, smax + 2), dn = f, dw = f - 1, dif = uw;
  }
  auto R = [&](int s) { return s == c ? var<4>(PS(CLV_DENSITY1, s, t), 1u) : cst<4>(PS(CLV_DENSITY1, s, t)); };
  auto E = [&](int s) { return s == c ? var<4>(PS(CLV_ENERGY1, s, t), 2u) : cst<4>(PS(CLV_ENERGY1, s, t)); };
  const double prev = pre_at(p, sweep, sw, dn, t);
  const Dn<4> pre_dn = dn == c ? var<4>(prev, 4u) : cst<4>(prev);
  const Dn<4> vf = seed_vf ? var<4>(vfv, 8u) : cst<4>(vfv);
  const int vd = sw.dir == 1 ? CLV_VERTEXDX : CLV_VERTEXDY;
  el_face<4>(vf, pre_dn, R(uw), R(dn), R(dw), E(uw), E(dn), E(dw), PM(vd, f), PM(vd, dif), mf, ef);
}

// mass and energy flux values of face f
__device__ void face_val(const Dev &p, int sweep, const Sweep &sw, int f, int t, double &mf, double &ef) {
  Dn<4> m, e;
  face_at(p, sweep, sw, f, t, -100, false, m, e);
  mf = m.v;
  ef = e.v;
}

// cell update of cell c on line t, seeded [rho, e, pre, mf_l, mf_r, ef_l, ef_r, vf_l, vf_r] selection:
// which = 0: own [rho(1), e(2), pre(4)]; which = 1: its left face [mf_l(1), ef_l(2), vf_l(4)];
// which = 2: its right face [mf_r(1), ef_r(2), vf_r(4)]
__device__ void update_at(const Dev &p, int sweep, const Sweep &sw, int c, int t, int which, Dn<3> &rn, Dn<3> &en) {
  const int vf_id = sw.dir == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y;
  double mfl, efl, mfr, efr;
  face_val(p, sweep, sw, c, t, mfl, efl);
  face_val(p, sweep, sw, c + 1, t, mfr, efr);
  const double pre = pre_at(p, sweep, sw, c, t);
  auto S = [&](double v, int w, unsigned bit) { return which == w ? var<3>(v, bit) : cst<3>(v); };
  el_cell_update<3>(S(PS(CLV_DENSITY1, c, t), 0, 1u), S(PS(CLV_ENERGY1, c, t), 0, 2u), S(pre, 0, 4u), S(mfl, 1, 1u), S(mfr, 2, 1u),
                    S(efl, 1, 2u), S(efr, 2, 2u), S(PS(vf_id, c, t), 1, 4u), S(PS(vf_id, c + 1, t), 2, 4u), rn, en);
}

// K1: faces f in [2, smax+2] of lines t in [2, tmax+1]: adjoints of mf, ef (total,
// into work1 / work2) and of vf from the cell updates (into work3); faces
// smax+3 carry only the external mass-flux adjoint.
__global__ void adj_cell_faces_k(Dev p, Dev a, int sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  const int f = 2 + blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (f > smax + 3 || t > tmax + 1) return;
  const int mf_id = dir == 1 ? CLV_MASS_FLUX_X : CLV_MASS_FLUX_Y;
  double mfb = AS(mf_id, f, t), efb = 0.0, vfb = 0.0;
  // f is the left face of cell f and the right face of cell f - 1
  if (f <= smax + 1) {
    Dn<3> rn, en;
    update_at(p, sweep, sw, f, t, 1, rn, en);
    const double rb = AS(CLV_DENSITY1, f, t), eb = AS(CLV_ENERGY1, f, t);
    mfb += rn.d[0] * rb + en.d[0] * eb;
    efb += rn.d[1] * rb + en.d[1] * eb;
    vfb += rn.d[2] * rb + en.d[2] * eb;
  }
  if (f - 1 >= 2 && f - 1 <= smax + 1) {
    Dn<3> rn, en;
    update_at(p, sweep, sw, f - 1, t, 2, rn, en);
    const double rb = AS(CLV_DENSITY1, f - 1, t), eb = AS(CLV_ENERGY1, f - 1, t);
    mfb += rn.d[0] * rb + en.d[0] * eb;
    efb += rn.d[1] * rb + en.d[1] * eb;
    vfb += rn.d[2] * rb + en.d[2] * eb;
  }
  AS(CLV_WORK1, f, t) = mfb;
  AS(CLV_WORK2, f, t) = efb;
  AS(CLV_WORK3, f, t) = vfb;
}

// K2: cells c in [0, smax+3] of lines t in [2, tmax+1]: new adjoints of rho1, e1
// (own update part for interior cells + every face that reads the cell) into
// work4 / work5, pre_vol adjoint into work6; the face's own vf term into work3.
__global__ void adj_cell_cells_k(Dev p, Dev a, int sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (c > smax + 3 || t > tmax + 1) return;
  double rb = 0.0, eb = 0.0, pb = 0.0;
  const bool interior = c >= 2 && c <= smax + 1;
  if (interior) {
    Dn<3> rn, en;
    update_at(p, sweep, sw, c, t, 0, rn, en);
    const double rnb = AS(CLV_DENSITY1, c, t), enb = AS(CLV_ENERGY1, c, t);
    rb = rn.d[0] * rnb + en.d[0] * enb;
    eb = rn.d[1] * rnb + en.d[1] * enb;
    pb = rn.d[2] * rnb + en.d[2] * enb;
  } else {
    rb = AS(CLV_DENSITY1, c, t);  // ghost cells: the kernel leaves them unchanged
    eb = AS(CLV_ENERGY1, c, t);
  }
  // faces f in [c - 1, c + 2] (within [2, smax + 3]) read cell c
  for (int f = max(2, c - 1); f <= min(smax + 3, c + 2); ++f) {
    const double mfb = AS(CLV_WORK1, f, t), efb = AS(CLV_WORK2, f, t);
    if (mfb == 0.0 && efb == 0.0) continue;
    Dn<4> mf, ef;
    face_at(p, sweep, sw, f, t, c, false, mf, ef);
    rb += mf.d[0] * mfb + ef.d[0] * efb;
    eb += mf.d[1] * mfb + ef.d[1] * efb;
    pb += mf.d[2] * mfb + ef.d[2] * efb;
  }
  AS(CLV_WORK4, c, t) = rb;
  AS(CLV_WORK5, c, t) = eb;
  AS(CLV_WORK6, c, t) = pb;
}

// K2b: faces f in [2, smax+3]: d(mf, ef)/d(vf) of the face itself, added to work3
__global__ void adj_cell_face_vf_k(Dev p, Dev a, int sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  const int f = 2 + blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (f > smax + 3 || t > tmax + 1) return;
  const double mfb = AS(CLV_WORK1, f, t), efb = AS(CLV_WORK2, f, t);
  Dn<4> mf, ef;
  face_at(p, sweep, sw, f, t, -100, true, mf, ef);
  AS(CLV_WORK3, f, t) += mf.d[3] * mfb + ef.d[3] * efb;
}

// K3: install the new rho1, e1 adjoints (cells [0, smax+3] of lines [2, tmax+1]),
// consume the mass-flux adjoint (faces [2, smax+3])
__global__ void adj_cell_install_k(Dev a, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? a.nx : a.ny, tmax = dir == 1 ? a.ny : a.nx;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (c > smax + 3 || t > tmax + 1) return;
  AS(CLV_DENSITY1, c, t) = AS(CLV_WORK4, c, t);
  AS(CLV_ENERGY1, c, t) = AS(CLV_WORK5, c, t);
  if (c >= 2) AS(dir == 1 ? CLV_MASS_FLUX_X : CLV_MASS_FLUX_Y, c, t) = 0.0;
}

// K4: pre_vol is computed on every cell [0, smax+3] x [0, tmax+3] and read where
// work6 (its adjoint, lines [2, tmax+1]) is nonzero.  vf along the sweep: faces
// [0, smax+4] of those lines: vf(f) += pb(f-1) - pb(f) + work3(f); sweep 1 also
// the other direction's flux: vf_t(s, t) += pb(s, t-1) - pb(s, t).
__global__ void adj_cell_vol_k(Dev a, int sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? a.nx : a.ny, tmax = dir == 1 ? a.ny : a.nx;
  const int s = blockIdx.x * blockDim.x + threadIdx.x, t = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (s > smax + 4 || t > tmax + 2) return;
  auto PB = [&](int ss, int tt) { return ss >= 0 && ss <= smax + 3 && tt >= 2 && tt <= tmax + 1 ? AS(CLV_WORK6, ss, tt) : 0.0; };
  const int vf_id = dir == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y, vo_id = dir == 1 ? CLV_VOL_FLUX_Y : CLV_VOL_FLUX_X;
  if (t >= 2 && t <= tmax + 1) {
    double g = PB(s - 1, t) - PB(s, t);
    if (s >= 2 && s <= smax + 3) g += AS(CLV_WORK3, s, t);
    AS(vf_id, s, t) += g;
  }
  if (sweep == 1 && s <= smax + 3) AS(vo_id, s, t) += PB(s, t - 1) - PB(s, t);
}

// ---- advec_mom (both velocities of one sweep) --------------------------------------------
// Along the sweep: node position s, line t (dir x: (s, t) = (i, j) ...).
// mom_sweep = dir + 2 (sweep - 1), src/serial/advec_mom.cpp:37
__device__ __forceinline__ double mom_post_vol(const Dev &p, int mom_sweep, int i, int j) {
  const double vol = PA(CLV_VOLUME, i, j);
  if (mom_sweep == 1) return vol + PA(CLV_VOL_FLUX_Y, i, j + 1) - PA(CLV_VOL_FLUX_Y, i, j);
  if (mom_sweep == 2) return vol + PA(CLV_VOL_FLUX_X, i + 1, j) - PA(CLV_VOL_FLUX_X, i, j);
  return vol;
}
// node flux of node (s, t): 0.25 (mf of the 4 faces around it)
__device__ __forceinline__ double node_flux_val(const Dev &p, const Sweep &sw, int s, int t) {
  if (sw.dir == 1) {
    const int i = s, j = t;
    return 0.25 * (PA(CLV_MASS_FLUX_X, i, j - 1) + PA(CLV_MASS_FLUX_X, i, j) + PA(CLV_MASS_FLUX_X, i + 1, j - 1) +
                   PA(CLV_MASS_FLUX_X, i + 1, j));
  }
  const int i = t, j = s;
  return 0.25 * (PA(CLV_MASS_FLUX_Y, i - 1, j) + PA(CLV_MASS_FLUX_Y, i, j) + PA(CLV_MASS_FLUX_Y, i - 1, j + 1) + PA(CLV_MASS_FLUX_Y, i, j + 1));
}
// node_mass_post of node (i, j) (storage indices)
__device__ __forceinline__ double node_mass_post_val(const Dev &p, int mom_sweep, int i, int j) {
  return 0.25 * (PA(CLV_DENSITY1, i, j - 1) * mom_post_vol(p, mom_sweep, i, j - 1) + PA(CLV_DENSITY1, i, j) * mom_post_vol(p, mom_sweep, i, j) +
                 PA(CLV_DENSITY1, i - 1, j - 1) * mom_post_vol(p, mom_sweep, i - 1, j - 1) +
                 PA(CLV_DENSITY1, i - 1, j) * mom_post_vol(p, mom_sweep, i - 1, j));
}
__device__ __forceinline__ double node_mass_pre_val(const Dev &p, const Sweep &sw, int mom_sweep, int s, int t) {
  const int i = sw.dir == 1 ? s : t, j = sw.dir == 1 ? t : s;
  return node_mass_post_val(p, mom_sweep, i, j) - node_flux_val(p, sw, s - 1, t) + node_flux_val(p, sw, s, t);
}

// momentum flux of node face f (between nodes f and f+1) of line t for velocity vel_id,
// seeding node c's velocity (bit 1) and node_mass_pre (bit 2), or the face's node flux (bit 4)
__device__ Dn<3> momflux_at(const Dev &p, const Sweep &sw, int mom_sweep, int vel_id, int f, int t, int c, bool seed_nf) {
  const double nf = node_flux_val(p, sw, f, t);
  int uw, dn, dw, dif;
  if (nf < 0.0) {
    uw = f + 2, dn = f + 1, dw = f, dif = dn;
  } else {
    uw = f - 1, dn = f, dw = f + 1, dif = uw;
  }
  const int cd = sw.dir == 1 ? CLV_CELLDX : CLV_CELLDY;
  auto Vv = [&](int s) { return s == c ? var<3>(PS(vel_id, s, t), 1u) : cst<3>(PS(vel_id, s, t)); };
  const double mpre = node_mass_pre_val(p, sw, mom_sweep, dn, t);
  return el_momflux<3>(seed_nf ? var<3>(nf, 4u) : cst<3>(nf), dn == c ? var<3>(mpre, 2u) : cst<3>(mpre), PM(cd, f), PM(cd, dif), Vv(uw),
                       Vv(dn), Vv(dw));
}

// M1: nodes s in [2, smax+2], lines t in [2, tmax+2]: for each velocity, the update's
// derivatives: own velocity -> work1 (u) / work2 (v) (new adjoint), node_mass_pre
// and node_mass_post adjoints -> work3 / work4; node-face momentum-flux adjoints are
// gathered in M2 from the two nodes that read each face.
__global__ void adj_mom_update_k(Dev p, Dev a, int mom_sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  const int s = 2 + blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (s > smax + 2 || t > tmax + 2) return;
  const int i = dir == 1 ? s : t, j = dir == 1 ? t : s;
  const double mpost = node_mass_post_val(p, mom_sweep, i, j);
  const double mpre = node_mass_pre_val(p, sw, mom_sweep, s, t);
  double mpreb = 0.0, mpostb = 0.0;
  for (int w = 0; w < 2; ++w) {
    const int vid = w == 0 ? CLV_XVEL1 : CLV_YVEL1;
    const double vb = AS(vid, s, t);
    const double mfl = momflux_at(p, sw, mom_sweep, vid, s - 1, t, -100, false).v, mfr = momflux_at(p, sw, mom_sweep, vid, s, t, -100, false).v;
    // seeds: vel (1), mpre (2), mpost (4)
    const Dn<3> out = el_vel_update<3>(var<3>(PS(vid, s, t), 1u), var<3>(mpre, 2u), cst<3>(mfl), cst<3>(mfr), var<3>(mpost, 4u));
    AS(w == 0 ? CLV_WORK1 : CLV_WORK2, s, t) = out.d[0] * vb;
    mpreb += out.d[1] * vb;
    mpostb += out.d[2] * vb;
  }
  AS(CLV_WORK3, s, t) = mpreb;
  AS(CLV_WORK4, s, t) = mpostb;
}

// M2: node faces f in [1, smax+2] of lines [2, tmax+2]: momentum-flux adjoint of
// each velocity from the updates of nodes f (as its right flux) and f + 1 (left).
// d(update)/d(mf_l) = 1 / mpost, d/d(mf_r) = -1 / mpost.  -> work5 (u), work6 (v)
__global__ void adj_mom_flux_k(Dev p, Dev a, int mom_sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  const int f = 1 + blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (f > smax + 2 || t > tmax + 2) return;
  for (int w = 0; w < 2; ++w) {
    const int vid = w == 0 ? CLV_XVEL1 : CLV_YVEL1;
    double g = 0.0;
    if (f >= 2) {  // right flux of node f
      const int i = dir == 1 ? f : t, j = dir == 1 ? t : f;
      g -= AS(vid, f, t) / node_mass_post_val(p, mom_sweep, i, j);
    }
    if (f + 1 <= smax + 2) {  // left flux of node f + 1
      const int i = dir == 1 ? f + 1 : t, j = dir == 1 ? t : f + 1;
      g += AS(vid, f + 1, t) / node_mass_post_val(p, mom_sweep, i, j);
    }
    AS(w == 0 ? CLV_WORK5 : CLV_WORK6, f, t) = g;
  }
}

// M3: nodes s in [0, smax+4] of lines [2, tmax+2]: through the momentum fluxes of
// faces f in [s-2, s+1] (within [1, smax+2]) the node's velocity (both) and its
// node_mass_pre; writes the velocities' new adjoints (work1/2 + flux part), adds to
// work3 (mpre adjoint); the node-flux adjoint of the face itself -> work7.
__global__ void adj_mom_nodes_k(Dev p, Dev a, int mom_sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  const int s = blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (s > smax + 4 || t > tmax + 2) return;
  const bool updated = s >= 2 && s <= smax + 2;
  double gv[2] = {0.0, 0.0}, gm = 0.0, gnf = 0.0;
  for (int w = 0; w < 2; ++w) {
    const int vid = w == 0 ? CLV_XVEL1 : CLV_YVEL1, fb_id = w == 0 ? CLV_WORK5 : CLV_WORK6;
    for (int f = max(1, s - 2); f <= min(smax + 2, s + 1); ++f) {
      const double fb = AS(fb_id, f, t);
      if (fb == 0.0) continue;
      const Dn<3> mf = momflux_at(p, sw, mom_sweep, vid, f, t, s, false);
      gv[w] += mf.d[0] * fb;
      gm += mf.d[1] * fb;
    }
    if (s >= 1 && s <= smax + 2) {  // the face s itself: its node flux
      const double fb = AS(fb_id, s, t);
      const Dn<3> mf = momflux_at(p, sw, mom_sweep, vid, s, t, -100, true);
      gnf += mf.d[2] * fb;
    }
  }
  const double u_old = updated ? AS(CLV_WORK1, s, t) : AS(CLV_XVEL1, s, t);
  const double v_old = updated ? AS(CLV_WORK2, s, t) : AS(CLV_YVEL1, s, t);
  // staged in work1/work2 (velocities) until M4 installs them
  AS(CLV_WORK1, s, t) = u_old + gv[0];
  AS(CLV_WORK2, s, t) = v_old + gv[1];
  AS(CLV_WORK3, s, t) = (updated ? AS(CLV_WORK3, s, t) : 0.0) + gm;
  AS(CLV_WORK7, s, t) = gnf;
}

// M4: install velocities (nodes [0, smax+4] of lines [2, tmax+2])
__global__ void adj_mom_install_k(Dev a, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? a.nx : a.ny, tmax = dir == 1 ? a.ny : a.nx;
  const int s = blockIdx.x * blockDim.x + threadIdx.x, t = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (s > smax + 4 || t > tmax + 2) return;
  AS(CLV_XVEL1, s, t) = AS(CLV_WORK1, s, t);
  AS(CLV_YVEL1, s, t) = AS(CLV_WORK2, s, t);
}

// M5: node masses.  node_mass_pre(s) = post(s) - nf(s-1) + nf(s) (nodes s in [1, smax+3])
// and post = 0.25 sum rho1 * post_vol over the node's 4 cells.  Per cell (i, j) in
// [0, nx+3] x [0, ny+3]: rho1 and post_vol adjoints from the nodes (i..i+1, j..j+1).
// mpost adjoint of a node = work4 + work3 (mpre's post term); nodes with a mass but
// no update have only the mpre part.  The post_vol adjoint of the cell goes to work5
// (free after M3) and M5b gathers it onto the volume fluxes.
__device__ __forceinline__ double mpost_bar(const Dev &a, const Sweep &sw, int smax, int tmax, int i, int j) {
  // node (i, j) in storage indices -> (s, t)
  const int s = sw.dir == 1 ? i : j, t = sw.dir == 1 ? j : i;
  if (t < 2 || t > tmax + 2 || s < 1 || s > smax + 3) return 0.0;
  double g = a.f[CLV_WORK3][sw.at(a, s, t)];  // mpre = mpost - ... -> mpost gets mpre's adjoint
  if (s >= 2 && s <= smax + 2) g += a.f[CLV_WORK4][sw.at(a, s, t)];
  return g;
}
__global__ void adj_mom_mass_k(Dev p, Dev a, int mom_sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > p.nx + 3 || j > p.ny + 3) return;
  double g = 0.0;
  for (int nj = j; nj <= j + 1; ++nj)
    for (int ni = i; ni <= i + 1; ++ni) g += mpost_bar(a, sw, smax, tmax, ni, nj);
  g *= 0.25;
  AA(CLV_DENSITY1, i, j) += mom_post_vol(p, mom_sweep, i, j) * g;
  AA(CLV_WORK5, i, j) = PA(CLV_DENSITY1, i, j) * g;  // post_vol adjoint of cell (i, j)
}

// M5b: post_vol = vol + vfy(j+1) - vfy(j) (mom_sweep 1) or vol + vfx(i+1) - vfx(i)
// (mom_sweep 2) on cells [0, nx+3] x [0, ny+3]: gather onto the fluxes
__global__ void adj_mom_postvol_k(Dev a, int mom_sweep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  auto PV = [&](int ci, int cj) { return ci >= 0 && ci <= a.nx + 3 && cj >= 0 && cj <= a.ny + 3 ? AA(CLV_WORK5, ci, cj) : 0.0; };
  if (mom_sweep == 1) {
    if (i > a.nx + 3 || j > a.ny + 4) return;
    AA(CLV_VOL_FLUX_Y, i, j) += PV(i, j - 1) - PV(i, j);
  } else if (mom_sweep == 2) {
    if (i > a.nx + 4 || j > a.ny + 3) return;
    AA(CLV_VOL_FLUX_X, i, j) += PV(i - 1, j) - PV(i, j);
  }
}

// M6: node fluxes.  nf(s) enters mpre(s) (+1), mpre(s+1) (-1) and the momentum flux
// of face s (work7).  nf(s, t) = 0.25 (mf of 4 faces) -> mass-flux adjoints.
__device__ __forceinline__ double nf_bar(const Dev &a, const Sweep &sw, int smax, int tmax, int s, int t) {
  if (t < 2 || t > tmax + 2 || s < 0 || s > smax + 3) return 0.0;
  double g = 0.0;
  if (s >= 1) g += a.f[CLV_WORK3][sw.at(a, s, t)];                     // mpre(s) = ... + nf(s)
  if (s + 1 <= smax + 3) g -= a.f[CLV_WORK3][sw.at(a, s + 1, t)];      // mpre(s+1) = ... - nf(s)
  if (s >= 1 && s <= smax + 2) g += a.f[CLV_WORK7][sw.at(a, s, t)];   // momentum flux of face s
  return g;
}
__global__ void adj_mom_nodeflux_k(Dev a, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? a.nx : a.ny, tmax = dir == 1 ? a.ny : a.nx;
  // dir x: mass_flux_x faces (i, j), i in [0, nx+4], j in [1, ny+2]: read by nodes (i-1, j), (i, j), (i-1, j+1), (i, j+1)
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (dir == 1) {
    if (i > a.nx + 4 || j > a.ny + 3) return;
    const double g = 0.25 * (nf_bar(a, sw, smax, tmax, i - 1, j) + nf_bar(a, sw, smax, tmax, i, j) + nf_bar(a, sw, smax, tmax, i - 1, j + 1) +
                             nf_bar(a, sw, smax, tmax, i, j + 1));
    AA(CLV_MASS_FLUX_X, i, j) += g;
  } else {
    // mass_flux_y faces (i, j) read by nodes (s = j', t = i') with nf(i', j') = 0.25 (mfy(i'-1, j') + mfy(i', j') + mfy(i'-1, j'+1) + mfy(i', j'+1))
    if (i > a.nx + 3 || j > a.ny + 4) return;
    const double g = 0.25 * (nf_bar(a, sw, smax, tmax, j, i + 1) + nf_bar(a, sw, smax, tmax, j, i) + nf_bar(a, sw, smax, tmax, j - 1, i + 1) +
                             nf_bar(a, sw, smax, tmax, j - 1, i));
    AA(CLV_MASS_FLUX_Y, i, j) += g;
  }
}

// ---- kinetic energy of field_summary, src/serial/field_summary.cpp:64-90 -------------------
// ke = sum over cells of vol * rho0 * 0.5 * sum over the 4 nodes 0.25 (u^2 + v^2)
__global__ void adj_ke_k(Dev p, Dev a, double keb) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i > nx + 2 || j > ny + 2) return;
  if (i <= nx + 1 && j <= ny + 1) {
    double vsqrd = 0.0;
    for (int kv = j; kv <= j + 1; ++kv)
      for (int jv = i; jv <= i + 1; ++jv)
        vsqrd += 0.25 * (PA(CLV_XVEL0, jv, kv) * PA(CLV_XVEL0, jv, kv) + PA(CLV_YVEL0, jv, kv) * PA(CLV_YVEL0, jv, kv));
    AA(CLV_DENSITY0, i, j) += keb * PA(CLV_VOLUME, i, j) * 0.5 * vsqrd;
  }
  double m = 0.0;  // cell masses around node (i, j)
  for (int cj = j - 1; cj <= j; ++cj)
    for (int ci = i - 1; ci <= i; ++ci)
      if (in(ci, 2, nx + 1) && in(cj, 2, ny + 1)) m += PA(CLV_VOLUME, ci, cj) * PA(CLV_DENSITY0, ci, cj);
  AA(CLV_XVEL0, i, j) += keb * m * 0.5 * 0.25 * 2.0 * PA(CLV_XVEL0, i, j);
  AA(CLV_YVEL0, i, j) += keb * m * 0.5 * 0.25 * 2.0 * PA(CLV_YVEL0, i, j);
}

}  // namespace

// ---- launchers ----------------------------------------------------------------------------
static dim3 grid2(int ni, int nj) { return dim3((ni + 63) / 64, (nj + 3) / 4, 1); }
static const dim3 kBlk(64, 4, 1);

static clv_status check(const char *what) { return hip_check(hipGetLastError(), what); }

static clv_status zero(clv_chunk *a, int fid, int i0, int i1, int j0, int j1) {
  hipLaunchKernelGGL(adj_zero_k, grid2(i1 - i0 + 1, j1 - j0 + 1), kBlk, 0, a->stream, a->d.f[fid], a->d.pitch, i0, i1, j0, j1);
  return check("adjoint zero");
}

clv_status adj_ideal_gas(clv_chunk *p, clv_chunk *a, int predict) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_ideal_gas_k, grid2(nx, ny), kBlk, 0, a->stream, p->d, a->d, predict ? CLV_DENSITY1 : CLV_DENSITY0,
                     predict ? CLV_ENERGY1 : CLV_ENERGY0);
  return check("adj ideal_gas");
}

clv_status adj_update_halo(clv_chunk *a, const int fields[CLV_NUM_HALO_FIELDS], int depth) {
  AdjHaloList list{};
  for (int f = 0; f < CLV_NUM_HALO_FIELDS; ++f)
    if (fields[f] == 1) list.ids[list.n++] = f;
  if (list.n == 0) return CLV_OK;
  const int rows = a->d.ny + 1 + 2 * depth, cols = a->d.nx + 1 + 2 * depth;
  hipLaunchKernelGGL(adj_halo_lr_k, dim3((rows + 255) / 256, list.n, 1), dim3(256, 1, 1), 0, a->stream, a->d, list, depth);
  hipLaunchKernelGGL(adj_halo_bt_k, dim3((cols + 255) / 256, list.n, 1), dim3(256, 1, 1), 0, a->stream, a->d, list, depth);
  return check("adj update_halo");
}

clv_status adj_viscosity(clv_chunk *p, clv_chunk *a) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_viscosity_k, grid2(nx + 2, ny + 2), kBlk, 0, a->stream, p->d, a->d);
  clv_status st = check("adj viscosity");
  if (st != CLV_OK) return st;
  return zero(a, CLV_VISCOSITY, 2, nx + 1, 2, ny + 1);
}

clv_status adj_PdV(clv_chunk *p, clv_chunk *a, int predict, double dt) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_pdv_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, p->d, a->d, predict, dt);
  clv_status st = check("adj PdV");
  if (st != CLV_OK) return st;
  if ((st = zero(a, CLV_DENSITY1, 2, nx + 1, 2, ny + 1)) != CLV_OK) return st;
  return zero(a, CLV_ENERGY1, 2, nx + 1, 2, ny + 1);
}

clv_status adj_revert(clv_chunk *p, clv_chunk *a) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_move_k, grid2(nx, ny), kBlk, 0, a->stream, a->d.f[CLV_DENSITY0], a->d.f[CLV_DENSITY1], a->d.pitch, 2, nx + 1, 2,
                     ny + 1);
  hipLaunchKernelGGL(adj_move_k, grid2(nx, ny), kBlk, 0, a->stream, a->d.f[CLV_ENERGY0], a->d.f[CLV_ENERGY1], a->d.pitch, 2, nx + 1, 2,
                     ny + 1);
  return check("adj revert");
}

clv_status adj_accelerate(clv_chunk *p, clv_chunk *a, double dt) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_accelerate_k, grid2(nx + 2, ny + 2), kBlk, 0, a->stream, p->d, a->d, dt);
  clv_status st = check("adj accelerate");
  if (st != CLV_OK) return st;
  if ((st = zero(a, CLV_XVEL1, 2, nx + 2, 2, ny + 2)) != CLV_OK) return st;
  return zero(a, CLV_YVEL1, 2, nx + 2, 2, ny + 2);
}

clv_status adj_flux_calc(clv_chunk *p, clv_chunk *a, double dt) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_flux_calc_k, grid2(nx + 2, ny + 2), kBlk, 0, a->stream, p->d, a->d, dt);
  clv_status st = check("adj flux_calc");
  if (st != CLV_OK) return st;
  if ((st = zero(a, CLV_VOL_FLUX_X, 2, nx + 2, 2, ny + 2)) != CLV_OK) return st;
  return zero(a, CLV_VOL_FLUX_Y, 2, nx + 2, 2, ny + 2);
}

clv_status adj_advec_cell(clv_chunk *p, clv_chunk *a, int sweep, int dir) {
  const int smax = dir == 1 ? p->d.nx : p->d.ny, tmax = dir == 1 ? p->d.ny : p->d.nx;
  // work arrays of the adjoint chunk are scratch: clear the ones read before written
  const dim3 gf = grid2(smax + 2, tmax), gc = grid2(smax + 4, tmax), gv = grid2(smax + 5, tmax + 2);
  hipLaunchKernelGGL(adj_cell_faces_k, gf, kBlk, 0, a->stream, p->d, a->d, sweep, dir);
  hipLaunchKernelGGL(adj_cell_cells_k, gc, kBlk, 0, a->stream, p->d, a->d, sweep, dir);
  hipLaunchKernelGGL(adj_cell_face_vf_k, gf, kBlk, 0, a->stream, p->d, a->d, sweep, dir);
  hipLaunchKernelGGL(adj_cell_install_k, gc, kBlk, 0, a->stream, a->d, dir);
  hipLaunchKernelGGL(adj_cell_vol_k, gv, kBlk, 0, a->stream, a->d, sweep, dir);
  return check("adj advec_cell");
}

clv_status adj_advec_mom(clv_chunk *p, clv_chunk *a, int sweep, int dir) {
  const int smax = dir == 1 ? p->d.nx : p->d.ny, tmax = dir == 1 ? p->d.ny : p->d.nx;
  const int mom_sweep = dir + 2 * (sweep - 1);
  hipLaunchKernelGGL(adj_mom_update_k, grid2(smax + 1, tmax + 1), kBlk, 0, a->stream, p->d, a->d, mom_sweep, dir);
  hipLaunchKernelGGL(adj_mom_flux_k, grid2(smax + 2, tmax + 1), kBlk, 0, a->stream, p->d, a->d, mom_sweep, dir);
  hipLaunchKernelGGL(adj_mom_nodes_k, grid2(smax + 5, tmax + 1), kBlk, 0, a->stream, p->d, a->d, mom_sweep, dir);
  hipLaunchKernelGGL(adj_mom_install_k, grid2(smax + 5, tmax + 1), kBlk, 0, a->stream, a->d, dir);
  hipLaunchKernelGGL(adj_mom_mass_k, grid2(p->d.nx + 4, p->d.ny + 4), kBlk, 0, a->stream, p->d, a->d, mom_sweep, dir);
  if (mom_sweep <= 2) hipLaunchKernelGGL(adj_mom_postvol_k, grid2(p->d.nx + 5, p->d.ny + 5), kBlk, 0, a->stream, a->d, mom_sweep);
  hipLaunchKernelGGL(adj_mom_nodeflux_k, grid2(p->d.nx + 5, p->d.ny + 5), kBlk, 0, a->stream, a->d, dir);
  return check("adj advec_mom");
}

clv_status adj_reset_field(clv_chunk *p, clv_chunk *a) {
  const int nx = p->d.nx, ny = p->d.ny;
  const int64_t P = a->d.pitch;
  hipLaunchKernelGGL(adj_move_k, grid2(nx, ny), kBlk, 0, a->stream, a->d.f[CLV_DENSITY1], a->d.f[CLV_DENSITY0], P, 2, nx + 1, 2, ny + 1);
  hipLaunchKernelGGL(adj_move_k, grid2(nx, ny), kBlk, 0, a->stream, a->d.f[CLV_ENERGY1], a->d.f[CLV_ENERGY0], P, 2, nx + 1, 2, ny + 1);
  hipLaunchKernelGGL(adj_move_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, a->d.f[CLV_XVEL1], a->d.f[CLV_XVEL0], P, 2, nx + 2, 2, ny + 2);
  hipLaunchKernelGGL(adj_move_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, a->d.f[CLV_YVEL1], a->d.f[CLV_YVEL0], P, 2, nx + 2, 2, ny + 2);
  return check("adj reset_field");
}

clv_status adj_kinetic_energy(clv_chunk *p, clv_chunk *a, double keb) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_ke_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, p->d, a->d, keb);
  return check("adj kinetic energy");
}

}  // namespace clv
