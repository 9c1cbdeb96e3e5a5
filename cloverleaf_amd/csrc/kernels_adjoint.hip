// Reverse-mode (adjoint) kernels of the fixed-dt hydro step, one group per
// reference kernel, for the gradient of a scalar of the final state with respect
// to the initial fields.
//
// Conventions.  p is the primal chunk holding the forward kernel's INPUT state;
// a is a second chunk of the same size holding adjoints.  On entry a holds the
// adjoints of the kernel's outputs (and whatever has accumulated on its
// inputs); on exit the outputs' adjoints are consumed (zeroed where the kernel
// overwrote them) and the inputs' adjoints are accumulated.  Geometry (volume,
// areas, mesh) is constant; soundspeed, which only calc_dt reads, carries no
// adjoint.  The work arrays of BOTH chunks are scratch: the primal's hold the
// forward's intermediate values (face fluxes, node masses), the adjoint's the
// per-element shares below.
//
// Structure.  Every element (cell, node or face) the forward kernel computes is
// differentiated once.  Its nonlinear core is evaluated on N-component dual
// numbers seeded at the core's inputs, on the same values in the same order as
// the forward, so every branch (upwind choice, van Leer limiter, viscosity
// switch, fmin/fmax, fabs) is differentiated on the piece the forward took; the
// element's output adjoints times those partials are stored per element and
// role.  A gather kernel then adds, for every input, the shares of the elements
// that read it, in a fixed order -- deterministic, no atomics.  Linear stages
// (sums of velocities, pressure differences, the total volume flux, node
// masses) are transposed by hand.  Ties follow the oracle's tangent-linear
// model: fmin/fmax return their first argument, fabs(0) has derivative +1.
#include "clv_internal.h"

namespace clv {
namespace {

// ---- N-component dual numbers (N = 0: values only) ---------------------------------------
template <int N>
struct Dn {
  double v;
  double d[N > 0 ? N : 1];
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
// the value is the correctly rounded quotient; the derivative uses one reciprocal
template <int N>
__device__ __forceinline__ Dn<N> operator/(Dn<N> a, Dn<N> b) {
  Dn<N> r;
  r.v = a.v / b.v;
  if constexpr (N > 0) {
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  }
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
__device__ __forceinline__ Dn<N> operator/(Dn<N> a, double b) {
  Dn<N> r;
  r.v = a.v / b;
  if constexpr (N > 0) {
    const double ib = 1.0 / b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * ib;
  }
  return r;
}
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
  if constexpr (N > 0) {
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
  }
  return r;
}

#define PA(fid, i, j) p.f[fid][(int64_t)(j) * p.pitch + (i)]
#define AA(fid, i, j) a.f[fid][(int64_t)(j) * a.pitch + (i)]
#define PM(mid, i) p.m[mid][i]
// adjoint and primal work values are stored once and read next by another launch:
// streaming (non-temporal) stores, 1.6 % off the backward pass at 1920^2
#define ST_STREAM(lhs, v) __builtin_nontemporal_store((v), &(lhs))

__device__ __forceinline__ bool in(int i, int lo, int hi) { return i >= lo && i <= hi; }

// ---- element cores: the reference expressions (oracle order) on dual numbers --------------
// viscosity of one cell, src/serial/viscosity.cpp:43-66, from its linear pre-stages:
// ugrad = (u10 + u11) - (u00 + u01), vgrad = (v01 + v11) - (v00 + v10),
// su = u01 + u11 - u00 - u10, sv = v10 + v11 - v00 - v01, dpx = p(i+1) - p(i-1),
// dpy = p(j+1) - p(j-1)
template <int N>
__device__ Dn<N> el_visc_core(Dn<N> ugrad, Dn<N> vgrad, Dn<N> su, Dn<N> sv, Dn<N> dpx, Dn<N> dpy, Dn<N> rho, double dxi, double dxi1,
                              double dyj, double dyj2) {
  Dn<N> div = (dxi * (ugrad) + dyj * (vgrad));
  Dn<N> strain2 = 0.5 * su / dyj + 0.5 * sv / dxi;
  Dn<N> pgradx = dpx / (dxi + dxi1);
  Dn<N> pgrady = dpy / (dyj + dyj2);
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

// PdV of one cell after its total volume flux, src/serial/PdV.cpp:61-63 / :81-83
template <int N>
__device__ void el_pdv_core(double vol, Dn<N> rho0, Dn<N> e0, Dn<N> pr, Dn<N> q, Dn<N> total_flux, Dn<N> &rho1, Dn<N> &e1) {
  Dn<N> volume_change_s = vol / (vol + total_flux);
  double recip_volume = 1.0 / vol;
  Dn<N> energy_change = (pr / rho0 + q / rho0) * total_flux * recip_volume;
  e1 = e0 - energy_change;
  rho1 = rho0 * volume_change_s;
}

// total volume flux of cell (i, j), src/serial/PdV.cpp:49-60 (predict) / :72-80 (correct)
__device__ __forceinline__ double pdv_total_flux(const Dev &p, bool predict, double dt, int i, int j) {
  const double xa0 = PA(CLV_XAREA, i, j), xa1 = PA(CLV_XAREA, i + 1, j), ya0 = PA(CLV_YAREA, i, j), ya1 = PA(CLV_YAREA, i, j + 1);
  const double u00 = PA(CLV_XVEL0, i, j), u01 = PA(CLV_XVEL0, i, j + 1), u10 = PA(CLV_XVEL0, i + 1, j), u11 = PA(CLV_XVEL0, i + 1, j + 1);
  const double v00 = PA(CLV_YVEL0, i, j), v10 = PA(CLV_YVEL0, i + 1, j), v01 = PA(CLV_YVEL0, i, j + 1), v11 = PA(CLV_YVEL0, i + 1, j + 1);
  double left_flux, right_flux, bottom_flux, top_flux;
  if (predict) {
    left_flux = (xa0 * (u00 + u01 + u00 + u01)) * 0.25 * dt * 0.5;
    right_flux = (xa1 * (u10 + u11 + u10 + u11)) * 0.25 * dt * 0.5;
    bottom_flux = (ya0 * (v00 + v10 + v00 + v10)) * 0.25 * dt * 0.5;
    top_flux = (ya1 * (v01 + v11 + v01 + v11)) * 0.25 * dt * 0.5;
  } else {
    const double w00 = PA(CLV_XVEL1, i, j), w01 = PA(CLV_XVEL1, i, j + 1), w10 = PA(CLV_XVEL1, i + 1, j), w11 = PA(CLV_XVEL1, i + 1, j + 1);
    const double z00 = PA(CLV_YVEL1, i, j), z10 = PA(CLV_YVEL1, i + 1, j), z01 = PA(CLV_YVEL1, i, j + 1), z11 = PA(CLV_YVEL1, i + 1, j + 1);
    left_flux = (xa0 * (u00 + u01 + w00 + w01)) * 0.25 * dt;
    right_flux = (xa1 * (u10 + u11 + w10 + w11)) * 0.25 * dt;
    bottom_flux = (ya0 * (v00 + v10 + z00 + z10)) * 0.25 * dt;
    top_flux = (ya1 * (v01 + v11 + z01 + z11)) * 0.25 * dt;
  }
  return right_flux - left_flux + top_flux - bottom_flux;
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

// ---- reverse sweeps of the limited fluxes ------------------------------------------------
// The same values in the same order as the forward; the branch each fmin / fabs /
// limiter switch took is replayed and differentiated (fmin(a, b) is b only when
// b < a; d|x|/dx = -1 for x < 0, else +1).
__device__ __forceinline__ double sgn_of(double x) { return x < 0.0 ? -1.0 : 1.0; }

struct VlBar {
  double sigma, s3, s4, uw, dn, dw;
};
// adjoint of el_vl for the output adjoint lb
__device__ __forceinline__ VlBar vl_bar(double sigma, double sigma3, double sigma4, double uw, double dn, double dw, double lb) {
  const double one_by_six = 1.0 / 6.0;
  VlBar r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double diffuw = dn - uw, diffdw = dw - dn;
  const double wind = diffdw <= 0.0 ? -1.0 : 1.0;
  if (!(diffuw * diffdw > 0.0)) return r;
  const double au = ::fabs(diffuw), ad = ::fabs(diffdw);
  const double c = one_by_six * (sigma3 * au + sigma4 * ad);
  const double m1 = ad < au ? ad : au;
  const double m = c < m1 ? c : m1;
  r.sigma = -wind * m * lb;
  const double mb = (1.0 - sigma) * wind * lb;
  double aub = 0.0, adb = 0.0;
  if (c < m1) {
    r.s3 = mb * one_by_six * au;
    r.s4 = mb * one_by_six * ad;
    aub = mb * one_by_six * sigma3;
    adb = mb * one_by_six * sigma4;
  } else if (ad < au) {
    adb = mb;
  } else {
    aub = mb;
  }
  const double dub = aub * sgn_of(diffuw), ddb = adb * sgn_of(diffdw);
  r.uw = -dub;
  r.dn = dub - ddb;
  r.dw = ddb;
  return r;
}

// adjoint of el_face: g = d(mf, ef)/d(vf, pre_dn, r_uw, r_dn, r_dw, e_uw, e_dn, e_dw)^T (mfb, efb)
__device__ __forceinline__ void face_bar(double vf, double pre, double r_uw, double r_dn, double r_dw, double e_uw, double e_dn, double e_dw, double vd_face,
                         double vd_dif, double mfb, double efb, double g[8]) {
  const double one_by_six = 1.0 / 6.0;
  // forward values
  const double sigmat = ::fabs(vf) / pre;
  const double ratio = vd_face / vd_dif;
  const double sigma3 = (1.0 + sigmat) * ratio;
  const double sigma4 = 2.0 - sigmat;
  double lim1 = 0.0;
  {
    const double diffuw = r_dn - r_uw, diffdw = r_dw - r_dn;
    const double wind = diffdw <= 0.0 ? -1.0 : 1.0;
    if (diffuw * diffdw > 0.0) {
      const double au = ::fabs(diffuw), ad = ::fabs(diffdw);
      const double m1 = ad < au ? ad : au, c = one_by_six * (sigma3 * au + sigma4 * ad);
      lim1 = (1.0 - sigmat) * wind * (c < m1 ? c : m1);
    }
  }
  const double mf = vf * (r_dn + lim1);
  const double den = r_dn * pre;
  const double sigmam = ::fabs(mf) / den;
  double lim2 = 0.0;
  {
    const double diffuw = e_dn - e_uw, diffdw = e_dw - e_dn;
    const double wind = diffdw <= 0.0 ? -1.0 : 1.0;
    if (diffuw * diffdw > 0.0) {
      const double au = ::fabs(diffuw), ad = ::fabs(diffdw);
      const double m1 = ad < au ? ad : au, c = one_by_six * (sigma3 * au + sigma4 * ad);
      lim2 = (1.0 - sigmam) * wind * (c < m1 ? c : m1);
    }
  }
  // reverse: ef = mf (e_dn + lim2)
  double mf_b = mfb + efb * (e_dn + lim2);
  const double lim2_b = efb * mf;
  const VlBar b2 = vl_bar(sigmam, sigma3, sigma4, e_uw, e_dn, e_dw, lim2_b);
  g[5] = b2.uw;
  g[6] = efb * mf + b2.dn;
  g[7] = b2.dw;
  double s3b = b2.s3, s4b = b2.s4;
  // sigmam = |mf| / (r_dn pre)
  mf_b += b2.sigma * sgn_of(mf) / den;
  const double den_b = -b2.sigma * sigmam / den;
  double rdn_b = den_b * pre, pre_b = den_b * r_dn;
  // mf = vf (r_dn + lim1)
  double vf_b = mf_b * (r_dn + lim1);
  rdn_b += mf_b * vf;
  const VlBar b1 = vl_bar(sigmat, sigma3, sigma4, r_uw, r_dn, r_dw, mf_b * vf);
  s3b += b1.s3;
  s4b += b1.s4;
  // sigmat feeds the first limiter's sigma, sigma3 and sigma4
  const double st_b = b1.sigma + s3b * ratio - s4b;
  vf_b += st_b * sgn_of(vf) / pre;
  pre_b -= st_b * sigmat / pre;
  g[0] = vf_b;
  g[1] = pre_b;
  g[2] = b1.uw;
  g[3] = rdn_b + b1.dn;
  g[4] = b1.dw;
}

// adjoint of el_momflux: g = d(flux)/d(node_flux, mass_pre_donor, v_uw, v_dn, v_dw) * gb;
// returns the momentum flux itself (el_momflux<0>'s value, the same operations)
__device__ double momflux_bar(double flux, double mpre, double width, double width_dif, double v_uw, double v_dn, double v_dw, double gb,
                              double g[5]) {
  const double sigma = ::fabs(flux) / mpre;
  const double vdiffuw = v_dn - v_uw, vdiffdw = v_dw - v_dn;
  double limiter = 0.0, wind = 1.0, auw = 0.0, adw = 0.0, x = 0.0, m1 = 0.0;
  const bool lim = vdiffuw * vdiffdw > 0.0;
  if (lim) {
    auw = ::fabs(vdiffuw);
    adw = ::fabs(vdiffdw);
    if (vdiffdw <= 0.0) wind = -1.0;
    x = width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / width_dif) / 6.0;
    m1 = auw < x ? auw : x;
    limiter = wind * (adw < m1 ? adw : m1);
  }
  const double advec_vel_s = v_dn + (1.0 - sigma) * limiter;
  // reverse: flux_out = advec_vel_s * flux
  const double av_b = gb * flux;
  double flux_b = gb * advec_vel_s;
  double vdn_b = av_b, vuw_b = 0.0, vdw_b = 0.0;
  double sigma_b = -av_b * limiter;
  if (lim) {
    const double mb = wind * av_b * (1.0 - sigma);
    double auw_b = 0.0, adw_b = 0.0;
    if (adw < m1) {
      adw_b = mb;
    } else if (auw < x) {
      auw_b = mb;
    } else {
      const double yb = mb * width / 6.0;
      sigma_b += yb * (-adw / width + auw / width_dif);
      adw_b = yb * (2.0 - sigma) / width;
      auw_b = yb * (1.0 + sigma) / width_dif;
    }
    const double duw_b = auw_b * sgn_of(vdiffuw), ddw_b = adw_b * sgn_of(vdiffdw);
    vuw_b -= duw_b;
    vdn_b += duw_b - ddw_b;
    vdw_b += ddw_b;
  }
  flux_b += sigma_b * sgn_of(flux) / mpre;
  g[0] = flux_b;
  g[1] = -sigma_b * sigma / mpre;
  g[2] = vuw_b;
  g[3] = vdn_b;
  g[4] = vdw_b;
  return advec_vel_s * flux;
}

// ---- ideal_gas, src/serial/ideal_gas.cpp:38-46: p = 0.4 rho e on the interior ---------------
__global__ void adj_ideal_gas_k(Dev p, Dev a, int rho_f, int e_f) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > p.nx + 1 || j > p.ny + 1) return;
  const double pb = AA(CLV_PRESSURE, i, j);
  AA(rho_f, i, j) += (1.4 - 1.0) * PA(e_f, i, j) * pb;
  AA(e_f, i, j) += (1.4 - 1.0) * PA(rho_f, i, j) * pb;
  ST_STREAM(AA(CLV_PRESSURE, i, j), 0.0);
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

// ---- viscosity ----------------------------------------------------------------------------
// V1, cells [2, nx+1]^2: the core's input adjoints, q_bar * dq/d(ugrad, vgrad, su,
// sv, dpx, dpy) -> a.work1..6; the density share directly; consumes q_bar.
__global__ void adj_visc_cell_k(Dev p, Dev a) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > p.nx + 1 || j > p.ny + 1) return;
  const double qb = AA(CLV_VISCOSITY, i, j);
  ST_STREAM(AA(CLV_VISCOSITY, i, j), 0.0);
  double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (qb != 0.0) {
    const double u00 = PA(CLV_XVEL0, i, j), u10 = PA(CLV_XVEL0, i + 1, j), u01 = PA(CLV_XVEL0, i, j + 1), u11 = PA(CLV_XVEL0, i + 1, j + 1);
    const double v00 = PA(CLV_YVEL0, i, j), v10 = PA(CLV_YVEL0, i + 1, j), v01 = PA(CLV_YVEL0, i, j + 1), v11 = PA(CLV_YVEL0, i + 1, j + 1);
    const double ugrad = (u10 + u11) - (u00 + u01), vgrad = (v01 + v11) - (v00 + v10);
    const double su = u01 + u11 - u00 - u10, sv = v10 + v11 - v00 - v01;
    const double dpx = PA(CLV_PRESSURE, i + 1, j) - PA(CLV_PRESSURE, i - 1, j);
    const double dpy = PA(CLV_PRESSURE, i, j + 1) - PA(CLV_PRESSURE, i, j - 1);
    const Dn<7> q = el_visc_core<7>(var<7>(ugrad, 1u), var<7>(vgrad, 2u), var<7>(su, 4u), var<7>(sv, 8u), var<7>(dpx, 16u), var<7>(dpy, 32u),
                                    var<7>(PA(CLV_DENSITY0, i, j), 64u), PM(CLV_CELLDX, i), PM(CLV_CELLDX, i + 1), PM(CLV_CELLDY, j),
                                    PM(CLV_CELLDY, j + 2));
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k] = q.d[k] * qb;
    AA(CLV_DENSITY0, i, j) += q.d[6] * qb;
  }
  ST_STREAM(AA(CLV_WORK1, i, j), g[0]);
  ST_STREAM(AA(CLV_WORK2, i, j), g[1]);
  ST_STREAM(AA(CLV_WORK3, i, j), g[2]);
  ST_STREAM(AA(CLV_WORK4, i, j), g[3]);
  ST_STREAM(AA(CLV_WORK5, i, j), g[4]);
  ST_STREAM(AA(CLV_WORK6, i, j), g[5]);
}

// V2, sites (i, j) in [1, nx+2]^2: node (i, j) velocities from its four cells,
// cell (i, j) pressure from the four cells whose pressure gradient reads it
__global__ void adj_visc_gather_k(Dev a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = a.nx, ny = a.ny;
  if (i > nx + 2 || j > ny + 2) return;
  auto W = [&](int fid, int ci, int cj) { return in(ci, 2, nx + 1) && in(cj, 2, ny + 1) ? AA(fid, ci, cj) : 0.0; };
  if (i >= 2 && j >= 2) {
    // node (i, j) is u00/v00 of cell (i, j), u10/v10 of (i-1, j), u01/v01 of (i, j-1), u11/v11 of (i-1, j-1)
    const double ug = -W(CLV_WORK1, i, j) + W(CLV_WORK1, i - 1, j) - W(CLV_WORK1, i, j - 1) + W(CLV_WORK1, i - 1, j - 1);
    const double us = -W(CLV_WORK3, i, j) - W(CLV_WORK3, i - 1, j) + W(CLV_WORK3, i, j - 1) + W(CLV_WORK3, i - 1, j - 1);
    const double vg = -W(CLV_WORK2, i, j) - W(CLV_WORK2, i - 1, j) + W(CLV_WORK2, i, j - 1) + W(CLV_WORK2, i - 1, j - 1);
    const double vs = -W(CLV_WORK4, i, j) + W(CLV_WORK4, i - 1, j) - W(CLV_WORK4, i, j - 1) + W(CLV_WORK4, i - 1, j - 1);
    AA(CLV_XVEL0, i, j) += ug + us;
    AA(CLV_YVEL0, i, j) += vg + vs;
  }
  // dpx(c) = p(c + x) - p(c - x), dpy(c) = p(c + y) - p(c - y)
  AA(CLV_PRESSURE, i, j) += W(CLV_WORK5, i - 1, j) - W(CLV_WORK5, i + 1, j) + W(CLV_WORK6, i, j - 1) - W(CLV_WORK6, i, j + 1);
}

// ---- PdV ----------------------------------------------------------------------------------
// P1, cells [2, nx+1]^2: own inputs (rho0, e0, p, q) and the total-flux adjoint
// (-> a.work1); consumes the rho1, e1 adjoints
__global__ void adj_pdv_cell_k(Dev p, Dev a, int predict, double dt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > p.nx + 1 || j > p.ny + 1) return;
  const double rb = AA(CLV_DENSITY1, i, j), eb = AA(CLV_ENERGY1, i, j);
  const double tf = pdv_total_flux(p, predict != 0, dt, i, j);
  Dn<5> r1, e1;
  el_pdv_core<5>(PA(CLV_VOLUME, i, j), var<5>(PA(CLV_DENSITY0, i, j), 1u), var<5>(PA(CLV_ENERGY0, i, j), 2u), var<5>(PA(CLV_PRESSURE, i, j), 4u),
                 var<5>(PA(CLV_VISCOSITY, i, j), 8u), var<5>(tf, 16u), r1, e1);
  AA(CLV_DENSITY0, i, j) += r1.d[0] * rb + e1.d[0] * eb;
  AA(CLV_ENERGY0, i, j) += r1.d[1] * rb + e1.d[1] * eb;
  AA(CLV_PRESSURE, i, j) += r1.d[2] * rb + e1.d[2] * eb;
  AA(CLV_VISCOSITY, i, j) += r1.d[3] * rb + e1.d[3] * eb;
  ST_STREAM(AA(CLV_WORK1, i, j), r1.d[4] * rb + e1.d[4] * eb);
  ST_STREAM(AA(CLV_DENSITY1, i, j), 0.0);
  ST_STREAM(AA(CLV_ENERGY1, i, j), 0.0);
}

// P2, nodes [2, nx+2]^2.  d(face flux)/d(u of a face node) = area * 0.25 dt in
// both forms (the predictor counts u0 twice and halves).  Node (i, j) is u00 of
// cell (i, j) (left face), u01 of (i, j-1) (left), u10 of (i-1, j) (right), u11 of
// (i-1, j-1) (right); v00 of (i, j) (bottom), v10 of (i-1, j) (bottom), v01 of
// (i, j-1) (top), v11 of (i-1, j-1) (top).
__global__ void adj_pdv_node_k(Dev p, Dev a, int predict, double dt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i > nx + 2 || j > ny + 2) return;
  auto T = [&](int ci, int cj) { return in(ci, 2, nx + 1) && in(cj, 2, ny + 1) ? AA(CLV_WORK1, ci, cj) : 0.0; };
  const double k = 0.25 * dt;
  const double gu = k * (PA(CLV_XAREA, i, j) * (T(i - 1, j) - T(i, j)) + PA(CLV_XAREA, i, j - 1) * (T(i - 1, j - 1) - T(i, j - 1)));
  const double gv = k * (PA(CLV_YAREA, i, j) * (T(i, j - 1) - T(i, j)) + PA(CLV_YAREA, i - 1, j) * (T(i - 1, j - 1) - T(i - 1, j)));
  AA(CLV_XVEL0, i, j) += gu;
  AA(CLV_YVEL0, i, j) += gv;
  if (!predict) {
    AA(CLV_XVEL1, i, j) += gu;
    AA(CLV_YVEL1, i, j) += gv;
  }
}

// ---- accelerate, src/serial/accelerate.cpp:43-53 ------------------------------------------
// A1, nodes [2, nx+2]^2: u1 = u0 - s (Pu + Qu), v1 = v0 - s (Pv + Qv) with
// s = halfdt / (0.25 M), M = sum rho vol over the node's cells.  u0 += u1_bar;
// a = -s u1_bar, b = -s v1_bar (pressure/viscosity shares) -> a.work1/2; the
// mass share R = s_bar ds/dM -> a.work3; consumes u1_bar, v1_bar.
__global__ void adj_accel_node_k(Dev p, Dev a, double dt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > p.nx + 2 || j > p.ny + 2) return;
  const double halfdt = 0.5 * dt;
  const double ub = AA(CLV_XVEL1, i, j), vb = AA(CLV_YVEL1, i, j);
  AA(CLV_XVEL0, i, j) += ub;
  AA(CLV_YVEL0, i, j) += vb;
  ST_STREAM(AA(CLV_XVEL1, i, j), 0.0);
  ST_STREAM(AA(CLV_YVEL1, i, j), 0.0);
  const double m = (PA(CLV_DENSITY0, i - 1, j - 1) * PA(CLV_VOLUME, i - 1, j - 1) + PA(CLV_DENSITY0, i - 1, j) * PA(CLV_VOLUME, i - 1, j) +
                    PA(CLV_DENSITY0, i, j) * PA(CLV_VOLUME, i, j) + PA(CLV_DENSITY0, i, j - 1) * PA(CLV_VOLUME, i, j - 1)) *
                   0.25;
  const double s = halfdt / m;
  auto D = [&](int fid) {  // x and y differences of a cell field around the node
    const double f00 = PA(fid, i, j), fm0 = PA(fid, i - 1, j), f0m = PA(fid, i, j - 1), fmm = PA(fid, i - 1, j - 1);
    return double2{PA(CLV_XAREA, i, j) * (f00 - fm0) + PA(CLV_XAREA, i, j - 1) * (f0m - fmm),
                   PA(CLV_YAREA, i, j) * (f00 - f0m) + PA(CLV_YAREA, i - 1, j) * (fm0 - fmm)};
  };
  const double2 dp = D(CLV_PRESSURE), dq = D(CLV_VISCOSITY);
  const double sb = -(ub * (dp.x + dq.x) + vb * (dp.y + dq.y));
  ST_STREAM(AA(CLV_WORK1, i, j), -s * ub);
  ST_STREAM(AA(CLV_WORK2, i, j), -s * vb);
  ST_STREAM(AA(CLV_WORK3, i, j), sb * (-s / m) * 0.25);  // ds/d(rho_c) = -s / m * 0.25 vol_c
}

// A2, cells [1, nx+2]^2: cell c is 00 of node c, m0 of node c+x, 0m of node c+y,
// mm of node c+x+y
__global__ void adj_accel_cell_k(Dev p, Dev a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i > nx + 2 || j > ny + 2) return;
  auto W = [&](int fid, int ni, int nj) { return in(ni, 2, nx + 2) && in(nj, 2, ny + 2) ? AA(fid, ni, nj) : 0.0; };
  const double r = W(CLV_WORK3, i, j) + W(CLV_WORK3, i + 1, j) + W(CLV_WORK3, i, j + 1) + W(CLV_WORK3, i + 1, j + 1);
  AA(CLV_DENSITY0, i, j) += PA(CLV_VOLUME, i, j) * r;
  const double xa0 = PA(CLV_XAREA, i, j), xa1 = PA(CLV_XAREA, i + 1, j), ya0 = PA(CLV_YAREA, i, j), ya1 = PA(CLV_YAREA, i, j + 1);
  const double g = W(CLV_WORK1, i, j) * xa0 + W(CLV_WORK2, i, j) * ya0 - W(CLV_WORK1, i + 1, j) * xa1 + W(CLV_WORK2, i + 1, j) * ya0 +
                   W(CLV_WORK1, i, j + 1) * xa0 - W(CLV_WORK2, i, j + 1) * ya1 - W(CLV_WORK1, i + 1, j + 1) * xa1 -
                   W(CLV_WORK2, i + 1, j + 1) * ya1;
  AA(CLV_PRESSURE, i, j) += g;
  AA(CLV_VISCOSITY, i, j) += g;
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
__global__ void adj_copy_k(double *dst, const double *src, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

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
// (s, t) of this thread from offsets (s0, t0): lanes run along x -- s for dir x, t for
// dir y -- so the sweeps of both directions read whole rows; the grid is sgrid()'s
__device__ __forceinline__ void sweep_pos(int dir, int s0, int t0, int &s, int &t) {
  const int gx = blockIdx.x * blockDim.x + threadIdx.x, gy = blockIdx.y * blockDim.y + threadIdx.y;
  s = s0 + (dir == 1 ? gx : gy);
  t = t0 + (dir == 1 ? gy : gx);
}
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

// the cells face f reads: upwind, donor, downwind (src/serial/advec_cell.cpp:74-83, the
// upwind index clamped at smax+2), and the mesh index of the limiter's width ratio
struct FaceRoles {
  int uw, dn, dw, dif;
};
__device__ __forceinline__ FaceRoles face_roles(double vf, int f, int smax) {
  if (vf > 0.0) return FaceRoles{f - 2, f - 1, f, f - 1};
  const int uw = min(f + 1, smax + 2);
  return FaceRoles{uw, f, f - 1, uw};
}

// C0 (primal), positions s in [0, smax+3] of lines [2, tmax+1]: pre_vol(s) -> p.work1;
// for faces s >= 2 the forward's mass and energy flux -> p.work2, p.work3
__global__ void adj_cell_values_k(Dev p, int sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  int s, t;
  sweep_pos(dir, 0, 2, s, t);
  if (s > smax + 3 || t > tmax + 1) return;
  ST_STREAM(PS(CLV_WORK1, s, t), pre_at(p, sweep, sw, s, t));
  if (s < 2) return;
  const int vf_id = dir == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y, vd = dir == 1 ? CLV_VERTEXDX : CLV_VERTEXDY;
  const double vfv = PS(vf_id, s, t);
  const FaceRoles r = face_roles(vfv, s, smax);
  Dn<0> mf, ef;
  el_face<0>(cst<0>(vfv), cst<0>(pre_at(p, sweep, sw, r.dn, t)), cst<0>(PS(CLV_DENSITY1, r.uw, t)), cst<0>(PS(CLV_DENSITY1, r.dn, t)),
             cst<0>(PS(CLV_DENSITY1, r.dw, t)), cst<0>(PS(CLV_ENERGY1, r.uw, t)), cst<0>(PS(CLV_ENERGY1, r.dn, t)),
             cst<0>(PS(CLV_ENERGY1, r.dw, t)), PM(vd, s), PM(vd, r.dif), mf, ef);
  ST_STREAM(PS(CLV_WORK2, s, t), mf.v);
  ST_STREAM(PS(CLV_WORK3, s, t), ef.v);
}

// One cell's update, src/serial/advec_cell.cpp:120-125 / :211-216:
// rho' = (rho pre + mf_l - mf_r) / (pre + vf_l - vf_r),
// e' = (e rho pre + ef_l - ef_r) / (rho pre + mf_l - mf_r),
// differentiated for the incoming adjoints rnb, enb of rho', e'.  Own shares of rho, e
// and pre (rb, eb, pb; the adjoints then accumulate the faces' shares), and the terms its
// faces take: d/d mf_l = M = rnb / advec_vol - enb e' / post_mass, d/d ef_l = E =
// enb / post_mass, d/d vf_l = -V = -rnb rho' / advec_vol (the right face's: negated).
// Ghost cells pass their adjoints through; cells outside [0, smax+3] are never read.
struct CellAdj {
  double rho, e, pre;  // primal values, read by the faces
  double rb, eb, pb;
  double M, E, V;
};
// the values cell c's update reads (zero outside [0, smax+3]; the fluxes only for
// interior cells)
struct CellIn {
  double rho, e, pre, rnb, enb, mf_l, mf_r, ef_l, ef_r, vf_l, vf_r;
};
__device__ __forceinline__ CellIn cell_in(const Dev &p, const Dev &a, const Sweep &sw, int smax, int vf_id, int c, int t) {
  CellIn v{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (c >= 0 && c <= smax + 3) {
    v.rho = PS(CLV_DENSITY1, c, t);
    v.e = PS(CLV_ENERGY1, c, t);
    v.pre = PS(CLV_WORK1, c, t);
    v.rnb = AS(CLV_DENSITY1, c, t);
    v.enb = AS(CLV_ENERGY1, c, t);
    if (c >= 2 && c <= smax + 1) {
      v.mf_l = PS(CLV_WORK2, c, t);
      v.mf_r = PS(CLV_WORK2, c + 1, t);
      v.ef_l = PS(CLV_WORK3, c, t);
      v.ef_r = PS(CLV_WORK3, c + 1, t);
      v.vf_l = PS(vf_id, c, t);
      v.vf_r = PS(vf_id, c + 1, t);
    }
  }
  return v;
}
__device__ __forceinline__ CellAdj cell_adj_of(const CellIn &v, bool interior) {
  // scalars, and one struct built at the end, so that the cell stays in registers
  double rb = v.rnb, eb = v.enb, pb = 0.0, M = 0.0, E = 0.0, V = 0.0;
  if (interior) {
    const double pre_mass_s = v.rho * v.pre;
    const double post_mass = pre_mass_s + v.mf_l - v.mf_r;
    const double e_new = (v.e * pre_mass_s + v.ef_l - v.ef_r) / post_mass;
    const double advec_vol = v.pre + v.vf_l - v.vf_r;
    const double rho_new = post_mass / advec_vol;
    // d rho'/d rho = pre / advec_vol, d rho'/d pre = (rho - rho') / advec_vol,
    // d e'/d e = rho pre / post_mass, d e'/d rho = (e - e') pre / post_mass,
    // d e'/d pre = (e - e') rho / post_mass
    const double de = (v.e - e_new) / post_mass;
    rb = v.rnb * v.pre / advec_vol + v.enb * de * v.pre;
    eb = v.enb * (v.rho * v.pre) / post_mass;
    pb = v.rnb * (v.rho - rho_new) / advec_vol + v.enb * de * v.rho;
    M = v.rnb / advec_vol - v.enb * e_new / post_mass;
    E = v.enb / post_mass;
    V = v.rnb * rho_new / advec_vol;
  }
  return CellAdj{v.rho, v.e, v.pre, rb, eb, pb, M, E, V};
}
__device__ __forceinline__ CellAdj cell_adj(const Dev &p, const Dev &a, const Sweep &sw, int smax, int vf_id, int c, int t) {
  return cell_adj_of(cell_in(p, a, sw, smax, vf_id, c, t), c >= 2 && c <= smax + 1);
}

// Face f's adjoint: its total mass / energy flux adjoints (the external mass-flux
// adjoint mfb0, plus the update terms of its right cell f and left cell f-1 where those
// are interior), then the face differentiated once (face_bar): g[1] = donor pre share,
// g[2..4] = rho of upwind / donor / downwind, g[5..7] = e of them.  Returns the face's
// whole vf share.
__device__ __forceinline__ double face_adj(double vfv, double mfb0, bool has_r, const CellAdj &cr, bool has_l, double Ml, double El, double Vl,
                                           double pre_dn, double r_uw, double r_dn, double r_dw, double e_uw, double e_dn, double e_dw,
                                           double vd_face, double vd_dif, double g[8]) {
  double mfb = mfb0, efb = 0.0, vfb = 0.0;
  if (has_r) {
    mfb += cr.M;
    efb += cr.E;
    vfb -= cr.V;
  }
  if (has_l) {
    mfb -= Ml;
    efb -= El;
    vfb += Vl;
  }
  face_bar(vfv, pre_dn, r_uw, r_dn, r_dw, e_uw, e_dn, e_dw, vd_face, vd_dif, mfb, efb, g);
  return vfb + g[0];
}
// a face's shares into cell w when w is the face's upwind / donor / downwind cell
__device__ __forceinline__ void take_shares(CellAdj &w, bool uw, bool dn, bool dw, const double g[8]) {
  if (uw) {
    w.rb += g[2];
    w.eb += g[5];
  }
  if (dn) {
    w.rb += g[3];
    w.eb += g[6];
    w.pb += g[1];
  }
  if (dw) {
    w.rb += g[4];
    w.eb += g[7];
  }
}

// C1 (x sweep), faces and cells of lines [2, ny+1] in one pass.  One wave per 58 cells of
// a line (blocks of 64 x 4, one line per wave): lane l sits at position s = s0 - 3 + l.
// Each lane takes its cell's own update (cell_adj), differentiates its face s (the
// cells the face reads come from lanes l-2..l+1), and gathers its cell's shares from the
// faces of lanes l-1..l+2, in face order (upwind, donor, downwind within a face): the
// order of a gather over faces c-1..c+2.  Lanes 3..60 own their cell and face.  New
// rho1 / e1 adjoints -> rn / en over the whole storage extent (lines the sweep leaves
// alone and column nx+4 pass through), pre adjoint -> p.work5 (cells [0, nx+3]), each
// face's vf share -> a.work3 (faces [2, nx+3]).  The mass-flux adjoint it consumes is
// cleared by C3.
constexpr int kCellXOwn = 58;
constexpr int kCellYSeg = 16;  // cells per thread of adj_cell_y_k
__global__ __launch_bounds__(256) void adj_cell_x_k(Dev p, Dev a, double *rn, double *en) {
  const Sweep sw{1};
  const int smax = p.nx, tmax = p.ny;
  const int lane = threadIdx.x, t = blockIdx.y * blockDim.y + threadIdx.y;
  const int s0 = blockIdx.x * kCellXOwn, s = s0 - 3 + lane;
  if (t > tmax + 4) return;  // whole waves
  const bool owned = lane >= 3 && lane < 3 + kCellXOwn && s <= smax + 4;
  const int64_t o = sw.at(a, s, t);
  if (owned && (t < 2 || t > tmax + 1 || s == smax + 4)) {
    rn[o] = AS(CLV_DENSITY1, s, t);
    en[o] = AS(CLV_ENERGY1, s, t);
  }
  if (t < 2 || t > tmax + 1) return;  // whole waves
  const CellAdj u = cell_adj(p, a, sw, smax, CLV_VOL_FLUX_X, s, t);
  const bool fval = s >= 2 && s <= smax + 3 && lane >= 2 && lane <= 62;
  const double vfv = fval ? PS(CLV_VOL_FLUX_X, s, t) : 0.0;
  const FaceRoles r = face_roles(vfv, s, smax);
  const int lu = lane + r.uw - s, ld = lane + r.dn - s, lw = lane + r.dw - s;
  const double pre_dn = __shfl(u.pre, ld), r_uw = __shfl(u.rho, lu), r_dn = __shfl(u.rho, ld), r_dw = __shfl(u.rho, lw);
  const double e_uw = __shfl(u.e, lu), e_dn = __shfl(u.e, ld), e_dw = __shfl(u.e, lw);
  const double Ml = __shfl(u.M, lane - 1), El = __shfl(u.E, lane - 1), Vl = __shfl(u.V, lane - 1);
  double g[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (fval) {
    const double vfs = face_adj(vfv, AS(CLV_MASS_FLUX_X, s, t), s <= smax + 1, u, s - 1 >= 2 && s - 1 <= smax + 1, Ml, El, Vl, pre_dn, r_uw, r_dn,
                                r_dw, e_uw, e_dn, e_dw, PM(CLV_VERTEXDX, s), PM(CLV_VERTEXDX, r.dif), g);
    if (owned) ST_STREAM(AS(CLV_WORK3, s, t), vfs);
  }
  // role cells as offsets + 2 from the face (uw in [-2, 1], dn and dw in [-1, 0]); bit 6: a face
  const int roles = fval ? (64 | (r.uw - s + 2) | ((r.dn - s + 2) << 2) | ((r.dw - s + 2) << 4)) : 0;
  CellAdj w = u;
#pragma unroll
  for (int j = -1; j <= 2; ++j) {
    const int src = lane + j, rr = __shfl(roles, src);
    double h[8];
#pragma unroll
    for (int q = 1; q < 8; ++q) h[q] = __shfl(g[q], src);
    const int want = 2 - j;  // this cell seen from face s + j
    if (rr & 64) take_shares(w, (rr & 3) == want, ((rr >> 2) & 3) == want, ((rr >> 4) & 3) == want, h);
  }
  if (owned && s <= smax + 3) {
    rn[o] = w.rb;
    en[o] = w.eb;
    ST_STREAM(PS(CLV_WORK5, s, t), w.pb);
  }
}

// C1 (y sweep), the same pass with lanes along x: each thread walks the cells of one
// segment [c0, c0 + SEG - 1] of its line (positions along y) in order, holding the
// adjoints of the four cells f-2..f+1 that face f reads in registers; cell c is complete
// after face c+2.
// Faces from c0-1 on are differentiated (the first few again by the segment before), each
// face's vf share written by the segment that owns position f.  Outputs as adj_cell_x_k.
template <int SEG>
__global__ __launch_bounds__(256) void adj_cell_y_k(Dev p, Dev a, double *rn, double *en) {
  const Sweep sw{2};
  const int smax = p.ny, tmax = p.nx;
  const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  const int c0 = k * SEG;
  if (t > tmax + 4 || c0 > smax + 4) return;
  const int c1 = min(c0 + SEG - 1, smax + 4);
  if (t < 2 || t > tmax + 1 || c1 == smax + 4) {
    for (int c = (t < 2 || t > tmax + 1) ? c0 : smax + 4; c <= c1; ++c) {
      rn[sw.at(a, c, t)] = AS(CLV_DENSITY1, c, t);
      en[sw.at(a, c, t)] = AS(CLV_ENERGY1, c, t);
    }
    if (t < 2 || t > tmax + 1) return;
  }
  const int cl = min(c1, smax + 3);
  if (c0 > cl) return;
  auto put = [&](const CellAdj &w, int c) {
    rn[sw.at(a, c, t)] = w.rb;
    en[sw.at(a, c, t)] = w.eb;
    ST_STREAM(PS(CLV_WORK5, c, t), w.pb);
  };
  const int fs = max(2, c0 - 1), fe = min(smax + 3, cl + 2);
  CellAdj w0 = cell_adj(p, a, sw, smax, CLV_VOL_FLUX_Y, fs - 2, t), w1 = cell_adj(p, a, sw, smax, CLV_VOL_FLUX_Y, fs - 1, t);
  CellAdj w2 = cell_adj(p, a, sw, smax, CLV_VOL_FLUX_Y, fs, t), w3 = cell_adj(p, a, sw, smax, CLV_VOL_FLUX_Y, fs + 1, t);
  // face values one face ahead, vertexdy of f-1..f+1
  double vfv = PS(CLV_VOL_FLUX_Y, fs, t), mfb0 = AS(CLV_MASS_FLUX_Y, fs, t);
  double vdm = PM(CLV_VERTEXDY, fs - 1), vd0 = PM(CLV_VERTEXDY, fs), vdp = PM(CLV_VERTEXDY, fs + 1);
  for (int f = fs; f <= fe; ++f) {
    // everything the next step reads, before this step's stores (which the compiler
    // cannot order loads past)
    const CellIn nin = cell_in(p, a, sw, smax, CLV_VOL_FLUX_Y, f + 2, t);
    const bool more = f + 1 <= smax + 3;
    const double vf_n = more ? PS(CLV_VOL_FLUX_Y, f + 1, t) : 0.0, mf_n = more ? AS(CLV_MASS_FLUX_Y, f + 1, t) : 0.0;
    const double vd_n = f + 2 <= smax + 4 ? PM(CLV_VERTEXDY, f + 2) : 0.0;
    const FaceRoles r = face_roles(vfv, f, smax);
    // role cells as window slots: slot q holds cell f - 2 + q
    const int qu = r.uw - f + 2, qd = r.dn - f + 2, qw = r.dw - f + 2;
    const double r_uw = qu == 0 ? w0.rho : qu == 1 ? w1.rho : qu == 2 ? w2.rho : w3.rho;
    const double e_uw = qu == 0 ? w0.e : qu == 1 ? w1.e : qu == 2 ? w2.e : w3.e;
    double g[8];
    const double vd_dif = r.dif == f - 1 ? vdm : r.dif == f ? vd0 : vdp;  // dif in [f-1, f+1]
    const double vfs = face_adj(vfv, mfb0, f <= smax + 1, w2, f - 1 >= 2 && f - 1 <= smax + 1, w1.M, w1.E, w1.V, qd == 1 ? w1.pre : w2.pre,
                                r_uw, qd == 1 ? w1.rho : w2.rho, qw == 1 ? w1.rho : w2.rho, e_uw, qd == 1 ? w1.e : w2.e,
                                qw == 1 ? w1.e : w2.e, vd0, vd_dif, g);
    if (f >= c0 && f <= c1) ST_STREAM(AS(CLV_WORK3, f, t), vfs);
    take_shares(w0, qu == 0, false, false, g);
    take_shares(w1, qu == 1, qd == 1, qw == 1, g);
    take_shares(w2, qu == 2, qd == 2, qw == 2, g);
    take_shares(w3, qu == 3, false, false, g);
    if (f - 2 >= c0) put(w0, f - 2);
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = cell_adj_of(nin, f + 2 >= 2 && f + 2 <= smax + 1);
    vfv = vf_n;
    mfb0 = mf_n;
    vdm = vd0;
    vd0 = vdp;
    vdp = vd_n;
  }
  // the window now holds cells fe-1..fe+2; those up to cl are complete
  if (fe - 1 >= c0 && fe - 1 <= cl) put(w0, fe - 1);
  if (fe >= c0 && fe <= cl) put(w1, fe);
  if (fe + 1 >= c0 && fe + 1 <= cl) put(w2, fe + 1);
}

// C3: pre_vol of cells [0, smax+3] x lines [2, tmax+1] (its adjoint in p.work5).  vf
// along the sweep, faces [0, smax+4] of those lines: vf(f) += pb(f-1) - pb(f) + the
// face's own share (a.work3); sweep 1 also the other direction's flux:
// vf_t(s, t) += pb(s, t-1) - pb(s, t), lines [1, tmax+2].
__global__ void adj_cell_vol_k(Dev p, Dev a, int sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? a.nx : a.ny, tmax = dir == 1 ? a.ny : a.nx;
  int s, t;
  sweep_pos(dir, 0, 1, s, t);
  if (s > smax + 4 || t > tmax + 2) return;
  auto PB = [&](int ss, int tt) { return ss >= 0 && ss <= smax + 3 && tt >= 2 && tt <= tmax + 1 ? PS(CLV_WORK5, ss, tt) : 0.0; };
  const int vf_id = dir == 1 ? CLV_VOL_FLUX_X : CLV_VOL_FLUX_Y, vo_id = dir == 1 ? CLV_VOL_FLUX_Y : CLV_VOL_FLUX_X;
  const int mf_id = dir == 1 ? CLV_MASS_FLUX_X : CLV_MASS_FLUX_Y;
  if (t >= 2 && t <= tmax + 1) {
    double g = PB(s - 1, t) - PB(s, t);
    if (s >= 2 && s <= smax + 3) {
      g += AS(CLV_WORK3, s, t);
      ST_STREAM(AS(mf_id, s, t), 0.0);  // consumed by C1
    }
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

// M0 (primal), nodes s in [0, smax+3] of lines [2, tmax+2]: node flux -> p.work1;
// for s >= 1 node_mass_post -> p.work2, node_mass_pre = post - nf(s-1) + nf(s) -> p.work3
__global__ void adj_mom_values_k(Dev p, int mom_sweep, int dir) {
  const Sweep sw{dir};
  const int smax = dir == 1 ? p.nx : p.ny, tmax = dir == 1 ? p.ny : p.nx;
  int s, t;
  sweep_pos(dir, 0, 2, s, t);
  if (s > smax + 3 || t > tmax + 2) return;
  const double nf = node_flux_val(p, sw, s, t);
  ST_STREAM(PS(CLV_WORK1, s, t), nf);
  if (s < 1) return;
  const int i = dir == 1 ? s : t, j = dir == 1 ? t : s;
  const double post = node_mass_post_val(p, mom_sweep, i, j);
  ST_STREAM(PS(CLV_WORK2, s, t), post);
  ST_STREAM(PS(CLV_WORK3, s, t), post - node_flux_val(p, sw, s - 1, t) + nf);
}

// the nodes node face f (between nodes f and f+1) reads, by the sign of its node flux
// (src/serial/advec_mom.cpp:114-125)
__device__ __forceinline__ FaceRoles mom_roles(double nf, int f) {
  if (nf < 0.0) return FaceRoles{f + 2, f + 1, f, f + 1};
  return FaceRoles{f - 1, f, f + 1, f - 1};
}
// MX (x sweep), one velocity of advec_mom (src/serial/advec_mom.cpp:114-146), faces and
// nodes of lines [2, ny+2] in one pass.  One wave per 58 nodes of a line (blocks of
// 64 x 4, one line per wave), lane l at node s = s0 - 3 + l; lanes 3..60 own their
// node and face.
// - Face s (momflux_bar): its momentum-flux adjoint comes from the updates of nodes s
//   (its right flux, -1 / mpost) and s+1 (its left flux, +1 / mpost); the nodes it reads
//   come from lanes l-1..l+2.  Node-flux share -> a.work5 (the first velocity writes,
//   the second adds); the first velocity's donor node_mass_pre share -> p.work5.
// - Node s: updated nodes (s in [2, nx+2]) differentiate
//   vel' = (vel mpre + mf(s-1) - mf(s)) / mpost: own share vb mpre / mpost, and the
//   node_mass_pre / node_mass_post adjoints vb vel / mpost and -vb vel' / mpost, started by
//   the first velocity in a.work3 / a.work4 and completed by the second; other nodes
//   pass their adjoint through.  Then the velocity shares of faces s-2..s+1 (lanes
//   l-2..l+1), in face order.  After the second velocity node_mass_pre's adjoint
//   (a.work3) gains the donor shares of faces s-1 and s, both velocities' summed per face.
// The new velocity adjoint -> vn over the whole storage extent (lines the sweep leaves
// alone pass through), for the caller to swap in.
constexpr int kMomXOwn = 58;
template <int VEL, bool LAST>
__global__ __launch_bounds__(256) void adj_mom_x_k(Dev p, Dev a, double *vn) {
  const Sweep sw{1};
  const int smax = p.nx, tmax = p.ny;
  const int lane = threadIdx.x, t = blockIdx.y * blockDim.y + threadIdx.y;
  const int s = blockIdx.x * kMomXOwn - 3 + lane;
  if (t > tmax + 4) return;  // whole waves
  const bool owned = lane >= 3 && lane < 3 + kMomXOwn && s <= smax + 4;
  const int64_t o = sw.at(a, s, t);
  if (t < 2 || t > tmax + 2) {  // whole waves
    if (owned) vn[o] = AS(VEL, s, t);
    return;
  }
  const bool node = s >= 0 && s <= smax + 4, upd = s >= 2 && s <= smax + 2;
  const double vb = node ? AS(VEL, s, t) : 0.0, vel = node ? PS(VEL, s, t) : 0.0;
  const double mpost = upd ? PS(CLV_WORK2, s, t) : 1.0;
  const double mpre = s >= 1 && s <= smax + 3 ? PS(CLV_WORK3, s, t) : 0.0;
  const double q = upd ? vb / mpost : 0.0;
  // face s
  const bool fval = s >= 1 && s <= smax + 2 && lane >= 1 && lane <= 61;
  const double nf = fval ? PS(CLV_WORK1, s, t) : 0.0;
  const FaceRoles r = mom_roles(nf, s);
  const int lu = lane + r.uw - s, ld = lane + r.dn - s, lw = lane + r.dw - s;
  const double v_uw = __shfl(vel, lu), v_dn = __shfl(vel, ld), v_dw = __shfl(vel, lw), mpre_dn = __shfl(mpre, ld);
  const double q_r = __shfl(q, lane + 1);
  double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double flux = 0.0, dsum = 0.0;
  if (fval) {
    double fb = 0.0;
    if (s >= 2) fb -= q;
    if (s + 1 <= smax + 2) fb += q_r;
    flux = momflux_bar(nf, mpre_dn, PM(CLV_CELLDX, s), PM(CLV_CELLDX, r.dif), v_uw, v_dn, v_dw, fb, g);
    if (LAST) dsum = PS(CLV_WORK5, s, t) + g[1];  // both velocities' donor shares of face s
    if (owned) {
      if (LAST) {
        AS(CLV_WORK5, s, t) += g[0];
      } else {
        ST_STREAM(AS(CLV_WORK5, s, t), g[0]);
        ST_STREAM(PS(CLV_WORK5, s, t), g[1]);
      }
    }
  }
  const double flux_l = __shfl(flux, lane - 1);
  double gv = vb, mpreb = 0.0;
  if (upd && owned) {
    const double vnew = (vel * mpre + flux_l - flux) / mpost;
    gv = vb * mpre / mpost;
    mpreb = (LAST ? AS(CLV_WORK3, s, t) : 0.0) + vb * vel / mpost;
    ST_STREAM(AS(CLV_WORK4, s, t), (LAST ? AS(CLV_WORK4, s, t) : 0.0) - vb * vnew / mpost);
    if (!LAST) ST_STREAM(AS(CLV_WORK3, s, t), mpreb);
  }
  // role nodes as offsets from the face: uw + 1 in {0, 3}, dn and dw in {0, 1}; bit 6: a face
  const int roles = fval ? (64 | (r.uw - s + 1) | ((r.dn - s) << 2) | ((r.dw - s) << 4)) : 0;
  double gm = 0.0;
#pragma unroll
  for (int j = -2; j <= 1; ++j) {  // face s + j
    const int src = lane + j, rr = __shfl(roles, src);
    const double g2 = __shfl(g[2], src), g3 = __shfl(g[3], src), g4 = __shfl(g[4], src);
    const double ds = LAST ? __shfl(dsum, src) : 0.0;
    if (rr & 64) {
      if ((rr & 3) == 1 - j) gv += g2;
      if (((rr >> 2) & 3) == -j) {
        gv += g3;
        if (LAST) gm += ds;
      }
      if (((rr >> 4) & 3) == -j) gv += g4;
    }
  }
  if (owned) {
    vn[o] = gv;
    if (LAST) ST_STREAM(AS(CLV_WORK3, s, t), mpreb + gm);
  }
}

// MY (y sweep), one velocity, the same pass with lanes along x: each thread walks the
// nodes of one segment [c0, c0 + SEG - 1] of its line in order.  It keeps the four
// nodes s..s+3 that face s+1 reads and the results of faces s-2..s+1 in registers and
// loads one node (and one face's node flux) ahead, so each step reads one new node;
// face s+1 is differentiated at step s and node s then completed.  Outputs as adj_mom_x_k.
struct MomNode {
  double vel, vb, mpost, mpre, q, w3, w4;
};
// node n's values (its mass adjoints, LAST, only where the segment owns it)
template <int VEL, bool LAST>
__device__ __forceinline__ MomNode mom_node_in(const Dev &p, const Dev &a, const Sweep &sw, int smax, int n, int t, bool own) {
  double vel = 0.0, vb = 0.0, mpost = 1.0, mpre = 0.0, q = 0.0, w3 = 0.0, w4 = 0.0;
  if (n >= 0 && n <= smax + 4) {
    vel = PS(VEL, n, t);
    vb = AS(VEL, n, t);
  }
  if (n >= 1 && n <= smax + 3) mpre = PS(CLV_WORK3, n, t);
  if (n >= 2 && n <= smax + 2) {
    mpost = PS(CLV_WORK2, n, t);
    q = vb / mpost;
    if (LAST && own) {
      w3 = AS(CLV_WORK3, n, t);
      w4 = AS(CLV_WORK4, n, t);
    }
  }
  return MomNode{vel, vb, mpost, mpre, q, w3, w4};
}
struct MomFace {
  double flux, g2, g3, g4, ds;
  int roles;  // as adj_mom_x_k's
};
// face f from its node flux nf, the nodes f-1..f+2 (n0..n3) and, LAST, the face's
// first-velocity donor share pw5 and node-flux share aw5; writes the face's shares
// where the segment owns it
template <int VEL, bool LAST>
__device__ __forceinline__ MomFace mom_face_y(const Dev &p, const Dev &a, const Sweep &sw, int smax, int f, int t, bool own, double nf,
                                              const MomNode &n0, const MomNode &n1, const MomNode &n2, const MomNode &n3, double pw5,
                                              double aw5) {
  double flux = 0.0, g2 = 0.0, g3 = 0.0, g4 = 0.0, ds = 0.0;
  int roles = 0;
  if (f >= 1 && f <= smax + 2) {
    const FaceRoles r = mom_roles(nf, f);
    const bool neg = nf < 0.0;  // uw = f+2, dn = f+1, dw = f; else uw = f-1, dn = f, dw = f+1
    double fb = 0.0;
    if (f >= 2) fb -= n1.q;
    if (f + 1 <= smax + 2) fb += n2.q;
    double g[5];
    flux = momflux_bar(nf, neg ? n2.mpre : n1.mpre, PM(CLV_CELLDY, f), PM(CLV_CELLDY, r.dif), neg ? n3.vel : n0.vel, neg ? n2.vel : n1.vel,
                       neg ? n1.vel : n2.vel, fb, g);
    g2 = g[2];
    g3 = g[3];
    g4 = g[4];
    if (LAST) ds = pw5 + g[1];
    if (own) {
      if (LAST) {
        ST_STREAM(AS(CLV_WORK5, f, t), aw5 + g[0]);
      } else {
        ST_STREAM(AS(CLV_WORK5, f, t), g[0]);
        ST_STREAM(PS(CLV_WORK5, f, t), g[1]);
      }
    }
    roles = 64 | (r.uw - f + 1) | ((r.dn - f) << 2) | ((r.dw - f) << 4);
  }
  return MomFace{flux, g2, g3, g4, ds, roles};
}
// node s's share from face s + j
template <bool LAST>
__device__ __forceinline__ void mom_take(double &gv, double &gm, const MomFace &m, int j) {
  if (m.roles & 64) {
    if ((m.roles & 3) == 1 - j) gv += m.g2;
    if (((m.roles >> 2) & 3) == -j) {
      gv += m.g3;
      if (LAST) gm += m.ds;
    }
    if (((m.roles >> 4) & 3) == -j) gv += m.g4;
  }
}
constexpr int kMomYSeg = 16;
template <int VEL, bool LAST>
__global__ __launch_bounds__(256) void adj_mom_y_k(Dev p, Dev a, double *vn) {
  const Sweep sw{2};
  const int smax = p.ny, tmax = p.nx;
  const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  const int c0 = k * kMomYSeg;
  if (t > tmax + 4 || c0 > smax + 4) return;
  const int c1 = min(c0 + kMomYSeg - 1, smax + 4);
  if (t < 2 || t > tmax + 2) {
    for (int s = c0; s <= c1; ++s) vn[sw.at(a, s, t)] = AS(VEL, s, t);
    return;
  }
  auto nf_at = [&](int f) { return f >= 1 && f <= smax + 2 ? PS(CLV_WORK1, f, t) : 0.0; };
  auto pw5_at = [&](int f) { return LAST && f >= 1 && f <= smax + 2 ? PS(CLV_WORK5, f, t) : 0.0; };
  auto aw5_at = [&](int f) { return LAST && f >= c0 && f <= c1 && f >= 1 && f <= smax + 2 ? AS(CLV_WORK5, f, t) : 0.0; };
  // nodes c0-3..c0+3 for the first faces c0-2..c0 and the window
  auto node = [&](int n) { return mom_node_in<VEL, LAST>(p, a, sw, smax, n, t, n >= c0 && n <= c1); };
  const MomNode m0 = node(c0 - 3), m1 = node(c0 - 2), m2 = node(c0 - 1);
  MomNode n0 = node(c0), n1 = node(c0 + 1), n2 = node(c0 + 2), n3 = node(c0 + 3);
  MomFace F0 = mom_face_y<VEL, LAST>(p, a, sw, smax, c0 - 2, t, false, nf_at(c0 - 2), m0, m1, m2, n0, pw5_at(c0 - 2), 0.0);
  MomFace F1 = mom_face_y<VEL, LAST>(p, a, sw, smax, c0 - 1, t, false, nf_at(c0 - 1), m1, m2, n0, n1, pw5_at(c0 - 1), 0.0);
  MomFace F2 = mom_face_y<VEL, LAST>(p, a, sw, smax, c0, t, true, nf_at(c0), m2, n0, n1, n2, pw5_at(c0), aw5_at(c0));
  double nf = nf_at(c0 + 1), pw5 = pw5_at(c0 + 1), aw5 = aw5_at(c0 + 1);
  for (int s = c0; s <= c1; ++s) {
    // window: nodes n0..n3 = s..s+3; faces F0..F2 = s-2..s; the next node and face
    // values load first, ahead of this step's stores
    const MomNode n4 = node(s + 4);
    const double nf_n = nf_at(s + 2), pw5_n = pw5_at(s + 2), aw5_n = aw5_at(s + 2);
    const MomFace F3 = mom_face_y<VEL, LAST>(p, a, sw, smax, s + 1, t, s + 1 <= c1, nf, n0, n1, n2, n3, pw5, aw5);
    // node s
    const bool upd = s >= 2 && s <= smax + 2;
    double gv = n0.vb, mpreb = 0.0;
    if (upd) {
      const double vnew = (n0.vel * n0.mpre + F1.flux - F2.flux) / n0.mpost;
      gv = n0.vb * n0.mpre / n0.mpost;
      mpreb = (LAST ? n0.w3 : 0.0) + n0.vb * n0.vel / n0.mpost;
      ST_STREAM(AS(CLV_WORK4, s, t), (LAST ? n0.w4 : 0.0) - n0.vb * vnew / n0.mpost);
      if (!LAST) ST_STREAM(AS(CLV_WORK3, s, t), mpreb);
    }
    double gm = 0.0;
    mom_take<LAST>(gv, gm, F0, -2);
    mom_take<LAST>(gv, gm, F1, -1);
    mom_take<LAST>(gv, gm, F2, 0);
    mom_take<LAST>(gv, gm, F3, 1);
    vn[sw.at(a, s, t)] = gv;
    if (LAST) ST_STREAM(AS(CLV_WORK3, s, t), mpreb + gm);
    n0 = n1;
    n1 = n2;
    n2 = n3;
    n3 = n4;
    F0 = F1;
    F1 = F2;
    F2 = F3;
    nf = nf_n;
    pw5 = pw5_n;
    aw5 = aw5_n;
  }
}

// M5: node masses.  node_mass_pre(s) = post(s) - nf(s-1) + nf(s) (nodes s in [1, smax+3])
// and post = 0.25 sum rho1 * post_vol over the node's 4 cells.  Per cell (i, j) in
// [0, nx+3] x [0, ny+3]: rho1 and post_vol adjoints from the nodes (i..i+1, j..j+1).
// mpost adjoint of a node = work4 + work3 (mpre's post term); nodes with a mass but
// no update have only the mpre part.  The post_vol adjoint of the cell goes to
// a.work6 and M5b gathers it onto the volume fluxes.
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
  ST_STREAM(AA(CLV_WORK6, i, j), PA(CLV_DENSITY1, i, j) * g);  // post_vol adjoint of cell (i, j)
}

// M5b: post_vol = vol + vfy(j+1) - vfy(j) (mom_sweep 1) or vol + vfx(i+1) - vfx(i)
// (mom_sweep 2) on cells [0, nx+3] x [0, ny+3]: gather onto the fluxes
__global__ void adj_mom_postvol_k(Dev a, int mom_sweep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  auto PV = [&](int ci, int cj) { return ci >= 0 && ci <= a.nx + 3 && cj >= 0 && cj <= a.ny + 3 ? AA(CLV_WORK6, ci, cj) : 0.0; };
  if (mom_sweep == 1) {
    if (i > a.nx + 3 || j > a.ny + 4) return;
    AA(CLV_VOL_FLUX_Y, i, j) += PV(i, j - 1) - PV(i, j);
  } else if (mom_sweep == 2) {
    if (i > a.nx + 4 || j > a.ny + 3) return;
    AA(CLV_VOL_FLUX_X, i, j) += PV(i - 1, j) - PV(i, j);
  }
}

// M6: node fluxes.  nf(s) enters mpre(s) (+1), mpre(s+1) (-1) and the momentum flux
// of face s (a.work5).  nf(s, t) = 0.25 (mf of 4 faces) -> mass-flux adjoints.
__device__ __forceinline__ double nf_bar(const Dev &a, const Sweep &sw, int smax, int tmax, int s, int t) {
  if (t < 2 || t > tmax + 2 || s < 0 || s > smax + 3) return 0.0;
  double g = 0.0;
  if (s >= 1) g += a.f[CLV_WORK3][sw.at(a, s, t)];                     // mpre(s) = ... + nf(s)
  if (s + 1 <= smax + 3) g -= a.f[CLV_WORK3][sw.at(a, s + 1, t)];      // mpre(s+1) = ... - nf(s)
  if (s >= 1 && s <= smax + 2) g += a.f[CLV_WORK5][sw.at(a, s, t)];   // momentum flux of face s
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
// grid of a sweep kernel over ns positions along the sweep and nt lines (sweep_pos)
static dim3 sgrid(int dir, int ns, int nt) { return dir == 1 ? grid2(ns, nt) : grid2(nt, ns); }

static clv_status check(const char *what) { return hip_check(hipGetLastError(), what); }

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
  hipLaunchKernelGGL(adj_visc_cell_k, grid2(nx, ny), kBlk, 0, a->stream, p->d, a->d);
  hipLaunchKernelGGL(adj_visc_gather_k, grid2(nx + 2, ny + 2), kBlk, 0, a->stream, a->d);
  return check("adj viscosity");
}

clv_status adj_PdV(clv_chunk *p, clv_chunk *a, int predict, double dt) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_pdv_cell_k, grid2(nx, ny), kBlk, 0, a->stream, p->d, a->d, predict, dt);
  hipLaunchKernelGGL(adj_pdv_node_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, p->d, a->d, predict, dt);
  return check("adj PdV");
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
  hipLaunchKernelGGL(adj_accel_node_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, p->d, a->d, dt);
  hipLaunchKernelGGL(adj_accel_cell_k, grid2(nx + 2, ny + 2), kBlk, 0, a->stream, p->d, a->d);
  return check("adj accelerate");
}

clv_status adj_flux_calc(clv_chunk *p, clv_chunk *a, double dt) {
  const int nx = p->d.nx, ny = p->d.ny;
  hipLaunchKernelGGL(adj_flux_calc_k, grid2(nx + 2, ny + 2), kBlk, 0, a->stream, p->d, a->d, dt);
  clv_status st = check("adj flux_calc");
  if (st != CLV_OK) return st;
  // the forward writes both fluxes over [2, nx+2] x [2, ny+2] (src/serial/flux_calc.cpp:34-41)
  hipLaunchKernelGGL(adj_zero_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, a->d.f[CLV_VOL_FLUX_X], a->d.pitch, 2, nx + 2, 2, ny + 2);
  hipLaunchKernelGGL(adj_zero_k, grid2(nx + 1, ny + 1), kBlk, 0, a->stream, a->d.f[CLV_VOL_FLUX_Y], a->d.pitch, 2, nx + 2, 2, ny + 2);
  return check("adj flux_calc zero");
}

clv_status adj_advec_cell(clv_chunk *p, clv_chunk *a, int sweep, int dir) {
  const int smax = dir == 1 ? p->d.nx : p->d.ny, tmax = dir == 1 ? p->d.ny : p->d.nx;
  hipLaunchKernelGGL(adj_cell_values_k, sgrid(dir, smax + 4, tmax), kBlk, 0, a->stream, p->d, sweep, dir);
  // the new rho1 / e1 adjoints go to a.work1 / a.work2 (the pass reads the old ones
  // around each cell), then replace the old: by buffer exchange where the chunk owns its
  // arrays, else copied back
  double *rn = a->d.f[CLV_WORK1], *en = a->d.f[CLV_WORK2];
  if (dir == 1)
    hipLaunchKernelGGL(adj_cell_x_k, dim3((smax + 5 + kCellXOwn - 1) / kCellXOwn, (tmax + 5 + 3) / 4), kBlk, 0, a->stream, p->d, a->d, rn, en);
  else
    hipLaunchKernelGGL(adj_cell_y_k<kCellYSeg>, dim3((tmax + 5 + 63) / 64, ((smax + 5 + kCellYSeg - 1) / kCellYSeg + 3) / 4), kBlk, 0, a->stream,
                       p->d, a->d, rn, en);
  if (a->owns) {
    std::swap(a->d.f[CLV_DENSITY1], a->d.f[CLV_WORK1]);
    std::swap(a->d.f[CLV_ENERGY1], a->d.f[CLV_WORK2]);
  } else {
    const size_t n = (size_t)a->d.pitch * (size_t)(a->d.ny + 5);
    hipLaunchKernelGGL(adj_copy_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, a->stream, a->d.f[CLV_DENSITY1], rn, n);
    hipLaunchKernelGGL(adj_copy_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, a->stream, a->d.f[CLV_ENERGY1], en, n);
  }
  hipLaunchKernelGGL(adj_cell_vol_k, sgrid(dir, smax + 5, tmax + 2), kBlk, 0, a->stream, p->d, a->d, sweep, dir);
  return check("adj advec_cell");
}

clv_status adj_advec_mom(clv_chunk *p, clv_chunk *a, int sweep, int dir) {
  const int smax = dir == 1 ? p->d.nx : p->d.ny, tmax = dir == 1 ? p->d.ny : p->d.nx;
  const int mom_sweep = dir + 2 * (sweep - 1);
  hipLaunchKernelGGL(adj_mom_values_k, sgrid(dir, smax + 4, tmax + 1), kBlk, 0, a->stream, p->d, mom_sweep, dir);
  {
    // each velocity's new adjoint goes to a.work1, then replaces the old (buffer
    // exchange where the chunk owns its arrays, else copied back)
    const dim3 gx((smax + 5 + kMomXOwn - 1) / kMomXOwn, (tmax + 5 + 3) / 4);
    const dim3 gy((tmax + 5 + 63) / 64, ((smax + 5 + kMomYSeg - 1) / kMomYSeg + 3) / 4);
    for (int w = 0; w < 2; ++w) {
      const int vid = w == 0 ? CLV_XVEL1 : CLV_YVEL1;
      double *vn = a->d.f[CLV_WORK1];
      if (dir == 1 && w == 0)
        hipLaunchKernelGGL((adj_mom_x_k<CLV_XVEL1, false>), gx, kBlk, 0, a->stream, p->d, a->d, vn);
      else if (dir == 1)
        hipLaunchKernelGGL((adj_mom_x_k<CLV_YVEL1, true>), gx, kBlk, 0, a->stream, p->d, a->d, vn);
      else if (w == 0)
        hipLaunchKernelGGL((adj_mom_y_k<CLV_XVEL1, false>), gy, kBlk, 0, a->stream, p->d, a->d, vn);
      else
        hipLaunchKernelGGL((adj_mom_y_k<CLV_YVEL1, true>), gy, kBlk, 0, a->stream, p->d, a->d, vn);
      if (a->owns) {
        std::swap(a->d.f[vid], a->d.f[CLV_WORK1]);
      } else {
        const size_t n = (size_t)a->d.pitch * (size_t)(a->d.ny + 5);
        hipLaunchKernelGGL(adj_copy_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, a->stream, a->d.f[vid], vn, n);
      }
    }
  }
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
