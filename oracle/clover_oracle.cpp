/*
 * CPU oracle for the CloverLeaf hydro hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference serial model (src/serial/<kernel>.cpp) expression by
 * expression, on x-contiguous arrays.  Every loop below keeps the reference's
 * iteration ranges and floating-point expression order; the OpenMP pragmas
 * only spread independent cells over threads (each iteration writes its own
 * cells and reads no cell written by another iteration), so results do not
 * depend on the thread count.  The one order-dependent loop (the
 * field_summary sums) runs serially in the reference's order.
 *
 * The hydro kernels are templates over the field type: T = double is the
 * oracle (the orc_* entry points); T = Dual, a forward-mode dual number, is
 * its tangent-linear model (the orc_tl_* entry points), which carries one
 * directional derivative through the same expressions.
 *
 * Build with -O2 -ffp-contract=off (oracle/Makefile): no fused multiply-add,
 * no fast-math, exactly like the reference built with g++ -O3.
 * How each function is pinned is stated in clover_oracle.h.
 */
#include "clover_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

namespace {

// Forward-mode dual number: a value and one directional derivative (the
// tangent-linear model of every expression it passes through).  Branches and
// comparisons follow the value, so a piecewise definition (van Leer limiter,
// upwind choice, viscosity switch, fmin/fmax, fabs) is differentiated on its
// active piece.  Ties: fmin/fmax return their first argument, fabs(0) returns
// its argument unchanged (derivative +1).
struct Dual {
  double v, d;
  Dual() = default;
  constexpr Dual(double x) : v(x), d(0.0) {}
  constexpr Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
inline Dual &operator-=(Dual &a, Dual b) { return a = a - b; }
// Kink census (orc_tl_set_kink_eps / orc_tl_kinks).  Every piecewise choice of the
// tangent-linear run goes through one of the comparisons below, fabs, fmin or fmax,
// and turns on the sign of a difference x = a - b.  Perturbing the inputs by
// +-eps along the run's direction moves x by +-eps * dx to first order, so the choice
// differs between the two runs of a central difference when |x| < eps |dx|.  With
// eps > 0 each such choice is counted (a tie that does not move, dx = 0, is not).
double g_kink_eps = 0.0;
long long g_kinks = 0;
inline void kink(double x, double dx) {
  if (g_kink_eps > 0.0 && ::fabs(x) < g_kink_eps * ::fabs(dx)) {
#pragma omp atomic
    ++g_kinks;
  }
}
inline bool operator<(Dual a, Dual b) { kink(a.v - b.v, a.d - b.d); return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { kink(a.v - b.v, a.d - b.d); return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { kink(a.v - b.v, a.d - b.d); return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { kink(a.v - b.v, a.d - b.d); return a.v >= b.v; }
inline Dual fabs(Dual a) { kink(a.v, a.d); return a.v < 0.0 ? -a : a; }
inline Dual fmin(Dual a, Dual b) { kink(a.v - b.v, a.d - b.d); return b.v < a.v ? b : a; }
inline Dual fmax(Dual a, Dual b) { kink(a.v - b.v, a.d - b.d); return b.v > a.v ? b : a; }
inline Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
static_assert(sizeof(Dual) == sizeof(orc_dual), "Dual and orc_dual share a layout");

// A chunk whose 2D arrays hold T (double, or Dual for the tangent-linear model);
// the layout of orc_chunk (T = double) and orc_tl_chunk (T = Dual).
template <class T>
struct View {
  int nx, ny, px;
  T *f[ORC_NUM_2D];
  double *m[ORC_NUM_1D];
};
inline View<double> *V(orc_chunk *c) { return reinterpret_cast<View<double> *>(c); }
inline View<Dual> *V(orc_tl_chunk *c) { return reinterpret_cast<View<Dual> *>(c); }

}  // namespace

#define G_SMALL 1.0e-16 /* driver/definitions.h:32 */
#define G_BIG 1.0e+21   /* driver/definitions.h:33 */

/* element (i, j) of 2D array p with pitch px */
#define AT(p, i, j) (p)[(size_t)(j) * (size_t)px + (size_t)(i)]

static inline int imin(int a, int b) { return a < b ? a : b; }

/* spread a loop over threads only when the grid is large enough to pay for it */
#define ORC_PAR(c) ((long)(c)->nx * (long)(c)->ny >= 65536L)

int orc_chunk_alloc(orc_chunk *c, int nx, int ny) {
  memset(c, 0, sizeof(*c));
  c->nx = nx;
  c->ny = ny;
  c->px = nx + 5;
  size_t n2 = (size_t)(nx + 5) * (size_t)(ny + 5);
  for (int a = 0; a < ORC_NUM_2D; ++a) {
    c->f[a] = (double *)calloc(n2, sizeof(double));
    if (!c->f[a]) return -1;
  }
  for (int a = 0; a < ORC_NUM_1D; ++a) {
    int n = (a == ORC_CELLX || a == ORC_VERTEXX || a == ORC_CELLDX || a == ORC_VERTEXDX) ? nx + 5 : ny + 5;
    c->m[a] = (double *)calloc((size_t)n, sizeof(double));
    if (!c->m[a]) return -1;
  }
  return 0;
}

void orc_chunk_free(orc_chunk *c) {
  for (int a = 0; a < ORC_NUM_2D; ++a) free(c->f[a]);
  for (int a = 0; a < ORC_NUM_1D; ++a) free(c->m[a]);
  memset(c, 0, sizeof(*c));
}

/* src/serial/ideal_gas.cpp:28-48 (kernel), :55-66 (driver: predict selects time level 1). */
template <class T>
static void k_ideal_gas(View<T> *c, int predict) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const T *density = c->f[predict ? ORC_DENSITY1 : ORC_DENSITY0];
  const T *energy = c->f[predict ? ORC_ENERGY1 : ORC_ENERGY0];
  T *pressure = c->f[ORC_PRESSURE], *soundspeed = c->f[ORC_SOUNDSPEED];
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 2; j++) {
    for (int i = x_min + 1; i < x_max + 2; i++) {
      T v = 1.0 / AT(density, i, j);
      AT(pressure, i, j) = (1.4 - 1.0) * AT(density, i, j) * AT(energy, i, j);
      T pressurebyenergy = (1.4 - 1.0) * AT(density, i, j);
      T pressurebyvolume = -AT(density, i, j) * AT(pressure, i, j);
      T sound_speed_squared = v * v * (AT(pressure, i, j) * pressurebyenergy - pressurebyvolume);
      AT(soundspeed, i, j) = sqrt(sound_speed_squared);
    }
  }
}

/* src/serial/viscosity.cpp:30-68.  Keeps the reference's diry test on pgradx (:57). */
template <class T>
static void k_viscosity(View<T> *c) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const double *celldx = c->m[ORC_CELLDX], *celldy = c->m[ORC_CELLDY];
  const T *density0 = c->f[ORC_DENSITY0], *pressure = c->f[ORC_PRESSURE];
  const T *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0];
  T *viscosity = c->f[ORC_VISCOSITY];
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 2; j++) {
    for (int i = x_min + 1; i < x_max + 2; i++) {
      T ugrad = (AT(xvel0, i + 1, j + 0) + AT(xvel0, i + 1, j + 1)) - (AT(xvel0, i, j) + AT(xvel0, i + 0, j + 1));
      T vgrad = (AT(yvel0, i + 0, j + 1) + AT(yvel0, i + 1, j + 1)) - (AT(yvel0, i, j) + AT(yvel0, i + 1, j + 0));
      T div = (celldx[i] * (ugrad) + celldy[j] * (vgrad));
      T strain2 = 0.5 * (AT(xvel0, i + 0, j + 1) + AT(xvel0, i + 1, j + 1) - AT(xvel0, i, j) - AT(xvel0, i + 1, j + 0)) / celldy[j] +
                       0.5 * (AT(yvel0, i + 1, j + 0) + AT(yvel0, i + 1, j + 1) - AT(yvel0, i, j) - AT(yvel0, i + 0, j + 1)) / celldx[i];
      T pgradx = (AT(pressure, i + 1, j + 0) - AT(pressure, i - 1, j + 0)) / (celldx[i] + celldx[i + 1]);
      T pgrady = (AT(pressure, i + 0, j + 1) - AT(pressure, i + 0, j - 1)) / (celldy[j] + celldy[j + 2]);
      T pgradx2 = pgradx * pgradx;
      T pgrady2 = pgrady * pgrady;
      T limiter = ((0.5 * (ugrad) / celldx[i]) * pgradx2 + (0.5 * (vgrad) / celldy[j]) * pgrady2 + strain2 * pgradx * pgrady) /
                       fmax(pgradx2 + pgrady2, G_SMALL);
      if ((limiter > 0.0) || (div >= 0.0)) {
        AT(viscosity, i, j) = 0.0;
      } else {
        T dirx = 1.0;
        if (pgradx < 0.0) dirx = -1.0;
        pgradx = dirx * fmax(G_SMALL, fabs(pgradx));
        T diry = 1.0;
        if (pgradx < 0.0) diry = -1.0;
        pgrady = diry * fmax(G_SMALL, fabs(pgrady));
        T pgrad = sqrt(pgradx * pgradx + pgrady * pgrady);
        T xgrad = fabs(celldx[i] * pgrad / pgradx);
        T ygrad = fabs(celldy[j] * pgrad / pgrady);
        T grad = fmin(xgrad, ygrad);
        T grad2 = grad * grad;
        AT(viscosity, i, j) = 2.0 * AT(density0, i, j) * grad2 * limiter * limiter;
      }
    }
  }
}

/* src/serial/calc_dt.cpp:30-111 (kernel) and :116-135 (driver).  The min is
 * order-independent, so the OpenMP min reduction returns the serial value. */
void orc_calc_dt(orc_chunk *c, double dtmin, double dtc_safe, double dtu_safe, double dtv_safe, double dtdiv_safe,
                 double *dt_min_val, int *dtl_control, int *jldt, int *kldt, int *small) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const double *celldx = c->m[ORC_CELLDX], *celldy = c->m[ORC_CELLDY];
  const double *xarea = c->f[ORC_XAREA], *yarea = c->f[ORC_YAREA], *volume = c->f[ORC_VOLUME];
  const double *density0 = c->f[ORC_DENSITY0], *viscosity_a = c->f[ORC_VISCOSITY], *soundspeed = c->f[ORC_SOUNDSPEED];
  const double *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0];
  *small = 0;
  double jk_control = 1.1;
  double dt_min_val0 = G_BIG;
#pragma omp parallel for reduction(min : dt_min_val0) if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 2; j++) {
    for (int i = x_min + 1; i < x_max + 2; i++) {
      double dsx = celldx[i];
      double dsy = celldy[j];
      double cc = AT(soundspeed, i, j) * AT(soundspeed, i, j);
      cc = cc + 2.0 * AT(viscosity_a, i, j) / AT(density0, i, j);
      cc = fmax(sqrt(cc), G_SMALL);
      double dtct = dtc_safe * fmin(dsx, dsy) / cc;
      double div = 0.0;
      double dv1 = (AT(xvel0, i, j) + AT(xvel0, i + 0, j + 1)) * AT(xarea, i, j);
      double dv2 = (AT(xvel0, i + 1, j + 0) + AT(xvel0, i + 1, j + 1)) * AT(xarea, i + 1, j + 0);
      div = div + dv2 - dv1;
      double dtut = dtu_safe * 2.0 * AT(volume, i, j) / fmax(fmax(fabs(dv1), fabs(dv2)), G_SMALL * AT(volume, i, j));
      dv1 = (AT(yvel0, i, j) + AT(yvel0, i + 1, j + 0)) * AT(yarea, i, j);
      dv2 = (AT(yvel0, i + 0, j + 1) + AT(yvel0, i + 1, j + 1)) * AT(yarea, i + 0, j + 1);
      div = div + dv2 - dv1;
      double dtvt = dtv_safe * 2.0 * AT(volume, i, j) / fmax(fmax(fabs(dv1), fabs(dv2)), G_SMALL * AT(volume, i, j));
      div = div / (2.0 * AT(volume, i, j));
      double dtdivt;
      if (div < -G_SMALL) {
        dtdivt = dtdiv_safe * (-1.0 / div);
      } else {
        dtdivt = G_BIG;
      }
      double mins = fmin(dtct, fmin(dtut, fmin(dtvt, fmin(dtdivt, G_BIG))));
      dt_min_val0 = fmin(mins, dt_min_val0);
    }
  }
  *dt_min_val = dt_min_val0;
  /* :80-83, constant control/location bookkeeping */
  *dtl_control = (int)(10.01 * (jk_control - (int)(jk_control)));
  jk_control = jk_control - (jk_control - (int)(jk_control));
  *jldt = ((int)jk_control) % x_max;
  *kldt = (int)(1.f + (jk_control / x_max));
  if (*dt_min_val < dtmin) *small = 1;
}

/* src/serial/PdV.cpp:36-87 */
template <class T>
static void k_PdV(View<T> *c, int predict, double dt) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const T *xarea = c->f[ORC_XAREA], *yarea = c->f[ORC_YAREA], *volume = c->f[ORC_VOLUME];
  const T *density0 = c->f[ORC_DENSITY0], *energy0 = c->f[ORC_ENERGY0];
  const T *pressure = c->f[ORC_PRESSURE], *viscosity = c->f[ORC_VISCOSITY];
  const T *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0];
  const T *xvel1 = c->f[ORC_XVEL1], *yvel1 = c->f[ORC_YVEL1];
  T *density1 = c->f[ORC_DENSITY1], *energy1 = c->f[ORC_ENERGY1];
  if (predict) {
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 2; j++) {
      for (int i = x_min + 1; i < x_max + 2; i++) {
        T left_flux = (AT(xarea, i, j) * (AT(xvel0, i, j) + AT(xvel0, i + 0, j + 1) + AT(xvel0, i, j) + AT(xvel0, i + 0, j + 1))) * 0.25 * dt * 0.5;
        T right_flux = (AT(xarea, i + 1, j + 0) * (AT(xvel0, i + 1, j + 0) + AT(xvel0, i + 1, j + 1) + AT(xvel0, i + 1, j + 0) +
                                                      AT(xvel0, i + 1, j + 1))) *
                            0.25 * dt * 0.5;
        T bottom_flux = (AT(yarea, i, j) * (AT(yvel0, i, j) + AT(yvel0, i + 1, j + 0) + AT(yvel0, i, j) + AT(yvel0, i + 1, j + 0))) * 0.25 * dt * 0.5;
        T top_flux = (AT(yarea, i + 0, j + 1) * (AT(yvel0, i + 0, j + 1) + AT(yvel0, i + 1, j + 1) + AT(yvel0, i + 0, j + 1) +
                                                    AT(yvel0, i + 1, j + 1))) *
                          0.25 * dt * 0.5;
        T total_flux = right_flux - left_flux + top_flux - bottom_flux;
        T volume_change_s = AT(volume, i, j) / (AT(volume, i, j) + total_flux);
        T recip_volume = 1.0 / AT(volume, i, j);
        T energy_change = (AT(pressure, i, j) / AT(density0, i, j) + AT(viscosity, i, j) / AT(density0, i, j)) * total_flux * recip_volume;
        AT(energy1, i, j) = AT(energy0, i, j) - energy_change;
        AT(density1, i, j) = AT(density0, i, j) * volume_change_s;
      }
    }
  } else {
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 2; j++) {
      for (int i = x_min + 1; i < x_max + 2; i++) {
        T left_flux = (AT(xarea, i, j) * (AT(xvel0, i, j) + AT(xvel0, i + 0, j + 1) + AT(xvel1, i, j) + AT(xvel1, i + 0, j + 1))) * 0.25 * dt;
        T right_flux = (AT(xarea, i + 1, j + 0) * (AT(xvel0, i + 1, j + 0) + AT(xvel0, i + 1, j + 1) + AT(xvel1, i + 1, j + 0) +
                                                      AT(xvel1, i + 1, j + 1))) *
                            0.25 * dt;
        T bottom_flux = (AT(yarea, i, j) * (AT(yvel0, i, j) + AT(yvel0, i + 1, j + 0) + AT(yvel1, i, j) + AT(yvel1, i + 1, j + 0))) * 0.25 * dt;
        T top_flux = (AT(yarea, i + 0, j + 1) * (AT(yvel0, i + 0, j + 1) + AT(yvel0, i + 1, j + 1) + AT(yvel1, i + 0, j + 1) +
                                                    AT(yvel1, i + 1, j + 1))) *
                          0.25 * dt;
        T total_flux = right_flux - left_flux + top_flux - bottom_flux;
        T volume_change_s = AT(volume, i, j) / (AT(volume, i, j) + total_flux);
        T recip_volume = 1.0 / AT(volume, i, j);
        T energy_change = (AT(pressure, i, j) / AT(density0, i, j) + AT(viscosity, i, j) / AT(density0, i, j)) * total_flux * recip_volume;
        AT(energy1, i, j) = AT(energy0, i, j) - energy_change;
        AT(density1, i, j) = AT(density0, i, j) * volume_change_s;
      }
    }
  }
}

/* src/serial/revert.cpp:29-41 */
template <class T>
static void k_revert(View<T> *c) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const T *density0 = c->f[ORC_DENSITY0], *energy0 = c->f[ORC_ENERGY0];
  T *density1 = c->f[ORC_DENSITY1], *energy1 = c->f[ORC_ENERGY1];
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 2; j++) {
    for (int i = x_min + 1; i < x_max + 2; i++) {
      AT(density1, i, j) = AT(density0, i, j);
      AT(energy1, i, j) = AT(energy0, i, j);
    }
  }
}

/* src/serial/accelerate.cpp:28-56.  Pressure then viscosity, two separate updates (:46-53). */
template <class T>
static void k_accelerate(View<T> *c, double dt) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const T *xarea = c->f[ORC_XAREA], *yarea = c->f[ORC_YAREA], *volume = c->f[ORC_VOLUME];
  const T *density0 = c->f[ORC_DENSITY0], *pressure = c->f[ORC_PRESSURE], *viscosity = c->f[ORC_VISCOSITY];
  const T *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0];
  T *xvel1 = c->f[ORC_XVEL1], *yvel1 = c->f[ORC_YVEL1];
  T halfdt = 0.5 * dt;
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
    for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
      T stepbymass_s = halfdt / ((AT(density0, i - 1, j - 1) * AT(volume, i - 1, j - 1) + AT(density0, i - 1, j + 0) * AT(volume, i - 1, j + 0) +
                                       AT(density0, i, j) * AT(volume, i, j) + AT(density0, i + 0, j - 1) * AT(volume, i + 0, j - 1)) *
                                      0.25);
      AT(xvel1, i, j) = AT(xvel0, i, j) - stepbymass_s * (AT(xarea, i, j) * (AT(pressure, i, j) - AT(pressure, i - 1, j + 0)) +
                                                          AT(xarea, i + 0, j - 1) * (AT(pressure, i + 0, j - 1) - AT(pressure, i - 1, j - 1)));
      AT(yvel1, i, j) = AT(yvel0, i, j) - stepbymass_s * (AT(yarea, i, j) * (AT(pressure, i, j) - AT(pressure, i + 0, j - 1)) +
                                                          AT(yarea, i - 1, j + 0) * (AT(pressure, i - 1, j + 0) - AT(pressure, i - 1, j - 1)));
      AT(xvel1, i, j) = AT(xvel1, i, j) - stepbymass_s * (AT(xarea, i, j) * (AT(viscosity, i, j) - AT(viscosity, i - 1, j + 0)) +
                                                          AT(xarea, i + 0, j - 1) * (AT(viscosity, i + 0, j - 1) - AT(viscosity, i - 1, j - 1)));
      AT(yvel1, i, j) = AT(yvel1, i, j) - stepbymass_s * (AT(yarea, i, j) * (AT(viscosity, i, j) - AT(viscosity, i + 0, j - 1)) +
                                                          AT(yarea, i - 1, j + 0) * (AT(viscosity, i - 1, j + 0) - AT(viscosity, i - 1, j - 1)));
    }
  }
}

/* src/serial/flux_calc.cpp:27-43 */
template <class T>
static void k_flux_calc(View<T> *c, double dt) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const T *xarea = c->f[ORC_XAREA], *yarea = c->f[ORC_YAREA];
  const T *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0];
  const T *xvel1 = c->f[ORC_XVEL1], *yvel1 = c->f[ORC_YVEL1];
  T *vol_flux_x = c->f[ORC_VOL_FLUX_X], *vol_flux_y = c->f[ORC_VOL_FLUX_Y];
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
    for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
      AT(vol_flux_x, i, j) = 0.25 * dt * AT(xarea, i, j) * (AT(xvel0, i, j) + AT(xvel0, i + 0, j + 1) + AT(xvel1, i, j) + AT(xvel1, i + 0, j + 1));
      AT(vol_flux_y, i, j) = 0.25 * dt * AT(yarea, i, j) * (AT(yvel0, i, j) + AT(yvel0, i + 1, j + 0) + AT(yvel1, i, j) + AT(yvel1, i + 1, j + 0));
    }
  }
}

/* One upwind limited flux of the van Leer remap, shared by both directions of
 * advec_cell (src/serial/advec_cell.cpp:86-110 for x, :178-202 for y).
 * a_* are density1 (or energy1) at upwind/donor/downwind. */
template <class T>
static inline T vl_limiter(T sigma, T sigma3, T sigma4, T uw, T dn, T dw) {
  const T one_by_six = 1.0 / 6.0;
  T diffuw = dn - uw;
  T diffdw = dw - dn;
  T wind = 1.0;
  if (diffdw <= 0.0) wind = -1.0;
  T limiter;
  if (diffuw * diffdw > 0.0) {
    limiter = (1.0 - sigma) * wind * fmin(fmin(fabs(diffuw), fabs(diffdw)), one_by_six * (sigma3 * fabs(diffuw) + sigma4 * fabs(diffdw)));
  } else {
    limiter = 0.0;
  }
  return limiter;
}

/* src/serial/advec_cell.cpp:28-220.  Work arrays as bound by the driver
 * (:228-231): pre_vol = work1, post_vol = work2, ener_flux = work7. */
template <class T>
static void k_advec_cell(View<T> *c, int sweep_number, int dir) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const double *vertexdx = c->m[ORC_VERTEXDX], *vertexdy = c->m[ORC_VERTEXDY];
  const T *volume = c->f[ORC_VOLUME];
  const T *vol_flux_x = c->f[ORC_VOL_FLUX_X], *vol_flux_y = c->f[ORC_VOL_FLUX_Y];
  T *density1 = c->f[ORC_DENSITY1], *energy1 = c->f[ORC_ENERGY1];
  T *mass_flux_x = c->f[ORC_MASS_FLUX_X], *mass_flux_y = c->f[ORC_MASS_FLUX_Y];
  T *pre_vol = c->f[ORC_WORK1], *post_vol = c->f[ORC_WORK2], *ener_flux = c->f[ORC_WORK7];

  if (dir == 1) {
    if (sweep_number == 1) {
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min - 2 + 1; j < y_max + 2 + 2; j++) {
        for (int i = x_min - 2 + 1; i < x_max + 2 + 2; i++) {
          AT(pre_vol, i, j) = AT(volume, i, j) + (AT(vol_flux_x, i + 1, j + 0) - AT(vol_flux_x, i, j) + AT(vol_flux_y, i + 0, j + 1) - AT(vol_flux_y, i, j));
          AT(post_vol, i, j) = AT(pre_vol, i, j) - (AT(vol_flux_x, i + 1, j + 0) - AT(vol_flux_x, i, j));
        }
      }
    } else {
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min - 2 + 1; j < y_max + 2 + 2; j++) {
        for (int i = x_min - 2 + 1; i < x_max + 2 + 2; i++) {
          AT(pre_vol, i, j) = AT(volume, i, j) + AT(vol_flux_x, i + 1, j + 0) - AT(vol_flux_x, i, j);
          AT(post_vol, i, j) = AT(volume, i, j);
        }
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 2; j++) {
      for (int i = x_min + 1; i < x_max + 2 + 2; i++) {
        int upwind, donor, downwind, dif;
        if (AT(vol_flux_x, i, j) > 0.0) {
          upwind = i - 2;
          donor = i - 1;
          downwind = i;
          dif = donor;
        } else {
          upwind = imin(i + 1, x_max + 2);
          donor = i;
          downwind = i - 1;
          dif = upwind;
        }
        T sigmat = fabs(AT(vol_flux_x, i, j)) / AT(pre_vol, donor, j);
        T sigma3 = (1.0 + sigmat) * (vertexdx[i] / vertexdx[dif]);
        T sigma4 = 2.0 - sigmat;
        T sigmav = sigmat;
        T limiter = vl_limiter<T>(sigmav, sigma3, sigma4, AT(density1, upwind, j), AT(density1, donor, j), AT(density1, downwind, j));
        AT(mass_flux_x, i, j) = AT(vol_flux_x, i, j) * (AT(density1, donor, j) + limiter);
        T sigmam = fabs(AT(mass_flux_x, i, j)) / (AT(density1, donor, j) * AT(pre_vol, donor, j));
        limiter = vl_limiter<T>(sigmam, sigma3, sigma4, AT(energy1, upwind, j), AT(energy1, donor, j), AT(energy1, downwind, j));
        AT(ener_flux, i, j) = AT(mass_flux_x, i, j) * (AT(energy1, donor, j) + limiter);
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 2; j++) {
      for (int i = x_min + 1; i < x_max + 2; i++) {
        T pre_mass_s = AT(density1, i, j) * AT(pre_vol, i, j);
        T post_mass_s = pre_mass_s + AT(mass_flux_x, i, j) - AT(mass_flux_x, i + 1, j + 0);
        T post_ener_s = (AT(energy1, i, j) * pre_mass_s + AT(ener_flux, i, j) - AT(ener_flux, i + 1, j + 0)) / post_mass_s;
        T advec_vol_s = AT(pre_vol, i, j) + AT(vol_flux_x, i, j) - AT(vol_flux_x, i + 1, j + 0);
        AT(density1, i, j) = post_mass_s / advec_vol_s;
        AT(energy1, i, j) = post_ener_s;
      }
    }
  } else if (dir == 2) {
    if (sweep_number == 1) {
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min - 2 + 1; j < y_max + 2 + 2; j++) {
        for (int i = x_min - 2 + 1; i < x_max + 2 + 2; i++) {
          AT(pre_vol, i, j) = AT(volume, i, j) + (AT(vol_flux_y, i + 0, j + 1) - AT(vol_flux_y, i, j) + AT(vol_flux_x, i + 1, j + 0) - AT(vol_flux_x, i, j));
          AT(post_vol, i, j) = AT(pre_vol, i, j) - (AT(vol_flux_y, i + 0, j + 1) - AT(vol_flux_y, i, j));
        }
      }
    } else {
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min - 2 + 1; j < y_max + 2 + 2; j++) {
        for (int i = x_min - 2 + 1; i < x_max + 2 + 2; i++) {
          AT(pre_vol, i, j) = AT(volume, i, j) + AT(vol_flux_y, i + 0, j + 1) - AT(vol_flux_y, i, j);
          AT(post_vol, i, j) = AT(volume, i, j);
        }
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 2 + 2; j++) {
      for (int i = x_min + 1; i < x_max + 2; i++) {
        int upwind, donor, downwind, dif;
        if (AT(vol_flux_y, i, j) > 0.0) {
          upwind = j - 2;
          donor = j - 1;
          downwind = j;
          dif = donor;
        } else {
          upwind = imin(j + 1, y_max + 2);
          donor = j;
          downwind = j - 1;
          dif = upwind;
        }
        T sigmat = fabs(AT(vol_flux_y, i, j)) / AT(pre_vol, i, donor);
        T sigma3 = (1.0 + sigmat) * (vertexdy[j] / vertexdy[dif]);
        T sigma4 = 2.0 - sigmat;
        T sigmav = sigmat;
        T limiter = vl_limiter<T>(sigmav, sigma3, sigma4, AT(density1, i, upwind), AT(density1, i, donor), AT(density1, i, downwind));
        AT(mass_flux_y, i, j) = AT(vol_flux_y, i, j) * (AT(density1, i, donor) + limiter);
        T sigmam = fabs(AT(mass_flux_y, i, j)) / (AT(density1, i, donor) * AT(pre_vol, i, donor));
        limiter = vl_limiter<T>(sigmam, sigma3, sigma4, AT(energy1, i, upwind), AT(energy1, i, donor), AT(energy1, i, downwind));
        AT(ener_flux, i, j) = AT(mass_flux_y, i, j) * (AT(energy1, i, donor) + limiter);
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 2; j++) {
      for (int i = x_min + 1; i < x_max + 2; i++) {
        T pre_mass_s = AT(density1, i, j) * AT(pre_vol, i, j);
        T post_mass_s = pre_mass_s + AT(mass_flux_y, i, j) - AT(mass_flux_y, i + 0, j + 1);
        T post_ener_s = (AT(energy1, i, j) * pre_mass_s + AT(ener_flux, i, j) - AT(ener_flux, i + 0, j + 1)) / post_mass_s;
        T advec_vol_s = AT(pre_vol, i, j) + AT(vol_flux_y, i, j) - AT(vol_flux_y, i + 0, j + 1);
        AT(density1, i, j) = post_mass_s / advec_vol_s;
        AT(energy1, i, j) = post_ener_s;
      }
    }
  }
}

/* Momentum limiter of src/serial/advec_mom.cpp:125-139 (x) / :197-211 (y). */
template <class T>
static inline T mom_flux_of(T flux, T mass_pre_donor, T width, T width_dif, T v_uw, T v_dn, T v_dw) {
  T sigma = fabs(flux) / (mass_pre_donor);
  T vdiffuw = v_dn - v_uw;
  T vdiffdw = v_dw - v_dn;
  T limiter = 0.0;
  if (vdiffuw * vdiffdw > 0.0) {
    T auw = fabs(vdiffuw);
    T adw = fabs(vdiffdw);
    T wind = 1.0;
    if (vdiffdw <= 0.0) wind = -1.0;
    limiter = wind * fmin(fmin(width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / width_dif) / 6.0, auw), adw);
  }
  T advec_vel_s = v_dn + (1.0 - sigma) * limiter;
  return advec_vel_s * flux;
}

/* src/serial/advec_mom.cpp:30-225.  Work arrays as bound by the driver
 * (:234-242): node_flux = work1, node_mass_post = work2, node_mass_pre = work3,
 * mom_flux = work4, pre_vol = work5, post_vol = work6. */
template <class T>
static void k_advec_mom(View<T> *c, int which_vel, int sweep_number, int direction) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  T *vel1 = c->f[which_vel == 1 ? ORC_XVEL1 : ORC_YVEL1];
  const T *mass_flux_x = c->f[ORC_MASS_FLUX_X], *vol_flux_x = c->f[ORC_VOL_FLUX_X];
  const T *mass_flux_y = c->f[ORC_MASS_FLUX_Y], *vol_flux_y = c->f[ORC_VOL_FLUX_Y];
  const T *volume = c->f[ORC_VOLUME], *density1 = c->f[ORC_DENSITY1];
  T *node_flux = c->f[ORC_WORK1], *node_mass_post = c->f[ORC_WORK2], *node_mass_pre = c->f[ORC_WORK3];
  T *mom_flux = c->f[ORC_WORK4], *pre_vol = c->f[ORC_WORK5], *post_vol = c->f[ORC_WORK6];
  const double *celldx = c->m[ORC_CELLDX], *celldy = c->m[ORC_CELLDY];

  int mom_sweep = direction + 2 * (sweep_number - 1);
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min - 2 + 1; j < y_max + 2 + 2; j++) {
    for (int i = x_min - 2 + 1; i < x_max + 2 + 2; i++) {
      if (mom_sweep == 1) { /* x 1 */
        AT(post_vol, i, j) = AT(volume, i, j) + AT(vol_flux_y, i + 0, j + 1) - AT(vol_flux_y, i, j);
        AT(pre_vol, i, j) = AT(post_vol, i, j) + AT(vol_flux_x, i + 1, j + 0) - AT(vol_flux_x, i, j);
      } else if (mom_sweep == 2) { /* y 1 */
        AT(post_vol, i, j) = AT(volume, i, j) + AT(vol_flux_x, i + 1, j + 0) - AT(vol_flux_x, i, j);
        AT(pre_vol, i, j) = AT(post_vol, i, j) + AT(vol_flux_y, i + 0, j + 1) - AT(vol_flux_y, i, j);
      } else if (mom_sweep == 3) { /* x 2 */
        AT(post_vol, i, j) = AT(volume, i, j);
        AT(pre_vol, i, j) = AT(post_vol, i, j) + AT(vol_flux_y, i + 0, j + 1) - AT(vol_flux_y, i, j);
      } else if (mom_sweep == 4) { /* y 2 */
        AT(post_vol, i, j) = AT(volume, i, j);
        AT(pre_vol, i, j) = AT(post_vol, i, j) + AT(vol_flux_x, i + 1, j + 0) - AT(vol_flux_x, i, j);
      }
    }
  }

  if (direction == 1) {
    if (which_vel == 1) {
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
        for (int i = x_min - 2 + 1; i < x_max + 2 + 2; i++) {
          AT(node_flux, i, j) = 0.25 * (AT(mass_flux_x, i + 0, j - 1) + AT(mass_flux_x, i, j) + AT(mass_flux_x, i + 1, j - 1) + AT(mass_flux_x, i + 1, j + 0));
        }
      }
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
        for (int i = x_min - 1 + 1; i < x_max + 2 + 2; i++) {
          AT(node_mass_post, i, j) = 0.25 * (AT(density1, i + 0, j - 1) * AT(post_vol, i + 0, j - 1) + AT(density1, i, j) * AT(post_vol, i, j) +
                                             AT(density1, i - 1, j - 1) * AT(post_vol, i - 1, j - 1) + AT(density1, i - 1, j + 0) * AT(post_vol, i - 1, j + 0));
          AT(node_mass_pre, i, j) = AT(node_mass_post, i, j) - AT(node_flux, i - 1, j + 0) + AT(node_flux, i, j);
        }
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
      for (int i = x_min - 1 + 1; i < x_max + 1 + 2; i++) {
        int upwind, donor, downwind, dif;
        if (AT(node_flux, i, j) < 0.0) {
          upwind = i + 2;
          donor = i + 1;
          downwind = i;
          dif = donor;
        } else {
          upwind = i - 1;
          donor = i;
          downwind = i + 1;
          dif = upwind;
        }
        AT(mom_flux, i, j) = mom_flux_of<T>(AT(node_flux, i, j), AT(node_mass_pre, donor, j), celldx[i], celldx[dif], AT(vel1, upwind, j),
                                          AT(vel1, donor, j), AT(vel1, downwind, j));
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
      for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
        AT(vel1, i, j) = (AT(vel1, i, j) * AT(node_mass_pre, i, j) + AT(mom_flux, i - 1, j + 0) - AT(mom_flux, i, j)) / AT(node_mass_post, i, j);
      }
    }
  } else if (direction == 2) {
    if (which_vel == 1) {
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min - 2 + 1; j < y_max + 2 + 2; j++) {
        for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
          AT(node_flux, i, j) = 0.25 * (AT(mass_flux_y, i - 1, j + 0) + AT(mass_flux_y, i, j) + AT(mass_flux_y, i - 1, j + 1) + AT(mass_flux_y, i + 0, j + 1));
        }
      }
#pragma omp parallel for if (ORC_PAR(c))
      for (int j = y_min - 1 + 1; j < y_max + 2 + 2; j++) {
        for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
          AT(node_mass_post, i, j) = 0.25 * (AT(density1, i + 0, j - 1) * AT(post_vol, i + 0, j - 1) + AT(density1, i, j) * AT(post_vol, i, j) +
                                             AT(density1, i - 1, j - 1) * AT(post_vol, i - 1, j - 1) + AT(density1, i - 1, j + 0) * AT(post_vol, i - 1, j + 0));
          AT(node_mass_pre, i, j) = AT(node_mass_post, i, j) - AT(node_flux, i + 0, j - 1) + AT(node_flux, i, j);
        }
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min - 1 + 1; j < y_max + 1 + 2; j++) {
      for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
        int upwind, donor, downwind, dif;
        if (AT(node_flux, i, j) < 0.0) {
          upwind = j + 2;
          donor = j + 1;
          downwind = j;
          dif = donor;
        } else {
          upwind = j - 1;
          donor = j;
          downwind = j + 1;
          dif = upwind;
        }
        AT(mom_flux, i, j) = mom_flux_of<T>(AT(node_flux, i, j), AT(node_mass_pre, i, donor), celldy[j], celldy[dif], AT(vel1, i, upwind),
                                          AT(vel1, i, donor), AT(vel1, i, downwind));
      }
    }
#pragma omp parallel for if (ORC_PAR(c))
    for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
      for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
        AT(vel1, i, j) = (AT(vel1, i, j) * AT(node_mass_pre, i, j) + AT(mom_flux, i + 0, j - 1) - AT(mom_flux, i, j)) / AT(node_mass_post, i, j);
      }
    }
  }
}

/* src/serial/reset_field.cpp:28-51 */
template <class T>
static void k_reset_field(View<T> *c) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  T *density0 = c->f[ORC_DENSITY0], *energy0 = c->f[ORC_ENERGY0], *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0];
  const T *density1 = c->f[ORC_DENSITY1], *energy1 = c->f[ORC_ENERGY1], *xvel1 = c->f[ORC_XVEL1], *yvel1 = c->f[ORC_YVEL1];
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 2; j++) {
    for (int i = x_min + 1; i < x_max + 2; i++) {
      AT(density0, i, j) = AT(density1, i, j);
      AT(energy0, i, j) = AT(energy1, i, j);
    }
  }
#pragma omp parallel for if (ORC_PAR(c))
  for (int j = y_min + 1; j < y_max + 1 + 2; j++) {
    for (int i = x_min + 1; i < x_max + 1 + 2; i++) {
      AT(xvel0, i, j) = AT(xvel1, i, j);
      AT(yvel0, i, j) = AT(yvel1, i, j);
    }
  }
}

/* Reflective halo of one field, src/serial/update_halo.cpp:33-684, for one
 * chunk and one tile (every face external).  The four passes run in the
 * reference's order (bottom, top, left, right); that order fixes the corners.
 * Per-field offsets, including the reference's irregular ones (soundspeed
 * bottom :305, the x-face top :530/:573 and the y-face right :636/:679), come
 * from one table:
 *   bsrc: bottom source row = bsrc + k       (dst row 1 - k)
 *   tdst/tsrc: top dst row y_max + tdst + k, source row y_max + tsrc - k
 *   lsrc: left source col = lsrc + j         (dst col 1 - j)
 *   rdst/rsrc: right dst col x_max + rdst + j, source col x_max + rsrc - j
 *   xext: 1 if the bottom/top passes run one column further (node in x)
 *   yext: 1 if the left/right passes run one row further (node in y)
 *   sy / sx: sign applied by the bottom/top and left/right passes. */
typedef struct { int bsrc, tdst, tsrc, lsrc, rdst, rsrc, xext, yext; double sy, sx; } orc_halo_rule;

static const orc_halo_rule ORC_HALO_RULES[15] = {
    /* density0   */ {2, 2, 1, 2, 2, 1, 0, 0, 1.0, 1.0},
    /* density1   */ {2, 2, 1, 2, 2, 1, 0, 0, 1.0, 1.0},
    /* energy0    */ {2, 2, 1, 2, 2, 1, 0, 0, 1.0, 1.0},
    /* energy1    */ {2, 2, 1, 2, 2, 1, 0, 0, 1.0, 1.0},
    /* pressure   */ {2, 2, 1, 2, 2, 1, 0, 0, 1.0, 1.0},
    /* viscosity  */ {2, 2, 1, 2, 2, 1, 0, 0, 1.0, 1.0},
    /* soundspeed */ {0, 2, 1, 2, 2, 1, 0, 0, 1.0, 1.0},
    /* xvel0      */ {3, 3, 1, 3, 3, 1, 1, 1, 1.0, -1.0},
    /* xvel1      */ {3, 3, 1, 3, 3, 1, 1, 1, 1.0, -1.0},
    /* yvel0      */ {3, 3, 1, 3, 3, 1, 1, 1, -1.0, 1.0},
    /* yvel1      */ {3, 3, 1, 3, 3, 1, 1, 1, -1.0, 1.0},
    /* vol_flux_x */ {3, 2, 0, 3, 3, 1, 1, 0, 1.0, -1.0},
    /* vol_flux_y */ {3, 3, 1, 3, 2, 0, 0, 1, -1.0, 1.0},
    /* mass_flux_x*/ {3, 2, 0, 3, 3, 1, 1, 0, 1.0, -1.0},
    /* mass_flux_y*/ {3, 3, 1, 3, 2, 0, 0, 1, -1.0, 1.0},
};

/* ext[face] (face: 0 left, 1 right, 2 bottom, 3 top, the order of the
 * reference's chunk_neighbour_type, driver/definitions.h:57): whether that face
 * is reflected.  The reference reflects a face of a tile only where both the
 * chunk's and the tile's neighbour on it are external
 * (src/serial/update_halo.cpp:41-71); faces shared with another tile or chunk
 * keep the ghosts their exchange wrote, corners included. */
template <class T>
static void k_update_halo_ext(View<T> *c, const int fields[15], int depth, const int ext[4]) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  for (int f = 0; f < 15; ++f) {
    if (fields[f] != 1) continue;
    const orc_halo_rule *r = &ORC_HALO_RULES[f];
    T *a = c->f[f];
    /* bottom then top: j over x_min-depth .. x_max(+1)+depth */
    if (ext[2])
      for (int j = x_min - depth + 1; j < x_max + r->xext + depth + 2; j++)
        for (int k = 0; k < depth; ++k) AT(a, j, 1 - k) = r->sy * AT(a, j, r->bsrc + k);
    if (ext[3])
      for (int j = x_min - depth + 1; j < x_max + r->xext + depth + 2; j++)
        for (int k = 0; k < depth; ++k) AT(a, j, y_max + r->tdst + k) = r->sy * AT(a, j, y_max + r->tsrc - k);
    /* left then right: k over y_min-depth .. y_max(+1)+depth */
    if (ext[0])
      for (int k = y_min - depth + 1; k < y_max + r->yext + depth + 2; k++)
        for (int j = 0; j < depth; ++j) AT(a, 1 - j, k) = r->sx * AT(a, r->lsrc + j, k);
    if (ext[1])
      for (int k = y_min - depth + 1; k < y_max + r->yext + depth + 2; k++)
        for (int j = 0; j < depth; ++j) AT(a, x_max + r->rdst + j, k) = r->sx * AT(a, x_max + r->rsrc - j, k);
  }
}

template <class T>
static void k_update_halo(View<T> *c, const int fields[15], int depth) {
  static const int all[4] = {1, 1, 1, 1};
  k_update_halo_ext(c, fields, depth, all);
}

/* Ghost layers of tile c on one face copied from the neighbouring tile n,
 * src/serial/update_tile_halo_kernel.cpp:30-804 (the l, r, t and b kernels).
 * The ghosts receive the neighbour's interior layers next to the shared face;
 * node and face fields normal to it skip the shared boundary node, which both
 * tiles compute (xext / yext of ORC_HALO_RULES).  Left/right copy rows
 * 2-depth .. ny+yext+depth+1 (the corner ghosts the bottom/top copies wrote
 * included), bottom/top columns 2-depth .. nx+xext+depth+1. */
template <class T>
static void k_tile_halo(View<T> *c, View<T> *n, int face, const int fields[15], int depth) {
  const size_t pc = c->px, pn = n->px;
  const int nx = c->nx, ny = c->ny;
  for (int f = 0; f < 15; ++f) {
    if (fields[f] != 1) continue;
    const orc_halo_rule *r = &ORC_HALO_RULES[f];
    T *a = c->f[f];
    const T *b = n->f[f];
    if (face == 0)
      for (int k = 2 - depth; k < ny + r->yext + depth + 2; k++)
        for (int j = 0; j < depth; ++j) a[(size_t)k * pc + (1 - j)] = b[(size_t)k * pn + (n->nx + 1 - j)];
    if (face == 1)
      for (int k = 2 - depth; k < ny + r->yext + depth + 2; k++)
        for (int j = 0; j < depth; ++j) a[(size_t)k * pc + (nx + 2 + r->xext + j)] = b[(size_t)k * pn + (2 + r->xext + j)];
    if (face == 2)
      for (int k = 0; k < depth; ++k)
        for (int j = 2 - depth; j < nx + r->xext + depth + 2; j++) a[(size_t)(1 - k) * pc + j] = b[(size_t)(n->ny + 1 - k) * pn + j];
    if (face == 3)
      for (int k = 0; k < depth; ++k)
        for (int j = 2 - depth; j < nx + r->xext + depth + 2; j++)
          a[(size_t)(ny + 2 + r->yext + k) * pc + j] = b[(size_t)(2 + r->yext + k) * pn + j];
  }
}

/* One field's exchange message for one face, src/serial/pack_kernel.cpp:27-303
 * (clover_pack_message_* / clover_unpack_message_*).  field_type 1 cell,
 * 2 vertex, 3 x face, 4 y face (driver/definitions.h:80) sets x_inc / y_inc.
 * Left/right messages hold element offset + j + k*depth for rows k in
 * 2-depth .. ny+y_inc+depth+1 and layers j < depth; bottom/top messages
 * offset + k + j*depth for columns j in 2-depth .. nx+x_inc+depth+1.  Packing
 * reads the interior layers next to the face (the left and bottom ones skip the
 * shared boundary node of node/face data); unpacking writes the ghost layers. */
template <class T>
static void k_pack(View<T> *c, T *a, int field_type, int face, int depth, T *buf, long offset, int unpack) {
  const int px = c->px, nx = c->nx, ny = c->ny;
  const int x_inc = (field_type == 2 || field_type == 3) ? 1 : 0;
  const int y_inc = (field_type == 2 || field_type == 4) ? 1 : 0;
  if (face == 0 || face == 1) {
    for (int k = 2 - depth; k < ny + y_inc + depth + 2; k++)
      for (int j = 0; j < depth; ++j) {
        const long index = offset + j + (long)k * depth;
        if (face == 0) {
          if (unpack) AT(a, 1 - j, k) = buf[index];
          else buf[index] = AT(a, 2 + x_inc + j, k);
        } else {
          if (unpack) AT(a, nx + x_inc + 2 + j, k) = buf[index];
          else buf[index] = AT(a, nx + 1 - j, k);
        }
      }
  } else {
    for (int k = 0; k < depth; ++k)
      for (int j = 2 - depth; j < nx + x_inc + depth + 2; j++) {
        const long index = offset + k + (long)j * depth;
        if (face == 2) {
          if (unpack) AT(a, j, 1 - k) = buf[index];
          else buf[index] = AT(a, j, 2 + y_inc + k);
        } else {
          if (unpack) AT(a, j, ny + y_inc + 2 + k) = buf[index];
          else buf[index] = AT(a, j, ny + 1 - k);
        }
      }
  }
}

/* src/serial/field_summary.cpp:60-92, serial order (linear index, x fastest). */
template <class T>
static void k_field_summary_kernel(View<T> *c, T out[5]) {
  const int px = c->px, xmin = 1, ymin = 1, xmax = c->nx, ymax = c->ny;
  const T *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0], *volume = c->f[ORC_VOLUME];
  const T *density0 = c->f[ORC_DENSITY0], *energy0 = c->f[ORC_ENERGY0], *pressure = c->f[ORC_PRESSURE];
  T vol = 0.0, mass = 0.0, ie = 0.0, ke = 0.0, press = 0.0;
  for (int idx = 0; idx < ((ymax - ymin + 1) * (xmax - xmin + 1)); idx++) {
    const int j = xmin + 1 + idx % (xmax - xmin + 1);
    const int k = ymin + 1 + idx / (xmax - xmin + 1);
    T vsqrd = 0.0;
    for (int kv = k; kv <= k + 1; ++kv) {
      for (int jv = j; jv <= j + 1; ++jv) {
        vsqrd += 0.25 * (AT(xvel0, jv, kv) * AT(xvel0, jv, kv) + AT(yvel0, jv, kv) * AT(yvel0, jv, kv));
      }
    }
    T cell_vol = AT(volume, j, k);
    T cell_mass = cell_vol * AT(density0, j, k);
    vol += cell_vol;
    mass += cell_mass;
    ie += cell_mass * AT(energy0, j, k);
    ke += cell_mass * 0.5 * vsqrd;
    press += cell_vol * AT(pressure, j, k);
  }
  out[0] = vol;
  out[1] = mass;
  out[2] = ie;
  out[3] = ke;
  out[4] = press;
}

void orc_field_summary(orc_chunk *c, double out[5]) {
  orc_ideal_gas(c, 0);
  orc_field_summary_kernel(c, out);
}

/* src/serial/build_field.cpp:29-180: zero fill (calloc already zeroes; kept for re-use). */
void orc_build_field(orc_chunk *c) {
  size_t n2 = (size_t)(c->nx + 5) * (size_t)(c->ny + 5);
  for (int a = 0; a < ORC_NUM_2D; ++a) memset(c->f[a], 0, n2 * sizeof(double));
  for (int a = 0; a < ORC_NUM_1D; ++a) {
    int n = (a == ORC_CELLX || a == ORC_VERTEXX || a == ORC_CELLDX || a == ORC_VERTEXDX) ? c->nx + 5 : c->ny + 5;
    memset(c->m[a], 0, (size_t)n * sizeof(double));
  }
}

/* src/serial/initialise_chunk.cpp:30-85 (one tile, t_left = t_bottom = 1). */
void orc_initialise_chunk(orc_chunk *c, const orc_config *cfg) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  double dx = (cfg->xmax - cfg->xmin) / (double)(cfg->x_cells);
  double dy = (cfg->ymax - cfg->ymin) / (double)(cfg->y_cells);
  double xmin = cfg->xmin + dx * (double)(1 - 1);
  double ymin = cfg->ymin + dy * (double)(1 - 1);
  double *vertexx = c->m[ORC_VERTEXX], *vertexdx = c->m[ORC_VERTEXDX], *vertexy = c->m[ORC_VERTEXY], *vertexdy = c->m[ORC_VERTEXDY];
  double *cellx = c->m[ORC_CELLX], *celldx = c->m[ORC_CELLDX], *celly = c->m[ORC_CELLY], *celldy = c->m[ORC_CELLDY];
  const int xrange = (x_max + 3) - (x_min - 2) + 1;
  const int yrange = (y_max + 3) - (y_min - 2) + 1;
  for (int j = 0; j < xrange; j++) {
    vertexx[j] = xmin + dx * (j - 1 - x_min);
    vertexdx[j] = dx;
  }
  for (int k = 0; k < yrange; k++) {
    vertexy[k] = ymin + dy * (k - 1 - y_min);
    vertexdy[k] = dy;
  }
  const int xrange1 = (x_max + 2) - (x_min - 2) + 1;
  const int yrange1 = (y_max + 2) - (y_min - 2) + 1;
  for (int j = 0; j < xrange1; j++) {
    cellx[j] = 0.5 * (vertexx[j] + vertexx[j + 1]);
    celldx[j] = dx;
  }
  for (int k = 0; k < yrange1; k++) {
    celly[k] = 0.5 * (vertexy[k] + vertexy[k + 1]);
    celldy[k] = dy;
  }
  double *volume = c->f[ORC_VOLUME], *xarea = c->f[ORC_XAREA], *yarea = c->f[ORC_YAREA];
  for (int j = 0; j < yrange1; j++) {
    for (int i = 0; i < xrange1; i++) {
      AT(volume, i, j) = dx * dy;
      AT(xarea, i, j) = celldy[j];
      AT(yarea, i, j) = celldx[i];
    }
  }
}

/* src/serial/generate_chunk.cpp:31-131 */
void orc_generate_chunk(orc_chunk *c, const orc_config *cfg) {
  const int px = c->px, x_min = 1, y_min = 1, x_max = c->nx, y_max = c->ny;
  const int xrange = (x_max + 2) - (x_min - 2) + 1;
  const int yrange = (y_max + 2) - (y_min - 2) + 1;
  double *energy0 = c->f[ORC_ENERGY0], *density0 = c->f[ORC_DENSITY0], *xvel0 = c->f[ORC_XVEL0], *yvel0 = c->f[ORC_YVEL0];
  const double *vertexx = c->m[ORC_VERTEXX], *vertexy = c->m[ORC_VERTEXY], *cellx = c->m[ORC_CELLX], *celly = c->m[ORC_CELLY];
  const orc_state *st = cfg->states;
  for (int j = 0; j < yrange; j++) {
    for (int i = 0; i < xrange; i++) {
      AT(energy0, i, j) = st[0].energy;
      AT(density0, i, j) = st[0].density;
      AT(xvel0, i, j) = st[0].xvel;
      AT(yvel0, i, j) = st[0].yvel;
    }
  }
  for (int s = 1; s < cfg->number_of_states; ++s) {
    for (int j = 0; j < yrange; j++) {
      for (int i = 0; i < xrange; i++) {
        double x_cent = st[s].xmin;
        double y_cent = st[s].ymin;
        int hit = 0;
        if (st[s].geometry == 1) {
          hit = (vertexx[i + 1] >= st[s].xmin && vertexx[i] < st[s].xmax) && (vertexy[j + 1] >= st[s].ymin && vertexy[j] < st[s].ymax);
        } else if (st[s].geometry == 2) {
          double radius = sqrt((cellx[i] - x_cent) * (cellx[i] - x_cent) + (celly[j] - y_cent) * (celly[j] - y_cent));
          hit = radius <= st[s].radius;
        } else if (st[s].geometry == 3) {
          hit = vertexx[i] == x_cent && vertexy[j] == y_cent;
        }
        if (hit) {
          AT(energy0, i, j) = st[s].energy;
          AT(density0, i, j) = st[s].density;
          for (int kt = j; kt <= j + 1; ++kt) {
            for (int jt = i; jt <= i + 1; ++jt) {
              AT(xvel0, jt, kt) = st[s].xvel;
              AT(yvel0, jt, kt) = st[s].yvel;
            }
          }
        }
      }
    }
  }
}

static void orc_mask(int fields[15], const int *ids, int n) {
  for (int i = 0; i < 15; ++i) fields[i] = 0;
  for (int i = 0; i < n; ++i) fields[ids[i]] = 1;
}

/* driver/start.cpp:69-127 for one chunk and one tile. */
void orc_start(orc_chunk *c, const orc_config *cfg, orc_globals *g, double summary0[5]) {
  orc_build_field(c);
  orc_initialise_chunk(c, cfg);
  orc_generate_chunk(c, cfg);
  orc_ideal_gas(c, 0);
  int fields[15];
  const int ids[] = {ORC_DENSITY0, ORC_ENERGY0, ORC_PRESSURE, ORC_VISCOSITY, ORC_DENSITY1,
                     ORC_ENERGY1,  ORC_XVEL0,   ORC_YVEL0,    ORC_XVEL1,     ORC_YVEL1};
  orc_mask(fields, ids, 10);
  orc_update_halo(c, fields, 2);
  g->step = 0;
  g->advect_x = 1;
  g->time = 0.0;
  g->dt = cfg->dtinit;    /* driver/definitions.cpp:54 */
  g->dtold = cfg->dtinit;
  if (summary0) orc_field_summary(c, summary0);
}

/* driver/timestep.cpp:37-109 */
int orc_timestep(orc_chunk *c, const orc_config *cfg, orc_globals *g) {
  int fields[15];
  g->dt = G_BIG;
  orc_ideal_gas(c, 0);
  const int ids1[] = {ORC_PRESSURE, ORC_ENERGY0, ORC_DENSITY0, ORC_XVEL0, ORC_YVEL0};
  orc_mask(fields, ids1, 5);
  orc_update_halo(c, fields, 1);
  orc_viscosity(c);
  const int ids2[] = {ORC_VISCOSITY};
  orc_mask(fields, ids2, 1);
  orc_update_halo(c, fields, 1);
  double dtlp;
  int l_control, jldt, kldt, small;
  orc_calc_dt(c, cfg->dtmin, cfg->dtc_safe, cfg->dtu_safe, cfg->dtv_safe, cfg->dtdiv_safe, &dtlp, &l_control, &jldt, &kldt, &small);
  if (dtlp <= g->dt) g->dt = dtlp;
  /* std::min(std::min(dt, dtold * dtrise), dtmax): std::min(a, b) = (b < a) ? b : a */
  double a = g->dt, b = g->dtold * cfg->dtrise;
  double m = (b < a) ? b : a;
  g->dt = (cfg->dtmax < m) ? cfg->dtmax : m;
  int too_small = g->dt < cfg->dtmin;
  g->dtold = g->dt;
  return too_small;
}

/* src/serial/PdV.cpp:92-133 */
template <class T>
static void k_PdV_driver(View<T> *c, int predict, double dt) {
  k_PdV(c, predict, dt);
  if (predict) {
    k_ideal_gas(c, 1);
    int fields[15];
    const int ids[] = {ORC_PRESSURE};
    orc_mask(fields, ids, 1);
    k_update_halo(c, fields, 1);
    k_revert(c);
  }
}

/* driver/advection.cpp:29-105 */
template <class T>
static void k_advection(View<T> *c, int advect_x) {
  int fields[15];
  int direction = advect_x ? 1 : 2;
  const int ids1[] = {ORC_ENERGY1, ORC_DENSITY1, ORC_VOL_FLUX_X, ORC_VOL_FLUX_Y};
  orc_mask(fields, ids1, 4);
  k_update_halo(c, fields, 2);
  k_advec_cell(c, 1, direction);
  const int ids2[] = {ORC_DENSITY1, ORC_ENERGY1, ORC_XVEL1, ORC_YVEL1, ORC_MASS_FLUX_X, ORC_MASS_FLUX_Y};
  orc_mask(fields, ids2, 6);
  k_update_halo(c, fields, 2);
  k_advec_mom(c, 1, 1, direction);
  k_advec_mom(c, 2, 1, direction);
  direction = advect_x ? 2 : 1;
  k_advec_cell(c, 2, direction);
  k_update_halo(c, fields, 2);
  k_advec_mom(c, 1, 2, direction);
  k_advec_mom(c, 2, 2, direction);
}

/* driver/hydro.cpp:57-89 */
int orc_hydro_step(orc_chunk *c, const orc_config *cfg, orc_globals *g) {
  g->step += 1;
  int small = orc_timestep(c, cfg, g);
  if (small) return 1;
  orc_PdV_driver(c, 1, g->dt);
  orc_accelerate(c, g->dt);
  orc_PdV_driver(c, 0, g->dt);
  orc_flux_calc(c, g->dt);
  orc_advection(c, g->advect_x);
  orc_reset_field(c);
  g->advect_x = !g->advect_x;
  g->time += g->dt;
  return 0;
}

/* ---- fixed-dt step ------------------------------------------------------------------------
 * driver/timestep.cpp:47-69 without calc_dt (the step's dt is given), then
 * driver/hydro.cpp:63-89.  The same sequence as orc_hydro_step whenever calc_dt
 * would not limit dt. */
namespace {
template <class T>
void k_step_fixed(View<T> *c, double dt, int advect_x) {
  int fields[15];
  k_ideal_gas(c, 0);
  const int ids1[] = {ORC_PRESSURE, ORC_ENERGY0, ORC_DENSITY0, ORC_XVEL0, ORC_YVEL0};
  orc_mask(fields, ids1, 5);
  k_update_halo(c, fields, 1);
  k_viscosity(c);
  const int ids2[] = {ORC_VISCOSITY};
  orc_mask(fields, ids2, 1);
  k_update_halo(c, fields, 1);
  k_PdV_driver(c, 1, dt);
  k_accelerate(c, dt);
  k_PdV_driver(c, 0, dt);
  k_flux_calc(c, dt);
  k_advection(c, advect_x);
  k_reset_field(c);
}
}  // namespace

/* ---- C entry points: double (the oracle) ------------------------------------------------ */
void orc_ideal_gas(orc_chunk *c, int predict) { k_ideal_gas(V(c), predict); }
void orc_viscosity(orc_chunk *c) { k_viscosity(V(c)); }
void orc_PdV(orc_chunk *c, int predict, double dt) { k_PdV(V(c), predict, dt); }
void orc_revert(orc_chunk *c) { k_revert(V(c)); }
void orc_accelerate(orc_chunk *c, double dt) { k_accelerate(V(c), dt); }
void orc_flux_calc(orc_chunk *c, double dt) { k_flux_calc(V(c), dt); }
void orc_advec_cell(orc_chunk *c, int sweep_number, int dir) { k_advec_cell(V(c), sweep_number, dir); }
void orc_advec_mom(orc_chunk *c, int which_vel, int sweep_number, int direction) { k_advec_mom(V(c), which_vel, sweep_number, direction); }
void orc_reset_field(orc_chunk *c) { k_reset_field(V(c)); }
void orc_update_halo(orc_chunk *c, const int fields[15], int depth) { k_update_halo(V(c), fields, depth); }
void orc_update_halo_ext(orc_chunk *c, const int fields[15], int depth, const int ext[4]) { k_update_halo_ext(V(c), fields, depth, ext); }
void orc_tile_halo(orc_chunk *c, orc_chunk *neighbour, int face, const int fields[15], int depth) {
  k_tile_halo(V(c), V(neighbour), face, fields, depth);
}
void orc_pack_message(orc_chunk *c, double *field, int field_type, int face, int depth, double *buf, long offset) {
  k_pack(V(c), field, field_type, face, depth, buf, offset, 0);
}
void orc_unpack_message(orc_chunk *c, double *field, int field_type, int face, int depth, const double *buf, long offset) {
  k_pack(V(c), field, field_type, face, depth, const_cast<double *>(buf), offset, 1);
}
void orc_field_summary_kernel(orc_chunk *c, double out[5]) { k_field_summary_kernel(V(c), out); }
void orc_PdV_driver(orc_chunk *c, int predict, double dt) { k_PdV_driver(V(c), predict, dt); }
void orc_advection(orc_chunk *c, int advect_x) { k_advection(V(c), advect_x); }
void orc_step_fixed(orc_chunk *c, double dt, int advect_x) { k_step_fixed(V(c), dt, advect_x); }

/* ---- C entry points: the tangent-linear model ------------------------------------------- */
int orc_tl_alloc(orc_tl_chunk *t, int nx, int ny) {
  memset(t, 0, sizeof(*t));
  t->nx = nx;
  t->ny = ny;
  t->px = nx + 5;
  const size_t n2 = (size_t)(nx + 5) * (size_t)(ny + 5);
  for (int a = 0; a < ORC_NUM_2D; ++a) {
    t->f[a] = (orc_dual *)calloc(n2, sizeof(orc_dual));
    if (!t->f[a]) return -1;
  }
  for (int a = 0; a < ORC_NUM_1D; ++a) {
    const int n = (a == ORC_CELLX || a == ORC_VERTEXX || a == ORC_CELLDX || a == ORC_VERTEXDX) ? nx + 5 : ny + 5;
    t->m[a] = (double *)calloc((size_t)n, sizeof(double));
    if (!t->m[a]) return -1;
  }
  return 0;
}

void orc_tl_free(orc_tl_chunk *t) {
  for (int a = 0; a < ORC_NUM_2D; ++a) free(t->f[a]);
  for (int a = 0; a < ORC_NUM_1D; ++a) free(t->m[a]);
  memset(t, 0, sizeof(*t));
}

void orc_tl_load(orc_tl_chunk *t, const orc_chunk *c) {
  const size_t n2 = (size_t)(c->nx + 5) * (size_t)(c->ny + 5);
  for (int a = 0; a < ORC_NUM_2D; ++a)
    for (size_t k = 0; k < n2; ++k) {
      t->f[a][k].v = c->f[a][k];
      t->f[a][k].d = 0.0;
    }
  for (int a = 0; a < ORC_NUM_1D; ++a) {
    const int n = (a == ORC_CELLX || a == ORC_VERTEXX || a == ORC_CELLDX || a == ORC_VERTEXDX) ? c->nx + 5 : c->ny + 5;
    memcpy(t->m[a], c->m[a], (size_t)n * sizeof(double));
  }
}

void orc_tl_ideal_gas(orc_tl_chunk *t, int predict) { k_ideal_gas(V(t), predict); }
void orc_tl_viscosity(orc_tl_chunk *t) { k_viscosity(V(t)); }
void orc_tl_PdV(orc_tl_chunk *t, int predict, double dt) { k_PdV(V(t), predict, dt); }
void orc_tl_revert(orc_tl_chunk *t) { k_revert(V(t)); }
void orc_tl_accelerate(orc_tl_chunk *t, double dt) { k_accelerate(V(t), dt); }
void orc_tl_flux_calc(orc_tl_chunk *t, double dt) { k_flux_calc(V(t), dt); }
void orc_tl_advec_cell(orc_tl_chunk *t, int sweep_number, int dir) { k_advec_cell(V(t), sweep_number, dir); }
void orc_tl_advec_mom(orc_tl_chunk *t, int which_vel, int sweep_number, int direction) {
  k_advec_mom(V(t), which_vel, sweep_number, direction);
}
void orc_tl_reset_field(orc_tl_chunk *t) { k_reset_field(V(t)); }
void orc_tl_update_halo(orc_tl_chunk *t, const int fields[15], int depth) { k_update_halo(V(t), fields, depth); }
void orc_tl_field_summary_kernel(orc_tl_chunk *t, orc_dual out[5]) { k_field_summary_kernel(V(t), reinterpret_cast<Dual *>(out)); }
void orc_tl_field_summary(orc_tl_chunk *t, orc_dual out[5]) {
  k_ideal_gas(V(t), 0);
  k_field_summary_kernel(V(t), reinterpret_cast<Dual *>(out));
}
void orc_tl_PdV_driver(orc_tl_chunk *t, int predict, double dt) { k_PdV_driver(V(t), predict, dt); }
void orc_tl_advection(orc_tl_chunk *t, int advect_x) { k_advection(V(t), advect_x); }
void orc_tl_step_fixed(orc_tl_chunk *t, double dt, int advect_x) { k_step_fixed(V(t), dt, advect_x); }
void orc_tl_set_kink_eps(double eps) {
  g_kink_eps = eps;
  g_kinks = 0;
}
long long orc_tl_kinks(void) { return g_kinks; }
