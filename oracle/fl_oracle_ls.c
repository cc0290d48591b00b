/*
 * fl_oracle_ls.c -- CPU restatement of fluidity's level-set / ghost-fluid step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker for the HIP path); see fl_oracle.h.
 * Build with -ffp-contract=off, no -ffast-math.
 *
 * Reference semantics restated here (paths relative to include/fluidity/):
 *  - interface predicates ............. levelset/levelset.hpp:166-217
 *  - level-set speed u.n ............... levelset/cuda/velocity_initialization.cuh:108-143,
 *                                        iterator/multidim_iterator.hpp:291-306 (norm),
 *                                        math/math.hpp:83-104 (dot)
 *  - WENO-HJ5 one-sided derivatives .... scheme/stencils/weno_hj_5.hpp:51-144
 *  - Godunov upwind |grad phi| ......... scheme/schemes/godunov_upwind.hpp:94-166
 *  - L = V |grad phi| .................. scheme/evaluators/levelset_hamiltonian.hpp:63-98
 *  - SSP-RK3 ........................... scheme/updaters/runge_kutta_3.hpp:117-148
 *  - projection (two materials) ........ levelset/cuda/levelset_projection.cuh:54-66
 *  - projection (n > 2 materials) ...... levelset/cuda/levelset_projection.cuh:87-118
 *  - eikonal sources, Godunov update ... scheme/cuda/fast_iterative_method.cuh:48-85,
 *                                        scheme/schemes/godunov_upwind_scheme.hpp:37-53
 *  - Riemann ghost fluid ............... ghost_fluid/riemann_ghost_fluid.hpp:50-195,
 *                                        flux_method/flux_hllc.hpp:301-335, 420-571,
 *                                        math/interp.hpp:108-165
 *  - ghost-cell extrapolation .......... ghost_fluid/ghost_fluid_extrapolator.hpp:54-202
 *  - multimaterial step ................ simulator/multimaterial_simulator.hpp:227-303, 367-483
 *
 * Decisions (the reference path is work in progress and undefined in places):
 *  L1. std::pow(x, 2) in WENO-HJ5 is x*x (equal whenever pow is correctly
 *      rounded, which glibc's is for this case).
 *  L2. RK3 stages are global: every stage sees the previous stage everywhere
 *      (the reference keeps stale LDS halos between stages).
 *  L3. Level-set ghost cells mirror the interior (transmissive copy); a zero
 *      gradient gives a zero normal (the reference divides 0 by 0).
 *  L4. Eikonal reinitialisation: sources as in fast_iterative_init with
 *      |phi| as their distance; every other cell solves the Godunov update
 *      by Jacobi iteration to the exact fixed point, restricted to a band of
 *      `band` cells (an extension); cells not reached are set to
 *      +-(band + 1) dh; signs restored from the input.  A neighbour on the
 *      other side of the interface (other sign of phi) enters the update
 *      with its distance negated, which is the offset the reference's
 *      grad_diff pass subtracts afterwards (fast_iterative_method.cuh:88-200,
 *      a racy global shift there); with it a linear phi reinitialises to its
 *      exact distance.
 *  L5. Riemann ghost fluid: states are interpolated bilinearly at -+1.5 dh
 *      along the normal (the reference's 2D interp writes through references
 *      to the centre cell), rotated into the normal frame through the
 *      primitive state (the reference's set_velocity leaves the energy of the
 *      discarded tangential velocity in E), and the star state is stored at
 *      the interfacial cell itself; the extrapolation then carries it into
 *      the ghost band (the reference writes it to a neighbour cell, racing
 *      with other interfacial cells).
 *  L6. Extrapolation: Jacobi iteration, a fixed number of sweeps, over the
 *      ghost band 0 < phi <= 6 dh / sqrt(2) (the reference iterates each
 *      cell in place until its own change is below 1e-8, over dh <= phi).
 *  L7. Multimaterial step order: ghost states of every material are computed
 *      from the same data before any extrapolation; dt from the maximum
 *      wavespeed over every material's whole grid (update_time_delta); each
 *      material's split solve; speed from the old level sets and new states;
 *      advection of each level set; projection; reinitialisation after
 *      every reinit_every-th step (the reference's every step is 1).
 *  L8. Level set m is advected with V_m = u . n_m, u the velocity of the
 *      material present and n_m the normal of level set m.  The reference
 *      advects every level set with u . n of the material present, which
 *      flips the sign of the speed across the interface for every level set
 *      but the present material's and holds the interface in place.
 *  L9. n > 2 materials: material m's ghost-fluid star state at a cell comes
 *      from the last other material (in material order) whose level set is
 *      outside-interfacial or on the boundary there, as the reference's
 *      loop over the other materials overwrites earlier ones
 *      (riemann_ghost_fluid.hpp:56-190); the level sets are projected by the
 *      reference's general rule (the average of the two smallest values is
 *      subtracted from every level set); two materials keep the two-level-set
 *      projection, as the reference specialises it.
 */
#include "fl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline int signum_d(double x) { return (0.0 < x) - (x < 0.0); }

typedef struct {
  int64_t nx, ny;
  int nd;
} lsgrid_t;

/* Mirrored ghost index, clamped into the grid. */
static inline int64_t mirror(int64_t i, int64_t n) {
  if (i < 0) i = -i - 1;
  if (i >= n) i = 2 * n - 1 - i;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

/* phi at (x, y) with mirrored ghosts (decision L3). */
static inline double phi_at(const double* phi, const lsgrid_t* g, int64_t x, int64_t y) {
  const int64_t xi = mirror(x, g->nx);
  const int64_t yi = (g->nd > 1) ? mirror(y, g->ny) : 0;
  return phi[yi * g->nx + xi];
}

/* ---- predicates (levelset/levelset.hpp:166-217) ---------------------------- */

uint8_t orc_ls_mask_value(double p, double dh) {
  const int inside = (p <= 0.0) || (p <= -0.0);
  const int outside = p > 0.0;
  const int on_boundary = (p == 0.0) || (p == -0.0);
  const int in_ifc = (inside || on_boundary) && (p >= -dh);
  const int out_ifc = outside && (p < fabs(dh));
  return (uint8_t)(inside | (outside << 1) | (on_boundary << 2) | (in_ifc << 3) | (out_ifc << 4));
}

void orc_ls_masks(const double* phi, const int64_t* n, int nd, double dh, uint8_t* masks) {
  const int64_t total = n[0] * (nd > 1 ? n[1] : 1);
  for (int64_t i = 0; i < total; ++i) masks[i] = orc_ls_mask_value(phi[i], dh);
}

/* ---- normal (MultidimIterator::norm, iterator/multidim_iterator.hpp:291-306) - */

static void ls_norm(const double* phi, const lsgrid_t* g, int64_t x, int64_t y, double dh,
                    double* nrm) {
  const double fact = 0.5 / dh;
  double mag = 0.0;
  for (int d = 0; d < g->nd; ++d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    nrm[d] = fact * (phi_at(phi, g, x + dx, y + dy) - phi_at(phi, g, x - dx, y - dy));
    mag += nrm[d] * nrm[d];
  }
  const double s = sqrt(mag);
  for (int d = 0; d < g->nd; ++d) nrm[d] = (mag == 0.0) ? 0.0 : nrm[d] / s;
}

/* ---- level-set speed (levelset/cuda/velocity_initialization.cuh:108-143) ----
 * target < 0: the reference's field, u.n with the normal of the level set of
 * the material present; target = m: the normal of level set m (decision L8). */

void orc_ls_velocity(int nmat, const double* const* phis, const double* const* states,
                     const int64_t* n, int nd, double dh, int target, double* V) {
  lsgrid_t g = {n[0], nd > 1 ? n[1] : 1, nd};
  const int S = 2 + nd;
  for (int64_t y = 0; y < g.ny; ++y) {
    for (int64_t x = 0; x < g.nx; ++x) {
      const int64_t i = y * g.nx + x;
      double v = 0.0;
      for (int m = 0; m < nmat; ++m) {
        const double p = phis[m][i];
        if ((p <= 0.0) || (p <= -0.0)) {
          double nrm[2];
          ls_norm(phis[target < 0 ? m : target], &g, x, y, dh, nrm);
          const double* U = states[m] + i * S;
          double sum = 0.0;
          for (int d = 0; d < nd; ++d) sum += (U[2 + d] / U[0]) * nrm[d];
          v = sum;
          break;
        }
      }
      V[i] = v;
    }
  }
}

/* ---- WENO-HJ5 (scheme/stencils/weno_hj_5.hpp:51-144) ------------------------ */

static double weno_diff(double d1, double d2, double d3, double d4, double d5) {
  const double phi_1 = 2.0 * d1 - 7.0 * d2 + 11.0 * d3;
  const double phi_2 = -d2 + 5.0 * d3 + 2.0 * d4;
  const double phi_3 = 2.0 * d3 + 5.0 * d4 - d5;
  const double f1 = 13.0 / 12.0;
  const double f2 = 1.0 / 4.0;
  const double s11 = d1 - 2.0 * d2 + d3, s12 = d1 - 4.0 * d2 + 3.0 * d3;
  const double s21 = d2 - 2 * d3 + d4, s22 = d2 - d4;
  const double s31 = d3 - 2.0 * d4 + d5, s32 = 3.0 * d3 - 4.0 * d4 + d5;
  double alpha_1 = f1 * (s11 * s11) + f2 * (s12 * s12);
  double alpha_2 = f1 * (s21 * s21) + f2 * (s22 * s22);
  double alpha_3 = f1 * (s31 * s31) + f2 * (s32 * s32);
  const double e =
      10.0e-6 * smax(smax(smax(smax(d1 * d1, d2 * d2), d3 * d3), d4 * d4), d5 * d5) + 10.0e-99;
  const double a1e = alpha_1 + e, a2e = alpha_2 + e, a3e = alpha_3 + e;
  alpha_1 = 0.1 / (a1e * a1e);
  alpha_2 = 0.6 / (a2e * a2e);
  alpha_3 = 0.3 / (a3e * a3e);
  const double w_denom = (alpha_1 + alpha_2 + alpha_3) * 6.0;
  return (alpha_1 * phi_1 + alpha_2 * phi_2 + alpha_3 * phi_3) / w_denom;
}

/* phi at offsets -3..+3 along d about (x, y): f[k+3] = phi(k). */
static void line7(const double* phi, const lsgrid_t* g, int64_t x, int64_t y, int d, double* f) {
  const int64_t dx = (d == 0), dy = (d == 1);
  for (int k = -3; k <= 3; ++k) f[k + 3] = phi_at(phi, g, x + k * dx, y + k * dy);
}

static double weno_backward(const double* f, double h) {
  const double h1 = 1.0 / h;
  const double d1 = h1 * (f[1] - f[0]);
  const double d2 = h1 * (f[2] - f[1]);
  const double d3 = h1 * (f[3] - f[2]);
  const double d4 = h1 * (f[4] - f[3]);
  const double d5 = h1 * (f[5] - f[4]);
  return weno_diff(d1, d2, d3, d4, d5);
}

static double weno_forward(const double* f, double h) {
  const double h1 = 1.0 / h;
  const double d1 = h1 * (f[6] - f[5]);
  const double d2 = h1 * (f[5] - f[4]);
  const double d3 = h1 * (f[4] - f[3]);
  const double d4 = h1 * (f[3] - f[2]);
  const double d5 = h1 * (f[2] - f[1]);
  return weno_diff(d1, d2, d3, d4, d5);
}

/* The two one-sided WENO-HJ5 derivatives (D-phi, D+phi) at cell i of a 1D
 * line with mirrored ghosts: the stencil on its own, for the pins of its
 * weights, smoothness indicators and epsilon. */
void orc_ls_weno(const double* phi, int64_t n, double dh, int64_t i, double* dminus,
                 double* dplus) {
  lsgrid_t g = {n, 1, 1};
  double f[7];
  line7(phi, &g, i, 0, 0, f);
  *dminus = weno_backward(f, dh);
  *dplus = weno_forward(f, dh);
}

/* L = V * H (levelset_hamiltonian.hpp:63-98, godunov_upwind.hpp:94-166). */
static double ls_rhs(const double* phi, const lsgrid_t* g, int64_t x, int64_t y, double v,
                     double dh) {
  double result = 0.0;
  for (int d = 0; d < g->nd; ++d) {
    double f[7];
    line7(phi, g, x, y, d, f);
    const double bd = weno_backward(f, dh);
    const double fd = weno_forward(f, dh);
    double back, fwrd;
    if (v <= 0.0) {            /* GodunovUpwind::forward */
      back = smax(bd, 0.0);
      fwrd = smin(fd, 0.0);
    } else {                   /* GodunovUpwind::backward */
      back = smin(bd, 0.0);
      fwrd = smax(fd, 0.0);
    }
    result += smax(back * back, fwrd * fwrd);
  }
  return v * sqrt(result);
}

/* RungeKutta3::update_impl (scheme/updaters/runge_kutta_3.hpp:117-148), with
 * global stages (decision L2). */
void orc_ls_advect(const double* phi, const double* V, const int64_t* n, int nd, double dt,
                   double dh, double* out) {
  lsgrid_t g = {n[0], nd > 1 ? n[1] : 1, nd};
  const int64_t total = g.nx * g.ny;
  double* s1 = (double*)malloc(sizeof(double) * total);
  double* s2 = (double*)malloc(sizeof(double) * total);
#pragma omp parallel for
  for (int64_t i = 0; i < total; ++i) {
    const int64_t x = i % g.nx, y = i / g.nx;
    s1[i] = phi[i] - dt * ls_rhs(phi, &g, x, y, V[i], dh);
  }
#pragma omp parallel for
  for (int64_t i = 0; i < total; ++i) {
    const int64_t x = i % g.nx, y = i / g.nx;
    s2[i] = 0.75 * phi[i] + 0.25 * (s1[i] - dt * ls_rhs(s1, &g, x, y, V[i], dh));
  }
  const double fact_13 = 1.0 / 3.0;
  const double fact_23 = 2.0 * fact_13;
#pragma omp parallel for
  for (int64_t i = 0; i < total; ++i) {
    const int64_t x = i % g.nx, y = i / g.nx;
    const double s3 = s2[i] - dt * ls_rhs(s2, &g, x, y, V[i], dh);
    out[i] = (fact_13 * phi[i]) + (fact_23 * s3);
  }
  free(s1);
  free(s2);
}

/* project_impl for two level sets (levelset/cuda/levelset_projection.cuh:54-66). */
void orc_ls_project2(double* phi0, double* phi1, int64_t count) {
  for (int64_t i = 0; i < count; ++i) {
    const double avg = 0.5 * (phi0[i] + phi1[i]);
    phi0[i] -= avg;
    phi1[i] = -phi0[i];
  }
}

/* levelset/cuda/levelset_projection.cuh:87-118 (n > 2 level sets): the
 * average of the two smallest values at a cell is subtracted from every
 * level set there.  The running pair starts at (max - 1, max) as there. */
void orc_ls_project_n(int nmat, double* const* phis, int64_t count) {
  for (int64_t i = 0; i < count; ++i) {
    double smallest = DBL_MAX - 1.0, next_smallest = DBL_MAX;
    for (int m = 0; m < nmat; ++m) {
      const double v = phis[m][i];
      if (v < smallest) {
        next_smallest = smallest;
        smallest = v;
      } else if (v < next_smallest) {
        next_smallest = v;
      }
    }
    const double avg = 0.5 * (smallest + next_smallest);
    for (int m = 0; m < nmat; ++m) phis[m][i] -= avg;
  }
}

/* ---- eikonal reinitialisation (decision L4) --------------------------------- */

/* GodunovUpwindScheme::solve_impl (scheme/schemes/godunov_upwind_scheme.hpp:37-53)
 * with f = 1; out-of-domain neighbours are DBL_MAX
 * (scheme/cuda/fast_iterative_method.cuh:277-287). */
/* Distance of neighbour j as seen from a cell of sign sc: negated across the
 * interface (decision L4) unless not yet reached. */
static inline double eik_nbr(const double* d, const int8_t* sg, int64_t j, int sc) {
  const double v = d[j];
  return (sg[j] == sc || v == DBL_MAX) ? v : -v;
}

static double eik_update(const double* d, const int8_t* sg, const lsgrid_t* g, int64_t x,
                         int64_t y, double dh) {
  const double fh = 1.0 * dh;
  const int64_t i = y * g->nx + x;
  const int sc = sg[i];
  const double xm = x > 0 ? eik_nbr(d, sg, i - 1, sc) : DBL_MAX;
  const double xp = x < g->nx - 1 ? eik_nbr(d, sg, i + 1, sc) : DBL_MAX;
  const double a = smin(xm, xp);
  if (g->nd == 1) return a + fh;
  const double ym = y > 0 ? eik_nbr(d, sg, i - g->nx, sc) : DBL_MAX;
  const double yp = y < g->ny - 1 ? eik_nbr(d, sg, i + g->nx, sc) : DBL_MAX;
  const double b = smin(ym, yp);
  const double amb = a - b;
  return fabs(amb) >= fh ? smin(a, b) + fh : 0.5 * (a + b + sqrt(2.0 * fh * fh - (amb * amb)));
}

int orc_ls_reinit(const double* phi, const int64_t* n, int nd, double dh, int band,
                  double* out, uint8_t* band_mask, int64_t* iterations) {
  lsgrid_t g = {n[0], nd > 1 ? n[1] : 1, nd};
  const int64_t total = g.nx * g.ny;
  double* d = (double*)malloc(sizeof(double) * total);
  double* e = (double*)malloc(sizeof(double) * total);
  uint8_t* src = (uint8_t*)malloc(total);
  int8_t* sg = (int8_t*)malloc(total);
  if (!d || !e || !src || !sg) return -3;
  const double limit = band * dh;
  /* sources: fast_iterative_init (fast_iterative_method.cuh:48-85) */
  for (int64_t y = 0; y < g.ny; ++y) {
    for (int64_t x = 0; x < g.nx; ++x) {
      const int64_t i = y * g.nx + x;
      const double c = phi[i];
      int source = 0;
      for (int dd = 0; dd < nd && !source; ++dd) {
        const int64_t dx = (dd == 0), dy = (dd == 1);
        for (int s = 0; s < 2 && !source; ++s) {
          const int64_t ox = x + (s == 0 ? dx : -dx), oy = y + (s == 0 ? dy : -dy);
          if (ox < 0 || ox >= g.nx || oy < 0 || oy >= g.ny) continue;
          const double o = phi[oy * g.nx + ox];
          if (signum_d(c) != signum_d(o) && fabs(c) < fabs(o)) source = 1;
        }
      }
      src[i] = (uint8_t)source;
      sg[i] = (int8_t)signum_d(c);
      d[i] = source ? fabs(c) : DBL_MAX;
    }
  }
  int64_t it = 0;
  for (;;) {
    int changed = 0;
#pragma omp parallel for reduction(| : changed)
    for (int64_t i = 0; i < total; ++i) {
      e[i] = d[i];
      if (src[i]) continue;
      const double q = eik_update(d, sg, &g, i % g.nx, i / g.nx, dh);
      if (q < d[i] && q <= limit) {
        e[i] = q;
        changed = 1;
      }
    }
    double* t = d; d = e; e = t;
    ++it;
    if (!changed) break;
  }
  const double clamp = (band + 1) * dh;
  for (int64_t i = 0; i < total; ++i) {
    const int in_band = d[i] <= limit;
    const double dist = in_band ? d[i] : clamp;
    out[i] = (double)signum_d(phi[i]) * dist;
    if (band_mask) band_mask[i] = (uint8_t)in_band;
  }
  if (iterations) *iterations = it;
  free(d);
  free(e);
  free(src);
  free(sg);
  return 0;
}

/* ---- Riemann ghost fluid (decision L5) --------------------------------------- */

static void state_at(const double* U, const lsgrid_t* g, int64_t x, int64_t y, double* u) {
  const int S = 2 + g->nd;
  const int64_t xi = mirror(x, g->nx);
  const int64_t yi = (g->nd > 1) ? mirror(y, g->ny) : 0;
  memcpy(u, U + (yi * g->nx + xi) * S, sizeof(double) * S);
}

/* math::interp (math/interp.hpp:56-165) as a proper (bi)linear interpolation. */
static void interp_state(const double* U, const lsgrid_t* g, int64_t x, int64_t y,
                         const double* w, double* out) {
  const int S = 2 + g->nd;
  int64_t o0[2] = {0, 0}, o1[2] = {0, 0};
  double fr[2] = {0.0, 0.0};
  for (int d = 0; d < g->nd; ++d) {
    const int sign = signum_d(w[d]);
    const double abs_w = fabs(w[d]);
    const double abs_off = floor(abs_w);
    o0[d] = (int64_t)(sign * abs_off);
    o1[d] = o0[d] + sign;
    fr[d] = abs_w - abs_off;
  }
  double a[8], b[8];
  if (g->nd == 1) {
    state_at(U, g, x + o0[0], y, a);
    state_at(U, g, x + o1[0], y, b);
    for (int k = 0; k < S; ++k) out[k] = a[k] * (1.0 - fr[0]) + b[k] * fr[0];
    return;
  }
  double c[8], e[8];
  state_at(U, g, x + o0[0], y + o0[1], a);
  state_at(U, g, x + o1[0], y + o0[1], b);
  state_at(U, g, x + o0[0], y + o1[1], c);
  state_at(U, g, x + o1[0], y + o1[1], e);
  const double factor = 1.0 - fr[0];
  for (int k = 0; k < S; ++k) {
    const double close = a[k] * factor + b[k] * fr[0];
    const double far = c[k] * factor + e[k] * fr[0];
    out[k] = close * (1.0 - fr[1]) + far * fr[1];
  }
}

/* Hllc::solve_star_conservative_left (flux_method/flux_hllc.hpp:301-335). */
static void hllc_star_left(const double* ul, const double* ur, int nd, const orc_scheme* el,
                           const orc_scheme* er, double* out) {
  const int S = 2 + nd;
  const double al = orc_sound_speed(ul, nd, 0, el), ar = orc_sound_speed(ur, nd, 0, er);
  const double pl = orc_pressure(ul, nd, 0, el), pr = orc_pressure(ur, nd, 0, er);
  const double vl = ul[2] / ul[0], vr = ur[2] / ur[0];
  const double p_pvrs = 0.5 * ((pl + pr) - (0.25 * (vr - vl) * (ur[0] + ul[0]) * (al + ar)));
  /* a corrected stiffened gas on both sides bounds p* by -min(p_inf) */
  const double p_star = (el->eos == ORC_EOS_STIFF_CORRECTED && er->eos == ORC_EOS_STIFF_CORRECTED)
                            ? smax(-smin(el->p_inf, er->p_inf), p_pvrs)
                            : smax(0.0, p_pvrs);
  const double ql = (el->gamma + 1.0) / (2.0 * el->gamma);
  const double qr = (er->gamma + 1.0) / (2.0 * er->gamma);
  /* the corrected stiffened gas takes pressure ratios of p + p_inf */
  const double rl = el->eos == ORC_EOS_STIFF_CORRECTED ? (p_star + el->p_inf) / (pl + el->p_inf)
                                                       : p_star / pl;
  const double rr = er->eos == ORC_EOS_STIFF_CORRECTED ? (p_star + er->p_inf) / (pr + er->p_inf)
                                                       : p_star / pr;
  const double wsl = (p_star <= pl) ? vl + -al : vl + -al * sqrt(1.0 + ql * (rl - 1.0));
  const double wsr = (p_star <= pr) ? vr + ar : vr + ar * sqrt(1.0 + qr * (rr - 1.0));
  const double factorl = ul[0] * (wsl - vl);
  const double factorr = ur[0] * (wsr - vr);
  const double ws_star = (pr - pl + vl * factorl - vr * factorr) / (factorl - factorr);
  const double state_factor = wsl - vl;
  const double scale_factor = state_factor / (wsl - ws_star);
  for (int k = 0; k < S; ++k) out[k] = ul[k] * scale_factor;
  out[2] = ws_star * ul[0] * scale_factor;
  out[1] = scale_factor * (ul[1] + ul[0] * (ws_star - vl) * (ws_star + pl / (ul[0] * state_factor)));
}

/* The star state of material m at cell (x, y) against material o (L5), or 0
 * where the cell does not qualify: m inside-interfacial with no axis
 * neighbour on its boundary, o outside-interfacial or on its boundary. */
static int gfm_star_cell(const double* phi_m, const double* phi_o, const double* U_m,
                         const double* U_o, const orc_scheme* eos_m, const orc_scheme* eos_o,
                         const lsgrid_t* gp, double dh, int64_t x, int64_t y, double* out) {
  const lsgrid_t g = *gp;
  const int nd = g.nd;
  const int64_t i = y * g.nx + x;
  const double p = phi_m[i];
  const uint8_t mk = orc_ls_mask_value(p, dh);
  if (!(mk & 8)) return 0;                       /* inside_interfacial_cell */
  int nb_on_boundary = 0;                         /* neighbour_not_on_boundary */
  for (int d = 0; d < nd; ++d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    const double pm = phi_at(phi_m, &g, x - dx, y - dy), pp = phi_at(phi_m, &g, x + dx, y + dy);
    if (pm == 0.0 || pp == 0.0) nb_on_boundary = 1;
  }
  if (nb_on_boundary) return 0;
  const uint8_t mo = orc_ls_mask_value(phi_o[i], dh);
  if (!((mo & 16) || (mo & 4))) return 0;         /* out_ifc_or_boundary */
  double nrm[2] = {0.0, 0.0};
  ls_norm(phi_m, &g, x, y, dh, nrm);
  const double abs_phi = fabs(p);
  const double dh_three_two = 1.5 * dh;
  double w[2] = {0.0, 0.0}, wl[8], wr[8];
  for (int d = 0; d < nd; ++d) w[d] = nrm[d] * (abs_phi - dh_three_two) / dh;
  interp_state(U_m, &g, x, y, w, wl);
  for (int d = 0; d < nd; ++d) w[d] = nrm[d] * (abs_phi + dh_three_two) / dh;
  interp_state(U_o, &g, x, y, w, wr);
  /* normal frame through the primitive state */
  double Wl[8], Wr[8], vl[2];
  orc_cons_to_prim(wl, nd, 0, eos_m, Wl);
  orc_cons_to_prim(wr, nd, 0, eos_o, Wr);
  double vl_n = 0.0, vr_n = 0.0;
  for (int d = 0; d < nd; ++d) {
    vl[d] = Wl[2 + d];
    vl_n += Wl[2 + d] * nrm[d];
    vr_n += Wr[2 + d] * nrm[d];
  }
  Wl[2] = vl_n;
  Wr[2] = vr_n;
  for (int d = 1; d < nd; ++d) { Wl[2 + d] = 0.0; Wr[2 + d] = 0.0; }
  orc_prim_to_cons(Wl, nd, 0, eos_m, wl);
  orc_prim_to_cons(Wr, nd, 0, eos_o, wr);
  double st[8], Ws[8];
  hllc_star_left(wl, wr, nd, eos_m, eos_o, st);
  orc_cons_to_prim(st, nd, 0, eos_m, Ws);
  const double vs_n = Ws[2];
  for (int d = 0; d < nd; ++d) Ws[2 + d] = vl[d] + (vs_n - vl_n) * nrm[d];
  orc_prim_to_cons(Ws, nd, 0, eos_m, out);
  return 1;
}

void orc_gfm_star(const double* phi_m, const double* phi_o, const double* U_m,
                  const double* U_o, const orc_scheme* eos_m, const orc_scheme* eos_o,
                  const int64_t* n, int nd, double dh, double* star, uint8_t* has_star) {
  lsgrid_t g = {n[0], nd > 1 ? n[1] : 1, nd};
  const int S = 2 + nd;
#pragma omp parallel for
  for (int64_t i = 0; i < g.nx * g.ny; ++i) {
    const int64_t x = i % g.nx, y = i / g.nx;
    has_star[i] = (uint8_t)gfm_star_cell(phi_m, phi_o, U_m, U_o, eos_m, eos_o, &g, dh, x, y,
                                         star + i * S);
  }
}

/* Material m's star states among nmat materials (L9): every other material
 * in order, a later qualifying one overwriting an earlier one. */
void orc_gfm_star_n(int m, int nmat, const double* const* phis, const double* const* Us,
                    const orc_scheme* const* eos, const int64_t* n, int nd, double dh,
                    double* star, uint8_t* has_star) {
  lsgrid_t g = {n[0], nd > 1 ? n[1] : 1, nd};
  const int S = 2 + nd;
#pragma omp parallel for
  for (int64_t i = 0; i < g.nx * g.ny; ++i) {
    const int64_t x = i % g.nx, y = i / g.nx;
    has_star[i] = 0;
    double st[8];
    for (int o = 0; o < nmat; ++o) {
      if (o == m) continue;
      if (gfm_star_cell(phis[m], phis[o], Us[m], Us[o], eos[m], eos[o], &g, dh, x, y, st)) {
        memcpy(star + i * S, st, sizeof(double) * S);
        has_star[i] = 1;
      }
    }
  }
}

/* GhostCellExtrapolator (ghost_fluid/ghost_fluid_extrapolator.hpp:54-202),
 * Jacobi form (decision L6).  Cells with phi_m <= 0 are sources: their star
 * state where has_star, else their own state.  Ghost cells
 * 0 < phi_m <= 6 dh / sqrt(2) take the |n|-weighted average of their upwind
 * neighbours (towards smaller phi).  The band includes the first ghost layer
 * 0 < phi_m < dh, which the reference fills with the star state directly. */
void orc_gfm_extrapolate(const double* phi_m, double* U_m, const double* star,
                         const uint8_t* has_star, const int64_t* n, int nd, double dh,
                         int sweeps) {
  lsgrid_t g = {n[0], nd > 1 ? n[1] : 1, nd};
  const int S = 2 + nd;
  const int64_t total = g.nx * g.ny;
  const double upper = 6.0 * dh * (1.0 / sqrt(2.0));
  double* A = (double*)malloc(sizeof(double) * total * S);
  double* B = (double*)malloc(sizeof(double) * total * S);
  memcpy(A, U_m, sizeof(double) * total * S);
  for (int64_t i = 0; i < total; ++i)
    if (has_star[i]) memcpy(A + i * S, star + i * S, sizeof(double) * S);
  for (int sw = 0; sw < sweeps; ++sw) {
    memcpy(B, A, sizeof(double) * total * S);
#pragma omp parallel for
    for (int64_t i = 0; i < total; ++i) {
      const int64_t x = i % g.nx, y = i / g.nx;
      const double p = phi_m[i];
      if (!(p > 0.0 && p <= upper)) continue;
      double nrm[2];
      ls_norm(phi_m, &g, x, y, dh, nrm);
      const int sg = signum_d(p);
      double l1 = 0.0;
      for (int d = 0; d < nd; ++d) {
        nrm[d] = -sg * nrm[d];
        l1 += fabs(nrm[d]);
      }
      double acc[8];
      for (int k = 0; k < S; ++k) acc[k] = 0.0;
      for (int d = 0; d < nd; ++d) {
        const int64_t dx = (d == 0), dy = (d == 1);
        const int dir = phi_at(phi_m, &g, x - dx, y - dy) < phi_at(phi_m, &g, x + dx, y + dy) ? -1 : 1;
        const int64_t ux = mirror(x + dir * dx, g.nx), uy = (nd > 1) ? mirror(y + dir * dy, g.ny) : 0;
        const double* src = A + (uy * g.nx + ux) * S;
        const double wgt = l1 > 1e-2 ? fabs(nrm[d]) : 0.5;
        for (int k = 0; k < S; ++k) acc[k] += src[k] * wgt;
      }
      for (int k = 0; k < S; ++k) B[i * S + k] = (l1 > 1e-2) ? acc[k] / l1 : acc[k];
    }
    double* t = A; A = B; B = t;
  }
  /* write back only the ghost band; real cells keep their own state */
  for (int64_t i = 0; i < total; ++i) {
    const double p = phi_m[i];
    if (p > 0.0 && p <= upper) memcpy(U_m + i * S, A + i * S, sizeof(double) * S);
  }
  free(A);
  free(B);
}

/* combine_into_first (material/cuda/combine_materials.cuh:39-63): the first
 * material whose phi <= 0 provides the state, else the second. */
void orc_mm_combine(const double* phi0, const double* U0, const double* U1, const int64_t* n,
                    int nd, double* out) {
  const int S = 2 + nd;
  const int64_t total = n[0] * (nd > 1 ? n[1] : 1);
  for (int64_t i = 0; i < total; ++i) {
    const double p = phi0[i];
    const double* src = ((p <= 0.0) || (p <= -0.0)) ? U0 : U1;
    memcpy(out + i * S, src + i * S, sizeof(double) * S);
  }
}

/* MultimaterialSimulator::simulate_mm (simulator/multimaterial_simulator.hpp:227-303)
 * for two materials, in the order of decision L7. */
int orc_mm_simulate(double* U0, double* U1, double* phi0, double* phi1, const int64_t* n,
                    int nd, const orc_scheme* s0, const orc_scheme* s1, double cfl, double dh,
                    double sim_time, int64_t max_iters, int band, int extrap_sweeps,
                    int reinit_every, int nthreads, double* run_time, int64_t* iters,
                    double* dt_hist) {
  const int S = 2 + nd;
  const int64_t total = n[0] * (nd > 1 ? n[1] : 1);
  double* W = (double*)malloc(sizeof(double) * total * S);
  double* star0 = (double*)malloc(sizeof(double) * total * S);
  double* star1 = (double*)malloc(sizeof(double) * total * S);
  uint8_t* has0 = (uint8_t*)malloc(total);
  uint8_t* has1 = (uint8_t*)malloc(total);
  double* V = (double*)malloc(sizeof(double) * total);
  double* V1 = (double*)malloc(sizeof(double) * total);
  double* tmp = (double*)malloc(sizeof(double) * total);
  int rc = 0;
  double t = 0.0;
  int64_t it = 0;
  if (!W || !star0 || !star1 || !has0 || !has1 || !V || !V1 || !tmp) { rc = -3; goto done; }
  while (t < sim_time && it < max_iters) {
    orc_gfm_star(phi0, phi1, U0, U1, s0, s1, n, nd, dh, star0, has0);
    orc_gfm_star(phi1, phi0, U1, U0, s1, s0, n, nd, dh, star1, has1);
    orc_gfm_extrapolate(phi0, U0, star0, has0, n, nd, dh, extrap_sweeps);
    orc_gfm_extrapolate(phi1, U1, star1, has1, n, nd, dh, extrap_sweeps);
    const double ws0 = orc_max_wavespeed_grid(U0, n, nd, s0, nthreads);
    const double ws1 = orc_max_wavespeed_grid(U1, n, nd, s1, nthreads);
    if (ws0 != ws0 || ws1 != ws1) { rc = -4; goto done; }
    const double maxws = ws1 > ws0 ? ws1 : ws0;
    const double dt = cfl * dh / maxws;
    const double dtdh = dt / dh;
    double* Us[2] = {U0, U1};
    const orc_scheme* ss[2] = {s0, s1};
    for (int m = 0; m < 2; ++m) {
      rc = orc_sweep(Us[m], W, n, nd, ss[m], dtdh, 0, nthreads);
      if (rc) goto done;
      if (nd > 1) rc = orc_sweep(W, Us[m], n, nd, ss[m], dtdh, 1, nthreads);
      else memcpy(Us[m], W, sizeof(double) * total * S);
      if (rc) goto done;
    }
    const double* phis[2] = {phi0, phi1};
    const double* states[2] = {U0, U1};
    orc_ls_velocity(2, phis, states, n, nd, dh, 0, V);
    orc_ls_velocity(2, phis, states, n, nd, dh, 1, V1);
    orc_ls_advect(phi0, V, n, nd, dt, dh, tmp);
    memcpy(phi0, tmp, sizeof(double) * total);
    orc_ls_advect(phi1, V1, n, nd, dt, dh, tmp);
    memcpy(phi1, tmp, sizeof(double) * total);
    orc_ls_project2(phi0, phi1, total);
    if (it % reinit_every == reinit_every - 1) {
      rc = orc_ls_reinit(phi0, n, nd, dh, band, tmp, NULL, NULL);
      if (rc) goto done;
      memcpy(phi0, tmp, sizeof(double) * total);
      rc = orc_ls_reinit(phi1, n, nd, dh, band, tmp, NULL, NULL);
      if (rc) goto done;
      memcpy(phi1, tmp, sizeof(double) * total);
    }
    if (dt_hist) dt_hist[it] = dt;
    t += dt;
    it += 1;
  }
done:
  if (run_time) *run_time = t;
  if (iters) *iters = it;
  free(W); free(star0); free(star1); free(has0); free(has1); free(V); free(V1); free(tmp);
  return rc;
}

/* MultimaterialSimulator::simulate_mm for nmat >= 2 materials (L7-L9): the
 * two-material loop above, generalised; nmat == 2 computes exactly what
 * orc_mm_simulate does. */
int orc_mm_simulate_n(int nmat, double* const* Us, double* const* phis, const int64_t* n,
                      int nd, const orc_scheme* const* ss, double cfl, double dh,
                      double sim_time, int64_t max_iters, int band, int extrap_sweeps,
                      int reinit_every, int nthreads, double* run_time, int64_t* iters,
                      double* dt_hist) {
  const int S = 2 + nd;
  const int64_t total = n[0] * (nd > 1 ? n[1] : 1);
  if (nmat < 2 || nmat > 8) return -2;
  double* W = (double*)malloc(sizeof(double) * total * S);
  double* stars = (double*)malloc(sizeof(double) * total * S * nmat);
  uint8_t* has = (uint8_t*)malloc(total * nmat);
  double* V = (double*)malloc(sizeof(double) * total * nmat);
  double* tmp = (double*)malloc(sizeof(double) * total);
  int rc = 0;
  double t = 0.0;
  int64_t it = 0;
  if (!W || !stars || !has || !V || !tmp) { rc = -3; goto done; }
  while (t < sim_time && it < max_iters) {
    const double* cphis[8];
    const double* cUs[8];
    for (int m = 0; m < nmat; ++m) { cphis[m] = phis[m]; cUs[m] = Us[m]; }
    for (int m = 0; m < nmat; ++m)
      orc_gfm_star_n(m, nmat, cphis, cUs, ss, n, nd, dh, stars + m * total * S, has + m * total);
    for (int m = 0; m < nmat; ++m)
      orc_gfm_extrapolate(phis[m], Us[m], stars + m * total * S, has + m * total, n, nd, dh,
                          extrap_sweeps);
    double maxws = 0.0;
    for (int m = 0; m < nmat; ++m) {
      const double ws = orc_max_wavespeed_grid(Us[m], n, nd, ss[m], nthreads);
      if (ws != ws) { rc = -4; goto done; }
      maxws = ws > maxws ? ws : maxws;
    }
    const double dt = cfl * dh / maxws;
    const double dtdh = dt / dh;
    for (int m = 0; m < nmat; ++m) {
      rc = orc_sweep(Us[m], W, n, nd, ss[m], dtdh, 0, nthreads);
      if (rc) goto done;
      if (nd > 1) rc = orc_sweep(W, Us[m], n, nd, ss[m], dtdh, 1, nthreads);
      else memcpy(Us[m], W, sizeof(double) * total * S);
      if (rc) goto done;
    }
    for (int m = 0; m < nmat; ++m) cUs[m] = Us[m];
    for (int m = 0; m < nmat; ++m)
      orc_ls_velocity(nmat, cphis, cUs, n, nd, dh, m, V + m * total);
    for (int m = 0; m < nmat; ++m) {
      orc_ls_advect(phis[m], V + m * total, n, nd, dt, dh, tmp);
      memcpy(phis[m], tmp, sizeof(double) * total);
    }
    if (nmat == 2) orc_ls_project2(phis[0], phis[1], total);
    else orc_ls_project_n(nmat, phis, total);
    if (it % reinit_every == reinit_every - 1) {
      for (int m = 0; m < nmat; ++m) {
        rc = orc_ls_reinit(phis[m], n, nd, dh, band, tmp, NULL, NULL);
        if (rc) goto done;
        memcpy(phis[m], tmp, sizeof(double) * total);
      }
    }
    if (dt_hist) dt_hist[it] = dt;
    t += dt;
    it += 1;
  }
done:
  if (run_time) *run_time = t;
  if (iters) *iters = it;
  free(W); free(stars); free(has); free(V); free(tmp);
  return rc;
}
