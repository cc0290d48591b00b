/*
 * fl_oracle.c -- CPU restatement of fluidity's finite-volume Euler step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and timed CPU baseline); see
 * fl_oracle.h.  Build with -ffp-contract=off and without -ffast-math: every
 * expression below keeps the operation order of the reference line it cites,
 * so contraction or reassociation would break bit parity.
 *
 * Reference semantics restated here (paths relative to include/fluidity/):
 *  - state accessors and Euler flux ........ state/state_impl.hpp:28-406
 *  - ideal / stiffened gas .................. material/ideal_gas.hpp:27-119,
 *                                             material/stiff_fluid.hpp:38-133
 *  - limiters ............................... limiting/limiter.hpp:79-97,
 *                                             limiting/limiting_impl.hpp:72-188,
 *                                             limiting/van_leer_limiter.hpp:65-74,
 *                                             limiting/superbee_limiter.hpp:72-89
 *  - MUSCL-Hancock evolve + chi flattening .. reconstruction/muscl_reconstructor.hpp:59-263
 *  - HLLC ................................... flux_method/flux_hllc.hpp:56-259
 *  - FORCE = 1/2 (LF + Richtmyer) ........... flux_method/flux_force.hpp:73-84,
 *                                             flux_method/flux_lax_friedrichs.hpp:62-78,
 *                                             flux_method/flux_richtmyer.hpp:78-88
 *  - face flux composition .................. solver/flux_solver.hpp:66-155
 *  - sweep update and ghost cells ........... solver/split_solver.hpp:138-162,
 *                                             solver/boundary_loader.hpp:42-199
 *  - wavespeed, dt, step loop ............... state/state_impl.hpp:167-178,
 *                                             simulator/parameters.hpp:60-99,
 *                                             simulator/simulator_impl.hpp:222-258
 *
 * Decisions where the reference is undefined or cannot compile:
 *  1. chi's reads across the non-sweep dimension (2D) read the full-grid
 *     current input, with the same ghost rule as the sweep dimension.  On the
 *     GPU the reference reads outside its LDS patch there.
 *  2. HLLC is the reference formula with its two compile fixes only (public
 *     get(), star_state returning all state elements).
 *  3. Stiffened gas keeps the reference's pressure, which ignores p_inf.
 *  4. Periodic boundaries are an extension (the reference has transmissive
 *     and reflective only).
 */
#include "fl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXS 8

/* std::min / std::max as the reference calls them: comparison based. */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* math::signum for a signed type (math/math.hpp:390-393). */
static inline int signum_d(double x) { return (0.0 < x) - (x < 0.0); }

/* ------------------------------------------------------------------------ */
/* single-state numerics                                                      */
/* ------------------------------------------------------------------------ */

/* v_squared_sum, conservative overload (state/state_impl.hpp:49-61). */
static inline double vss_cons(const double* U, int nd) {
  double sum = 0.0;
  for (int i = 0; i < nd; ++i) {
    const double v = U[2 + i] / U[0];
    sum += v * v;
  }
  return sum;
}

/* v_squared_sum, primitive overload (state/state_impl.hpp:32-43). */
static inline double vss_prim(const double* W, int nd) {
  double sum = 0.0;
  for (int i = 0; i < nd; ++i) {
    const double v = W[2 + i];
    sum += v * v;
  }
  return sum;
}

/* velocity, conservative overload (state/state_impl.hpp:92-98). */
static inline double vel_cons(const double* U, int d) { return U[2 + d] / U[0]; }

static inline int is_stiff(const orc_scheme* s) {
  return s->eos == ORC_EOS_STIFF || s->eos == ORC_EOS_STIFF_CORRECTED;
}

/* pressure, conservative overload (state/state_impl.hpp:117-128).  Uses the
 * adiabatic index only, also for the stiffened gas; the corrected stiffened
 * gas subtracts gamma p_inf (the inverse of StiffFluid::eos). */
static inline double pres_cons(const double* U, int nd, const orc_scheme* s) {
  const double p = (s->gamma - 1.0) * (U[1] - 0.5 * U[0] * vss_cons(U, nd));
  if (s->eos == ORC_EOS_STIFF_CORRECTED) return p - s->gamma * s->p_inf;
  return p;
}

/* IdealGas::sound_speed / StiffFluid::sound_speed. */
static inline double sound_from_p(double p, double rho, const orc_scheme* s) {
  if (is_stiff(s))
    return sqrt(s->gamma * (p + s->p_inf) / rho);   /* stiff_fluid.hpp:95-98 */
  return sqrt(s->gamma * p / rho);                   /* ideal_gas.hpp:75-78 */
}

/* IdealGas::eos / StiffFluid::eos (specific internal energy). */
static inline double eos_energy(double p, double rho, const orc_scheme* s) {
  if (is_stiff(s))
    return (p + s->gamma * s->p_inf) / ((s->gamma - 1.0) * rho);
  return p / ((s->gamma - 1.0) * rho);
}

double orc_pressure(const double* U, int nd, int na, const orc_scheme* s) {
  (void)na;
  return pres_cons(U, nd, s);
}

double orc_sound_speed(const double* U, int nd, int na, const orc_scheme* s) {
  (void)na;
  return sound_from_p(pres_cons(U, nd, s), U[0], s);
}

/* max_wavespeed (state/state_impl.hpp:167-178): max_d |u_d| + a. */
double orc_max_wavespeed(const double* U, int nd, int na, const orc_scheme* s) {
  (void)na;
  double mv = fabs(vel_cons(U, 0));
  for (int i = 1; i < nd; ++i) mv = smax(mv, fabs(vel_cons(U, i)));
  return mv + sound_from_p(pres_cons(U, nd, s), U[0], s);
}

/* flux, conservative overload (state/state_impl.hpp:293-310) with
 * make_other_fluxes (:228-244). */
void orc_flux(const double* U, int nd, int na, const orc_scheme* s, int dim, double* F) {
  const double v = vel_cons(U, dim);
  const double p = pres_cons(U, nd, s);
  F[0]       = U[2 + dim];
  F[2 + dim] = U[2 + dim] * v + p;
  F[1]       = v * (U[1] + p);
  const int iters = nd + na - 1;
  for (int i = 0; i < iters; ++i) {
    const int idx = 2 + i + (i >= dim ? 1 : 0);
    F[idx] = U[idx] * v;
  }
}

/* energy, primitive overload (state/state_impl.hpp:135-142). */
static inline double energy_prim(const double* W, int nd, const orc_scheme* s) {
  return W[0] * (0.5 * vss_prim(W, nd) + eos_energy(W[1], W[0], s));
}

/* flux, primitive overload (state/state_impl.hpp:263-283). */
void orc_flux_prim(const double* W, int nd, int na, const orc_scheme* s, int dim, double* F) {
  const double v = W[2 + dim];
  F[0]       = W[0] * v;
  F[1]       = v * (energy_prim(W, nd, s) + W[1]);
  F[2 + dim] = F[0] * v + W[1];
  const int iters = nd + na - 1;
  for (int i = 0; i < iters; ++i) {
    const int idx = 2 + i + (i >= dim ? 1 : 0);
    F[idx] = W[idx] * v;
    F[idx] *= W[0];
  }
}

/* primitive(), conservative overload (state/state_impl.hpp:340-358). */
void orc_cons_to_prim(const double* U, int nd, int na, const orc_scheme* s, double* W) {
  W[0] = U[0];
  W[1] = pres_cons(U, nd, s);
  for (int i = 0; i < nd + na; ++i) W[2 + i] = U[2 + i] / U[0];
}

/* conservative(), primitive overload (state/state_impl.hpp:388-406). */
void orc_prim_to_cons(const double* W, int nd, int na, const orc_scheme* s, double* U) {
  U[0] = W[0];
  U[1] = energy_prim(W, nd, s);
  for (int i = 0; i < nd + na; ++i) U[2 + i] = W[2 + i] * W[0];
}

/* BoundarySetter::operator() (solver/boundary_loader.hpp:64-84): the ghost is
 * a copy; reflective negates the normal velocity through set_velocity, which
 * for a conservative state stores (-(m/rho)) * rho (state/state.hpp:435-441). */
void orc_apply_setter(double* U, int nd, int dim, int kind) {
  (void)nd;
  if (kind == ORC_BC_REFLECTIVE) {
    double v = -(U[2 + dim] / U[0]);
    v *= U[0];
    U[2 + dim] = v;
  }
}

/* ------------------------------------------------------------------------ */
/* limiters                                                                   */
/* ------------------------------------------------------------------------ */

double orc_limit_single(int limiter, double left, double right) {
  const double r = left / right;
  if (limiter == ORC_LIM_SUPERBEE) {            /* superbee_limiter.hpp:72-89 */
    if (r <= 0.0 || right == 0.0) return 0.0;
    else if (r <= 0.5) return 2.0 * r;
    else if (r <= 1.0) return 1.0;
    const double Er = 2.0 / (1.0 + r);
    return smin(Er, smin(r, 2.0));
  }
  /* van_leer_limiter.hpp:65-74 */
  return (r <= 0.0 || right == 0.0) ? 0.0 : 2.0 * smin(r, 1.0) / (1.0 + r);
}

/* Limiter::limit_generic (limiting/limiter.hpp:79-97).  um, u0, up are the
 * states at offsets -1, 0, +1 in the sweep dimension. */
static void limit_state(const double* um, const double* u0, const double* up,
                        int nd, const orc_scheme* s, double* limited) {
  const int S = 2 + nd;
  double fd[ORC_MAXS], bd[ORC_MAXS];
  if (s->form == ORC_FORM_PRIM) {
    double wm[ORC_MAXS], w0[ORC_MAXS], wp[ORC_MAXS];
    orc_cons_to_prim(um, nd, 0, s, wm);
    orc_cons_to_prim(u0, nd, 0, s, w0);
    orc_cons_to_prim(up, nd, 0, s, wp);
    for (int k = 0; k < S; ++k) { fd[k] = wp[k] - w0[k]; bd[k] = w0[k] - wm[k]; }
  } else {
    for (int k = 0; k < S; ++k) { fd[k] = up[k] - u0[k]; bd[k] = u0[k] - um[k]; }
  }
  for (int k = 0; k < S; ++k) {
    const double back = u0[k] - um[k];
    const double fwrd = up[k] - u0[k];
    limited[k] = (back + fwrd) * 0.5;
  }
  for (int k = 0; k < S; ++k) limited[k] *= orc_limit_single(s->limiter, bd[k], fd[k]);
}

/* ------------------------------------------------------------------------ */
/* numerical fluxes                                                           */
/* ------------------------------------------------------------------------ */

/* Hllc::Impl::wavespeed (flux_method/flux_hllc.hpp:82-99); the corrected
 * stiffened gas takes the ratio of p + p_inf. */
static inline double hllc_wavespeed(double u, double p, double pstar,
                                    double sound_speed, double adi_scale, const orc_scheme* s) {
  if (pstar <= p) return u + sound_speed;
  if (s->eos == ORC_EOS_STIFF_CORRECTED)
    return u + sound_speed * sqrt(1.0 + adi_scale * ((pstar + s->p_inf) / (p + s->p_inf) - 1.0));
  return u + sound_speed * sqrt(1.0 + adi_scale * (pstar / p - 1.0));
}

/* Hllc::Impl::star_state (flux_method/flux_hllc.hpp:150-188). */
static void hllc_star_state(const double* st, int nd, const orc_scheme* s, int dim,
                            double state_speed, double star_speed, double* u) {
  const int S = 2 + nd;
  const double state_factor = state_speed - vel_cons(st, dim);
  const double scale_factor = state_factor / (state_speed - star_speed);
  for (int k = 0; k < S; ++k) u[k] = st[k] * scale_factor;
  u[2 + dim] = star_speed * st[0] * scale_factor;
  u[1] = scale_factor *
         (st[1] + st[0] * (star_speed - vel_cons(st, dim)) *
                      (star_speed + pres_cons(st, nd, s) / (st[0] * state_factor)));
}

/* Hllc::Impl::operator() (flux_method/flux_hllc.hpp:202-245). */
static void flux_hllc(const double* ul, const double* ur, int nd,
                      const orc_scheme* s, int dim, double* F) {
  const int S = 2 + nd;
  const double pl = pres_cons(ul, nd, s), pr = pres_cons(ur, nd, s);
  const double al = sound_from_p(pl, ul[0], s);
  const double ar = sound_from_p(pr, ur[0], s);
  const double adi = (s->gamma + 1.0) / (2.0 * s->gamma);
  const double vl = vel_cons(ul, dim), vr = vel_cons(ur, dim);

  const double p_pvrs = 0.5 * ((pl + pr) - (0.25 * (vr - vl) * (ur[0] + ul[0]) * (al + ar)));
  /* the corrected stiffened gas bounds p* by -p_inf */
  const double p_star = s->eos == ORC_EOS_STIFF_CORRECTED ? smax(-s->p_inf, p_pvrs)
                                                          : smax(0.0, p_pvrs);

  const double wsl = hllc_wavespeed(vl, pl, p_star, -al, adi, s);
  double fl[ORC_MAXS];
  orc_flux(ul, nd, 0, s, dim, fl);
  if (0.0 <= wsl) { memcpy(F, fl, sizeof(double) * S); return; }

  const double wsr = hllc_wavespeed(vr, pr, p_star, ar, adi, s);
  double fr[ORC_MAXS];
  orc_flux(ur, nd, 0, s, dim, fr);
  if (0.0 >= wsr) { memcpy(F, fr, sizeof(double) * S); return; }

  /* star_speed (flux_method/flux_hllc.hpp:113-137) */
  const double factorl = ul[0] * (wsl - vl);
  const double factorr = ur[0] * (wsr - vr);
  const double ws_star = (pr - pl + vl * factorl - vr * factorr) / (factorl - factorr);

  double us[ORC_MAXS];
  if (0.0 <= ws_star) {
    hllc_star_state(ul, nd, s, dim, wsl, ws_star, us);
    for (int k = 0; k < S; ++k) F[k] = fl[k] + (us[k] - ul[k]) * wsl;
    return;
  }
  hllc_star_state(ur, nd, s, dim, wsr, ws_star, us);
  for (int k = 0; k < S; ++k) F[k] = fr[k] + (us[k] - ur[k]) * wsr;
}

/* LaxFriedrichs::Impl::operator() (flux_method/flux_lax_friedrichs.hpp:62-78). */
static void flux_lf(const double* ul, const double* ur, int nd, const orc_scheme* s,
                    double dtdh, int dim, double* F) {
  const int S = 2 + nd;
  const double dhdt = 1.0 / dtdh;
  double fl[ORC_MAXS], fr[ORC_MAXS];
  orc_flux(ul, nd, 0, s, dim, fl);
  orc_flux(ur, nd, 0, s, dim, fr);
  for (int k = 0; k < S; ++k) F[k] = (fl[k] + fr[k] + (ul[k] - ur[k]) * dhdt) * 0.5;
}

/* Richtmyer::Impl::operator() (flux_method/flux_richtmyer.hpp:78-88). */
static void flux_ri(const double* ul, const double* ur, int nd, const orc_scheme* s,
                    double dtdh, int dim, double* F) {
  const int S = 2 + nd;
  double fl[ORC_MAXS], fr[ORC_MAXS], u[ORC_MAXS];
  orc_flux(ul, nd, 0, s, dim, fl);
  orc_flux(ur, nd, 0, s, dim, fr);
  for (int k = 0; k < S; ++k) u[k] = (ul[k] + ur[k] + (fl[k] - fr[k]) * dtdh) * 0.5;
  orc_flux(u, nd, 0, s, dim, F);
}

void orc_numerical_flux(const double* ul, const double* ur, int nd,
                        const orc_scheme* s, double dtdh, int dim, double* F) {
  const int S = 2 + nd;
  switch (s->flux) {
    case ORC_FLUX_HLLC: flux_hllc(ul, ur, nd, s, dim, F); return;
    case ORC_FLUX_LF: flux_lf(ul, ur, nd, s, dtdh, dim, F); return;
    case ORC_FLUX_RICHTMYER: flux_ri(ul, ur, nd, s, dtdh, dim, F); return;
    default: {  /* Force::Impl::operator() (flux_method/flux_force.hpp:73-84) */
      double a[ORC_MAXS], b[ORC_MAXS];
      flux_lf(ul, ur, nd, s, dtdh, dim, a);
      flux_ri(ul, ur, nd, s, dtdh, dim, b);
      for (int k = 0; k < S; ++k) F[k] = (a[k] + b[k]) * 0.5;
      return;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* grid access with ghost cells                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
  const double* U;
  int64_t n[2];
  int nd, S;
  const orc_scheme* s;
} grid_t;

/* Ghost index map.  Transmissive and reflective ghosts mirror the interior:
 * ghost(-k-1) = cell(k), ghost(N+k) = cell(N-1-k)
 * (solver/boundary_loader.hpp:180-198).  Periodic wraps. */
static inline int64_t map_index(int64_t i, int64_t n, int kind, int* flip) {
  *flip = 0;
  if (i >= 0 && i < n) return i;
  if (kind == ORC_BC_PERIODIC) return ((i % n) + n) % n;
  if (kind == ORC_BC_REFLECTIVE) *flip = 1;
  return (i < 0) ? (-i - 1) : (2 * n - 1 - i);
}

/* State at (x, y), which may be a ghost in either dimension. */
static inline void grid_state(const grid_t* g, int64_t x, int64_t y, double* out) {
  int fx = 0, fy = 0;
  const int64_t xi = map_index(x, g->n[0], g->s->bc[0], &fx);
  int64_t yi = 0;
  if (g->nd > 1) yi = map_index(y, g->n[1], g->s->bc[1], &fy);
  const double* src = g->U + (yi * g->n[0] + xi) * g->S;
  for (int k = 0; k < g->S; ++k) out[k] = src[k];
  if (fx) orc_apply_setter(out, g->nd, 0, ORC_BC_REFLECTIVE);
  if (fy) orc_apply_setter(out, g->nd, 1, ORC_BC_REFLECTIVE);
}

static inline double grid_pressure(const grid_t* g, int64_t x, int64_t y) {
  double u[ORC_MAXS];
  grid_state(g, x, y, u);
  return pres_cons(u, g->nd, g->s);
}

static inline double grid_velocity(const grid_t* g, int64_t x, int64_t y, int d) {
  double u[ORC_MAXS];
  grid_state(g, x, y, u);
  return vel_cons(u, d);
}

/* MHReconstructor::chi (reconstruction/muscl_reconstructor.hpp:59-87) about
 * cell (x, y) in dimension d. */
static double chi_at(const grid_t* g, int64_t x, int64_t y, int d) {
  const int64_t dx = (d == 0), dy = (d == 1);
  const double p_next  = grid_pressure(g, x + dx, y + dy);
  const double p_prev  = grid_pressure(g, x - dx, y - dy);
  const double p_nnext = grid_pressure(g, x + 2 * dx, y + 2 * dy);
  const double p_pprev = grid_pressure(g, x - 2 * dx, y - 2 * dy);
  const double u_factor = grid_velocity(g, x - dx, y - dy, d) -
                          grid_velocity(g, x + dx, y + dy, d);
  /* the corrected stiffened gas measures the jump against min(p) + p_inf */
  const double pmin = g->s->eos == ORC_EOS_STIFF_CORRECTED ? smin(p_next, p_prev) + g->s->p_inf
                                                           : smin(p_next, p_prev);
  const int flatten = u_factor > 0.0 && (fabs(p_next - p_prev) / pmin) > 0.33;
  if (!flatten) return 1.0;
  const double z = fabs(p_next - p_prev) / fabs(p_nnext - p_pprev);
  const double temp = 1.0 - ((z - 0.75) / (0.85 - 0.75));
  return smin(1.0, smax(0.0, temp));
}

/* The chi_v minimum of MHReconstructor::evolve
 * (reconstruction/muscl_reconstructor.hpp:106-114, 149-157). */
static double chi_min(const grid_t* g, int64_t x, int64_t y) {
  double chi_v = DBL_MAX;
  for (int d = 0; d < g->nd; ++d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    const int sign = signum_d(grid_pressure(g, x + dx, y + dy) -
                              grid_pressure(g, x - dx, y - dy));
    chi_v = smin(chi_v, smin(chi_at(g, x, y, d),
                             chi_at(g, x + sign * dx, y + sign * dy, d)));
  }
  return chi_v;
}

/* Both MUSCL-Hancock evolved face states of cell (x, y) in dimension dim:
 * R = evolve(.., right_face) (muscl_reconstructor.hpp:99-130) and
 * L = evolve(.., left_face) (:142-172).  Both use the same chi_v and delta. */
static void evolve_both(const grid_t* g, int64_t x, int64_t y, int dim,
                        double dtdh, double* R, double* L, double* chi_out) {
  const int S = g->S, nd = g->nd;
  const int64_t dx = (dim == 0), dy = (dim == 1);
  double um[ORC_MAXS], u0[ORC_MAXS], up[ORC_MAXS], lim[ORC_MAXS];
  grid_state(g, x - dx, y - dy, um);
  grid_state(g, x, y, u0);
  grid_state(g, x + dx, y + dy, up);

  const double chi_v = chi_min(g, x, y);
  if (chi_out) *chi_out = chi_v;
  limit_state(um, u0, up, nd, g->s, lim);
  const double cscale = chi_v * 0.5;
  double plus[ORC_MAXS], minus[ORC_MAXS], fp[ORC_MAXS], fm[ORC_MAXS];
  for (int k = 0; k < S; ++k) {
    const double delta = lim[k] * cscale;
    plus[k]  = u0[k] + delta;
    minus[k] = u0[k] - delta;
  }
  orc_flux(minus, nd, 0, g->s, dim, fm);
  orc_flux(plus, nd, 0, g->s, dim, fp);
  const double h = 0.5 * dtdh;
  for (int k = 0; k < S; ++k) {
    const double dflux = (fm[k] - fp[k]) * h;
    R[k] = plus[k] + dflux;
    L[k] = minus[k] + dflux;
  }
}

void orc_face_states_at(const double* U, const int64_t* n, int nd,
                        const orc_scheme* s, double dtdh, int dim,
                        int64_t x, int64_t y, double* R, double* L, double* chi) {
  grid_t g = {U, {n[0], nd > 1 ? n[1] : 1}, nd, 2 + nd, s};
  evolve_both(&g, x, y, dim, dtdh, R, L, chi);
}

/* FaceFlux::flux_delta (solver/flux_solver.hpp:66-72): backward - forward. */
void orc_flux_delta_at(const double* U, const int64_t* n, int nd,
                       const orc_scheme* s, double dtdh, int dim,
                       int64_t x, int64_t y, double* out) {
  grid_t g = {U, {n[0], nd > 1 ? n[1] : 1}, nd, 2 + nd, s};
  const int S = 2 + nd;
  const int64_t dx = (dim == 0), dy = (dim == 1);
  double Rm[ORC_MAXS], Lm[ORC_MAXS], R0[ORC_MAXS], L0[ORC_MAXS], Rp[ORC_MAXS], Lp[ORC_MAXS];
  double fb[ORC_MAXS], ff[ORC_MAXS];
  evolve_both(&g, x - dx, y - dy, dim, dtdh, Rm, Lm, NULL);
  evolve_both(&g, x, y, dim, dtdh, R0, L0, NULL);
  evolve_both(&g, x + dx, y + dy, dim, dtdh, Rp, Lp, NULL);
  orc_numerical_flux(Rm, L0, nd, s, dtdh, dim, fb);
  orc_numerical_flux(R0, Lp, nd, s, dtdh, dim, ff);
  for (int k = 0; k < S; ++k) out[k] = fb[k] - ff[k];
}

/* One split sweep in dimension dim (solver/split_solver.hpp:138-162):
 * out = U + dtdh * (F_{i-1/2} - F_{i+1/2}), each face flux computed once from
 * the evolved states of its two cells. */
int orc_sweep(const double* in, double* out, const int64_t* n, int nd,
              const orc_scheme* s, double dtdh, int dim, int nthreads) {
  if (nd < 1 || nd > 2 || dim < 0 || dim >= nd) return -1;
  const int64_t nx = n[0], ny = nd > 1 ? n[1] : 1;
  if (nx < 4 || (nd > 1 && ny < 4)) return -2;
  grid_t g = {in, {nx, ny}, nd, 2 + nd, s};
  const int S = 2 + nd;
  const int64_t N = n[dim];
  const int64_t nlines = (dim == 0) ? ny : nx;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  int err = 0;
#pragma omp parallel
  {
    double* RL = (double*)malloc(sizeof(double) * (size_t)(N + 2) * 2 * S);
    double* G = (double*)malloc(sizeof(double) * (size_t)(N + 1) * S);
    if (!RL || !G) {
#pragma omp atomic write
      err = -3;
    } else {
#pragma omp for schedule(dynamic, 4)
      for (int64_t line = 0; line < nlines; ++line) {
        /* evolved states for cells -1..N along the line */
        for (int64_t c = -1; c <= N; ++c) {
          const int64_t x = (dim == 0) ? c : line;
          const int64_t y = (dim == 0) ? line : c;
          double* R = RL + (size_t)(c + 1) * 2 * S;
          evolve_both(&g, x, y, dim, dtdh, R, R + S, NULL);
        }
        /* face c+1/2 for c in -1..N-1 */
        for (int64_t c = -1; c < N; ++c) {
          const double* R = RL + (size_t)(c + 1) * 2 * S;
          const double* L = RL + (size_t)(c + 2) * 2 * S + S;
          orc_numerical_flux(R, L, nd, s, dtdh, dim, G + (size_t)(c + 1) * S);
        }
        for (int64_t c = 0; c < N; ++c) {
          const int64_t x = (dim == 0) ? c : line;
          const int64_t y = (dim == 0) ? line : c;
          const double* u = in + (y * nx + x) * S;
          double* o = out + (y * nx + x) * S;
          const double* gb = G + (size_t)c * S;
          const double* gf = G + (size_t)(c + 1) * S;
          for (int k = 0; k < S; ++k) o[k] = u[k] + (gb[k] - gf[k]) * dtdh;
        }
      }
    }
    free(RL);
    free(G);
  }
  return err;
}

/* set_wavespeeds + max_element (simulator/cuda/wavespeed_initialization.cuh:41-53,
 * algorithm/reduce.cuh:44-131).  The maximum is exact, hence order free; a NaN
 * wavespeed makes the result NaN. */
double orc_max_wavespeed_grid(const double* U, const int64_t* n, int nd,
                              const orc_scheme* s, int nthreads) {
  const int64_t total = n[0] * (nd > 1 ? n[1] : 1);
  const int S = 2 + nd;
  double m = 0.0;
  int nan_seen = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for reduction(max : m) reduction(| : nan_seen)
  for (int64_t i = 0; i < total; ++i) {
    const double w = orc_max_wavespeed(U + i * S, nd, 0, s);
    if (w != w) nan_seen = 1;
    else if (w > m) m = w;
  }
  return nan_seen ? NAN : m;
}

/* SimulatorImpl::simulate (simulator/simulator_impl.hpp:222-258) with
 * Parameters::update_time_delta / dt_dh / update_simulation_info
 * (simulator/parameters.hpp:60-99): no clip of the final dt. */
int orc_simulate(double* U, double* work, const int64_t* n, int nd,
                 const orc_scheme* s, double cfl, double dh,
                 double sim_time, int64_t max_iters, int nthreads,
                 double* run_time, int64_t* iters, double* dt_hist) {
  double t = 0.0;
  int64_t it = 0;
  double* a = U;
  double* b = work;
  while (t < sim_time && it < max_iters) {
    const double maxws = orc_max_wavespeed_grid(a, n, nd, s, nthreads);
    const double dt = cfl * dh / maxws;
    const double dtdh = dt / dh;
    int rc = orc_sweep(a, b, n, nd, s, dtdh, 0, nthreads);
    if (rc) return rc;
    if (nd > 1) {
      rc = orc_sweep(b, a, n, nd, s, dtdh, 1, nthreads);
      if (rc) return rc;
    } else {
      double* tmp = a; a = b; b = tmp;
    }
    if (dt_hist) dt_hist[it] = dt;
    t += dt;
    it += 1;
  }
  if (a != U) {
    const int64_t total = n[0] * (nd > 1 ? n[1] : 1) * (2 + nd);
    memcpy(U, a, sizeof(double) * (size_t)total);
  }
  if (run_time) *run_time = t;
  if (iters) *iters = it;
  return 0;
}
