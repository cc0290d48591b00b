/*
 * fl_oracle.h -- CPU restatement of fluidity's finite-volume Euler step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  No product code links it.
 *
 * Every function restates the reference header named beside it, keeping the
 * reference's floating-point operation order so that an FP64 build with
 * -ffp-contract=off reproduces the reference bit for bit.
 *
 * State vectors are the reference's conservative layout
 * (state/state.hpp:88-98): [rho, E, m_0, .., m_{nd-1}, add_0, ..] and the
 * primitive layout [rho, p, v_0, .., v_{nd-1}, add_0, ..].
 */
#ifndef FL_ORACLE_H
#define FL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LIM_VANLEER = 0, ORC_LIM_SUPERBEE = 1 };
enum { ORC_FORM_CONS = 0, ORC_FORM_PRIM = 1 };
enum { ORC_FLUX_HLLC = 0, ORC_FLUX_FORCE = 1, ORC_FLUX_LF = 2, ORC_FLUX_RICHTMYER = 3 };
/* ORC_EOS_STIFF: the reference's StiffFluid with its pressure, which ignores
 * p_inf.  ORC_EOS_STIFF_CORRECTED: the pressure of the stiffened-gas energy,
 * p = (gamma-1)(E - rho |v|^2/2) - gamma p_inf, and HLLC pressure ratios of
 * p + p_inf (an extension beside the parity default). */
enum { ORC_EOS_IDEAL = 0, ORC_EOS_STIFF = 1, ORC_EOS_STIFF_CORRECTED = 2 };
enum { ORC_BC_TRANSMISSIVE = 0, ORC_BC_REFLECTIVE = 1, ORC_BC_PERIODIC = 2 };

typedef struct orc_scheme {
  int32_t limiter;   /* ORC_LIM_*  */
  int32_t form;      /* ORC_FORM_* : variables the limiter ratio is taken on */
  int32_t flux;      /* ORC_FLUX_* */
  int32_t eos;       /* ORC_EOS_*  */
  double  gamma;     /* adiabatic index */
  double  p_inf;     /* stiffened-gas constant (ORC_EOS_STIFF only) */
  int32_t bc[3];     /* ORC_BC_* per dimension, applied at both ends */
  int32_t pad;
} orc_scheme;

/* ---- single-state numerics (state/state_impl.hpp, material/ideal_gas.hpp, stiff_fluid.hpp) ---- */
double orc_pressure(const double* U, int nd, int na, const orc_scheme* s);
double orc_sound_speed(const double* U, int nd, int na, const orc_scheme* s);
double orc_max_wavespeed(const double* U, int nd, int na, const orc_scheme* s);
void   orc_flux(const double* U, int nd, int na, const orc_scheme* s, int dim, double* F);
void   orc_flux_prim(const double* W, int nd, int na, const orc_scheme* s, int dim, double* F);
void   orc_cons_to_prim(const double* U, int nd, int na, const orc_scheme* s, double* W);
void   orc_prim_to_cons(const double* W, int nd, int na, const orc_scheme* s, double* U);
void   orc_apply_setter(double* U, int nd, int dim, int kind);

/* ---- limiters (limiting/van_leer_limiter.hpp, superbee_limiter.hpp) ---- */
double orc_limit_single(int limiter, double left, double right);

/* ---- numerical fluxes (flux_method/flux_hllc.hpp, flux_force.hpp, ..) ---- */
void orc_numerical_flux(const double* ul, const double* ur, int nd,
                        const orc_scheme* s, double dtdh, int dim, double* F);

/* ---- whole-grid operations on interior-only AoS arrays ---- */
/* n[0] = cells in x (fastest), n[1] = cells in y (1 for 1D). */
int    orc_sweep(const double* in, double* out, const int64_t* n, int nd,
                 const orc_scheme* s, double dtdh, int dim, int nthreads);
double orc_max_wavespeed_grid(const double* U, const int64_t* n, int nd,
                              const orc_scheme* s, int nthreads);
void   orc_flux_delta_at(const double* U, const int64_t* n, int nd,
                         const orc_scheme* s, double dtdh, int dim,
                         int64_t x, int64_t y, double* out);
void   orc_face_states_at(const double* U, const int64_t* n, int nd,
                          const orc_scheme* s, double dtdh, int dim,
                          int64_t x, int64_t y, double* R, double* L,
                          double* chi);
/* Runs the simulator loop (simulator/simulator_impl.hpp:222-258).  U holds the
 * initial state and receives the final one; work is a same-sized buffer.  When
 * dt_hist is not NULL it receives the dt of each step (max_iters entries). */
int    orc_simulate(double* U, double* work, const int64_t* n, int nd,
                    const orc_scheme* s, double cfl, double dh,
                    double sim_time, int64_t max_iters, int nthreads,
                    double* run_time, int64_t* iters, double* dt_hist);

/* ---- level set / ghost fluid (fl_oracle_ls.c) ---- */
/* mask bits: 1 inside, 2 outside, 4 on boundary, 8 inside interfacial,
 * 16 outside interfacial (levelset/levelset.hpp:166-217). */
uint8_t orc_ls_mask_value(double phi, double dh);
void orc_ls_masks(const double* phi, const int64_t* n, int nd, double dh, uint8_t* masks);
void orc_ls_velocity(int nmat, const double* const* phis, const double* const* states,
                     const int64_t* n, int nd, double dh, int target, double* V);
void orc_ls_advect(const double* phi, const double* V, const int64_t* n, int nd, double dt,
                   double dh, double* out);
void orc_ls_weno(const double* phi, int64_t n, double dh, int64_t i, double* dminus,
                 double* dplus);
void orc_ls_project2(double* phi0, double* phi1, int64_t count);
void orc_ls_project_n(int nmat, double* const* phis, int64_t count);
int  orc_ls_reinit(const double* phi, const int64_t* n, int nd, double dh, int band,
                   double* out, uint8_t* band_mask, int64_t* iterations);
void orc_gfm_star(const double* phi_m, const double* phi_o, const double* U_m,
                  const double* U_o, const orc_scheme* eos_m, const orc_scheme* eos_o,
                  const int64_t* n, int nd, double dh, double* star, uint8_t* has_star);
void orc_gfm_extrapolate(const double* phi_m, double* U_m, const double* star,
                         const uint8_t* has_star, const int64_t* n, int nd, double dh,
                         int sweeps);
void orc_mm_combine(const double* phi0, const double* U0, const double* U1, const int64_t* n,
                    int nd, double* out);
int  orc_mm_simulate(double* U0, double* U1, double* phi0, double* phi1, const int64_t* n,
                     int nd, const orc_scheme* s0, const orc_scheme* s1, double cfl, double dh,
                     double sim_time, int64_t max_iters, int band, int extrap_sweeps,
                     int reinit_every, int nthreads, double* run_time, int64_t* iters,
                     double* dt_hist);
void orc_gfm_star_n(int m, int nmat, const double* const* phis, const double* const* Us,
                    const orc_scheme* const* eos, const int64_t* n, int nd, double dh,
                    double* star, uint8_t* has_star);
int  orc_mm_simulate_n(int nmat, double* const* Us, double* const* phis, const int64_t* n,
                       int nd, const orc_scheme* const* ss, double cfl, double dh,
                       double sim_time, int64_t max_iters, int band, int extrap_sweeps,
                       int reinit_every, int nthreads, double* run_time, int64_t* iters,
                       double* dt_hist);

/* ---- initial conditions and result files (fl_oracle_io.cpp) ---- */
int  orc_write_ascii(const double* U, const int64_t* n, int nd, const orc_scheme* s,
                     const double* phi, double run_time, double resolution, const char* path);
int  orc_write_separate_raw(const double* U, const int64_t* n, int nd, const orc_scheme* s,
                            double sim_time, const char* prefix);
void orc_set_states(const double* phi, const double* W, const orc_scheme* s, double* U,
                    int64_t cells, int nd);

#ifdef __cplusplus
}
#endif
#endif
