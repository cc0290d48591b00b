// fl_numerics.hpp -- per-cell Euler numerics for the gfx950 sweep kernels.
//
// Device functions for fluidity's MUSCL-Hancock / HLLC / FORCE finite-volume
// step.  For T = double every expression keeps the operation order of the
// reference header cited beside it; the library is compiled with
// -ffp-contract=off, and f64 division (for operands away from the ends of the
// exponent range, see dv) and square root are correctly rounded, so the FP64
// path reproduces the reference bit for bit.  For
// T = float the same formulas run in single precision (an extension: the
// reference templates do not compile for float).
//
// State vectors use the reference's conservative element order
// (state/state.hpp:88-98): [rho, E, m_0, .., m_{ND-1}].
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace fl {

enum : int { kLimVanLeer = 0, kLimSuperbee = 1 };
enum : int { kFormCons = 0, kFormPrim = 1 };
enum : int { kFluxHllc = 0, kFluxForce = 1, kFluxLaxFriedrichs = 2, kFluxRichtmyer = 3 };
enum : int { kEosIdeal = 0, kEosStiff = 1, kEosStiffCorrected = 2 };
enum : int { kBcTransmissive = 0, kBcReflective = 1, kBcPeriodic = 2 };

template <typename T> struct Limits;
template <> struct Limits<double> { static __device__ __forceinline__ double max() { return DBL_MAX; } };
template <> struct Limits<float>  { static __device__ __forceinline__ float max() { return FLT_MAX; } };

// std::min / std::max exactly as the reference calls them (comparison based,
// so NaN operands select the same side as in the reference).
template <typename T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }

// math::signum for signed types (math/math.hpp:390-393).
template <typename T> __device__ __forceinline__ int signum(T x) {
  return (T(0) < x) - (x < T(0));
}

// Division and square root.  FP64 division: the reciprocal of b refined by
// two Newton steps, the quotient a*r corrected by one residual step, then
// v_div_fixup for zero, infinite and NaN operands.  That is the compiler's
// IEEE a / b sequence without its operand rescaling (v_div_scale and
// v_div_fmas), which only acts when an operand or the quotient is near the
// ends of the exponent range (b denormal or |b| >= 2^1021, a nonzero with
// |a| < 2^-969, a quotient that would be denormal, or exponents of a and b
// 768 or more apart).  Outside those cases the result is the correctly
// rounded quotient, bit for bit what a / b gives, and since the refined
// reciprocal depends on b alone, quotients with one denominator (m_x/rho,
// m_y/rho, gamma p/rho) share it.  FP64 square root: the compiler's IEEE
// sqrt sequence (v_rsq_f64, one Goldschmidt step for root and half-inverse
// root, two residual corrections, x itself for +-0 and +inf) without its
// rescaling of arguments below 2^-767; for every other argument, negative,
// infinite and NaN ones included, it is the same computation and gives the
// correctly rounded root.  FP32 (bulk-precision extension): v_rcp_f32 / v_sqrt_f32 (about 1 ulp), far
// inside the FP32 error budget.
__device__ __forceinline__ double dv(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = a * r;
  return __builtin_amdgcn_div_fixup(__builtin_fma(__builtin_fma(-b, q, a), r, q), b, a);
}
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sq(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return __builtin_amdgcn_class(x, 0x260) ? x : g;  // class bits: -0, +0, +inf
}
__device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }

// Material parameters; derived constants are computed once on the host with
// the same expressions the reference evaluates per call.
template <typename T> struct Material {
  T gamma;   // adiabatic index
  T p_inf;   // stiffened-gas constant
  T gm1;     // gamma - 1           (state/state_impl.hpp:125)
  T q;       // (gamma+1)/(2 gamma) (flux_method/flux_hllc.hpp:64-67)
  T gpinf;   // gamma p_inf (corrected stiffened gas)
  int stiff; // either stiffened gas: StiffFluid::sound_speed and eos
  int corr;  // kEosStiffCorrected: the pressure of the stiffened-gas
             // energy, p = (gamma-1)(E - rho |v|^2 / 2) - gamma p_inf, and
             // wave-speed ratios (p* + p_inf)/(p + p_inf)
};

template <typename T, int ND> struct Cell { T u[2 + ND]; };

// Primitive quantities of one conservative state, computed once and reused by
// every accessor the reference calls on that state.
template <typename T, int ND> struct Prim {
  T vel[ND];  // m_d / rho                       (state/state_impl.hpp:92-98)
  T p;        // (gamma-1)(E - 1/2 rho |v|^2)    (state/state_impl.hpp:117-128)
};

template <typename T, int ND>
__device__ __forceinline__ Prim<T, ND> prim_of(const T* U, const Material<T>& m) {
  Prim<T, ND> w;
  T sum = T(0);
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    w.vel[i] = dv(U[2 + i], U[0]);
    sum += w.vel[i] * w.vel[i];
  }
  w.p = m.gm1 * (U[1] - T(0.5) * U[0] * sum);
  if (m.corr) w.p = w.p - m.gpinf;
  return w;
}

// IdealGas::sound_speed (material/ideal_gas.hpp:75-78) and
// StiffFluid::sound_speed (material/stiff_fluid.hpp:95-98).
template <typename T>
__device__ __forceinline__ T sound_speed(T p, T rho, const Material<T>& m) {
  if (m.stiff) return sq(dv(m.gamma * (p + m.p_inf), rho));
  return sq(dv(m.gamma * p, rho));
}

// Conservative Euler flux in dimension D (state/state_impl.hpp:293-310 with
// make_other_fluxes :228-244).
template <typename T, int ND, int D>
__device__ __forceinline__ void euler_flux(const T* U, const Prim<T, ND>& w, T* F) {
  const T v = w.vel[D];
  F[0] = U[2 + D];
  F[2 + D] = U[2 + D] * v + w.p;
  F[1] = v * (U[1] + w.p);
#pragma unroll
  for (int o = 0; o < ND; ++o)
    if (o != D) F[2 + o] = U[2 + o] * v;
}

// max_wavespeed (state/state_impl.hpp:167-178): max_d |u_d| + a.
template <typename T, int ND>
__device__ __forceinline__ T max_wavespeed(const T* U, const Prim<T, ND>& w,
                                           const Material<T>& m) {
  T mv = fabs(w.vel[0]);
#pragma unroll
  for (int i = 1; i < ND; ++i) mv = smax(mv, T(fabs(w.vel[i])));
  return mv + sound_speed(w.p, U[0], m);
}

// VanLeer::limit_single (limiting/van_leer_limiter.hpp:65-74) and
// Superbee::limit_single (limiting/superbee_limiter.hpp:72-89).
template <typename T>
__device__ __forceinline__ T limit_single(int limiter, T left, T right) {
  const T r = dv(left, right);
  if (limiter == kLimSuperbee) {
    if (r <= T(0) || right == T(0)) return T(0);
    if (r <= T(0.5)) return T(2) * r;
    if (r <= T(1)) return T(1);
    const T Er = dv(T(2), T(1) + r);
    return smin(Er, smin(r, T(2)));
  }
  return (r <= T(0) || right == T(0)) ? T(0) : dv(T(2) * smin(r, T(1)), T(1) + r);
}

// MHReconstructor::chi's flattening factor about a cell in one dimension
// (reconstruction/muscl_reconstructor.hpp:59-87), from the pressures at
// offsets -2,-1,+1,+2 and the velocities at -1,+1 in that dimension.  The
// corrected stiffened gas measures the jump against min(p) + p_inf (its
// shock relations are the ideal gas's in p + p_inf).
template <typename T>
__device__ __forceinline__ T chi_factor(T p_pprev, T p_prev, T p_next, T p_nnext,
                                        T v_prev, T v_next, const Material<T>& m) {
  const T u_factor = v_prev - v_next;
  const T dp = fabs(p_next - p_prev);
  T pmin = smin(p_next, p_prev);
  if (m.corr) pmin = pmin + m.p_inf;
  const bool flatten = u_factor > T(0) && dv(dp, pmin) > T(0.33);
  if (!flatten) return T(1);
  const T z = dv(dp, T(fabs(p_nnext - p_pprev)));
  const T temp = T(1) - dv(z - T(0.75), T(0.85) - T(0.75));
  return smin(T(1), smax(T(0), temp));
}

// Both MUSCL-Hancock evolved face states of one cell in sweep dimension D
// (reconstruction/muscl_reconstructor.hpp:99-172): with
// delta = (chi*1/2) * limit(U) (limiting/limiter.hpp:79-97),
//   R = (U+delta) + (1/2 dtdh)(F(U-delta) - F(U+delta))   right face
//   L = (U-delta) + (1/2 dtdh)(F(U-delta) - F(U+delta))   left face
template <typename T, int ND, int D>
__device__ __forceinline__ void evolve_faces(const T* um, const T* u0, const T* up,
                                             T chi, T half_dtdh, int limiter, int form,
                                             const Material<T>& m, T* R, T* L) {
  constexpr int S = 2 + ND;
  T bd[S], fd[S];
  if (form == kFormPrim) {
    // backward/forward differences of primitive states
    // (limiting/limiting_impl.hpp:77-81, 142-146)
    const Prim<T, ND> wm = prim_of<T, ND>(um, m);
    const Prim<T, ND> w0 = prim_of<T, ND>(u0, m);
    const Prim<T, ND> wp = prim_of<T, ND>(up, m);
    bd[0] = u0[0] - um[0];
    fd[0] = up[0] - u0[0];
    bd[1] = w0.p - wm.p;
    fd[1] = wp.p - w0.p;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      bd[2 + i] = w0.vel[i] - wm.vel[i];
      fd[2 + i] = wp.vel[i] - w0.vel[i];
    }
  }
  const T cscale = chi * T(0.5);
  T plus[S], minus[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const T back = u0[k] - um[k];
    const T fwrd = up[k] - u0[k];
    T lim = (back + fwrd) * T(0.5);
    lim *= (form == kFormPrim) ? limit_single(limiter, bd[k], fd[k])
                               : limit_single(limiter, back, fwrd);
    const T delta = lim * cscale;
    plus[k] = u0[k] + delta;
    minus[k] = u0[k] - delta;
  }
  T fm[S], fp[S];
  const Prim<T, ND> wmi = prim_of<T, ND>(minus, m);
  const Prim<T, ND> wpl = prim_of<T, ND>(plus, m);
  euler_flux<T, ND, D>(minus, wmi, fm);
  euler_flux<T, ND, D>(plus, wpl, fp);
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const T dflux = (fm[k] - fp[k]) * half_dtdh;
    R[k] = plus[k] + dflux;
    L[k] = minus[k] + dflux;
  }
}

// Hllc::Impl::wavespeed (flux_method/flux_hllc.hpp:82-99).  The corrected
// stiffened gas takes the pressure ratio of p + p_inf (its shock relations
// are the ideal gas's in p + p_inf).
template <typename T>
__device__ __forceinline__ T hllc_wavespeed(T u, T p, T pstar, T sound, const Material<T>& m) {
  if (pstar <= p) return u + sound;
  if (m.corr) return u + sound * sq(T(1) + m.q * (dv(pstar + m.p_inf, p + m.p_inf) - T(1)));
  return u + sound * sq(T(1) + m.q * (dv(pstar, p) - T(1)));
}

// Hllc::Impl::operator() (flux_method/flux_hllc.hpp:202-245) with
// star_speed (:113-137) and star_state (:150-188).
template <typename T, int ND, int D>
__device__ __forceinline__ void hllc_flux(const T* ul, const T* ur, const Material<T>& m,
                                          T* F) {
  constexpr int S = 2 + ND;
  const Prim<T, ND> wl = prim_of<T, ND>(ul, m);
  const Prim<T, ND> wr = prim_of<T, ND>(ur, m);
  const T al = sound_speed(wl.p, ul[0], m);
  const T ar = sound_speed(wr.p, ur[0], m);
  const T vl = wl.vel[D], vr = wr.vel[D];
  const T p_pvrs = T(0.5) * ((wl.p + wr.p) - (T(0.25) * (vr - vl) * (ur[0] + ul[0]) * (al + ar)));
  // the corrected stiffened gas bounds p* by -p_inf (p + p_inf >= 0)
  const T p_star = m.corr ? smax(-m.p_inf, p_pvrs) : smax(T(0), p_pvrs);

  const T wsl = hllc_wavespeed(vl, wl.p, p_star, -al, m);
  if (T(0) <= wsl) {
    euler_flux<T, ND, D>(ul, wl, F);
    return;
  }
  const T wsr = hllc_wavespeed(vr, wr.p, p_star, ar, m);
  if (T(0) >= wsr) {
    euler_flux<T, ND, D>(ur, wr, F);
    return;
  }
  const T factorl = ul[0] * (wsl - vl);
  const T factorr = ur[0] * (wsr - vr);
  const T ws_star = dv(wr.p - wl.p + vl * factorl - vr * factorr, factorl - factorr);

  // One side of the star region; the reference's two branches are the same
  // expressions on (ul, wsl) or (ur, wsr).
  const bool left = T(0) <= ws_star;
  const T* st = left ? ul : ur;
  const Prim<T, ND>& w = left ? wl : wr;
  const T ws = left ? wsl : wsr;
  T f[S];
  euler_flux<T, ND, D>(st, w, f);
  const T state_factor = ws - w.vel[D];
  const T scale_factor = dv(state_factor, ws - ws_star);
  T us[S];
#pragma unroll
  for (int k = 0; k < S; ++k) us[k] = st[k] * scale_factor;
  us[2 + D] = ws_star * st[0] * scale_factor;
  us[1] = scale_factor *
          (st[1] + st[0] * (ws_star - w.vel[D]) * (ws_star + dv(w.p, st[0] * state_factor)));
#pragma unroll
  for (int k = 0; k < S; ++k) F[k] = f[k] + (us[k] - st[k]) * ws;
}

// Force::Impl::operator() (flux_method/flux_force.hpp:73-84):
// 1/2 (LaxFriedrichs (flux_lax_friedrichs.hpp:62-78) + Richtmyer
// (flux_richtmyer.hpp:78-88)).
template <typename T, int ND, int D>
__device__ __forceinline__ void force_flux(const T* ul, const T* ur, T dtdh, T dhdt,
                                           const Material<T>& m, T* F) {
  constexpr int S = 2 + ND;
  const Prim<T, ND> wl = prim_of<T, ND>(ul, m);
  const Prim<T, ND> wr = prim_of<T, ND>(ur, m);
  T fl[S], fr[S], uri[S];
  euler_flux<T, ND, D>(ul, wl, fl);
  euler_flux<T, ND, D>(ur, wr, fr);
#pragma unroll
  for (int k = 0; k < S; ++k) uri[k] = (ul[k] + ur[k] + (fl[k] - fr[k]) * dtdh) * T(0.5);
  const Prim<T, ND> wri = prim_of<T, ND>(uri, m);
  T fri[S];
  euler_flux<T, ND, D>(uri, wri, fri);
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const T lf = (fl[k] + fr[k] + (ul[k] - ur[k]) * dhdt) * T(0.5);
    F[k] = (lf + fri[k]) * T(0.5);
  }
}

// LaxFriedrichs::Impl::operator() (flux_method/flux_lax_friedrichs.hpp:62-78).
template <typename T, int ND, int D>
__device__ __forceinline__ void lax_friedrichs_flux(const T* ul, const T* ur, T dhdt,
                                                    const Material<T>& m, T* F) {
  constexpr int S = 2 + ND;
  const Prim<T, ND> wl = prim_of<T, ND>(ul, m);
  const Prim<T, ND> wr = prim_of<T, ND>(ur, m);
  T fl[S], fr[S];
  euler_flux<T, ND, D>(ul, wl, fl);
  euler_flux<T, ND, D>(ur, wr, fr);
#pragma unroll
  for (int k = 0; k < S; ++k) F[k] = (fl[k] + fr[k] + (ul[k] - ur[k]) * dhdt) * T(0.5);
}

// Richtmyer::Impl::operator() (flux_method/flux_richtmyer.hpp:78-88).
template <typename T, int ND, int D>
__device__ __forceinline__ void richtmyer_flux(const T* ul, const T* ur, T dtdh,
                                               const Material<T>& m, T* F) {
  constexpr int S = 2 + ND;
  const Prim<T, ND> wl = prim_of<T, ND>(ul, m);
  const Prim<T, ND> wr = prim_of<T, ND>(ur, m);
  T fl[S], fr[S], uri[S];
  euler_flux<T, ND, D>(ul, wl, fl);
  euler_flux<T, ND, D>(ur, wr, fr);
#pragma unroll
  for (int k = 0; k < S; ++k) uri[k] = (ul[k] + ur[k] + (fl[k] - fr[k]) * dtdh) * T(0.5);
  const Prim<T, ND> wri = prim_of<T, ND>(uri, m);
  euler_flux<T, ND, D>(uri, wri, F);
}

template <typename T, int ND, int D, int FLUX>
__device__ __forceinline__ void numerical_flux(const T* ul, const T* ur, T dtdh, T dhdt,
                                               const Material<T>& m, T* F) {
  if constexpr (FLUX == kFluxHllc) {
    hllc_flux<T, ND, D>(ul, ur, m, F);
  } else if constexpr (FLUX == kFluxLaxFriedrichs) {
    lax_friedrichs_flux<T, ND, D>(ul, ur, dhdt, m, F);
  } else if constexpr (FLUX == kFluxRichtmyer) {
    richtmyer_flux<T, ND, D>(ul, ur, dtdh, m, F);
  } else {
    force_flux<T, ND, D>(ul, ur, dtdh, dhdt, m, F);
  }
}

// BoundarySetter::set_velocity for a reflective ghost
// (solver/boundary_loader.hpp:76-84, state/state.hpp:435-441):
// m_d <- (-(m_d / rho)) * rho.
template <typename T, int ND>
__device__ __forceinline__ void reflect(T* U, int d) {
  T v = -dv(U[2 + d], U[0]);
  v *= U[0];
  U[2 + d] = v;
}

}  // namespace fl
