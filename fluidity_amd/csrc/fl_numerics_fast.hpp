// fl_numerics_fast.hpp -- the per-cell Euler numerics in their fast form.
//
// The same MUSCL-Hancock / HLLC / FORCE formulas as fl_numerics.hpp, written
// for fewer instructions instead of the reference's rounding:
//   - each state's 1/rho is formed once (v_rcp_f64 and one Newton step,
//     within about an ulp) and every quotient by rho becomes a product; other
//     quotients are a * (1/b) the same way (FP32: v_rcp_f32);
//   - square roots take one Goldschmidt step and one correction after
//     v_rsq_f64 (FP32: v_sqrt_f32);
//   - the van Leer slope is its closed form: with r = b/f,
//     (b + f)/2 * 2 min(r, 1)/(1 + r) equals b when 0 < r <= 1 and f when
//     r > 1, i.e. minmod(b, f), so it needs no division at all;
//   - multiply-adds contract to FMAs.
// Results differ from the reference's in the last bits of each operation; the
// FP64 fast path is held to 1e-12 relative L-inf of the reference's results
// over a 100-step run.  FP32 uses these forms always.  Branches (limiter
// signs, flattening threshold, HLLC wave ordering) are the reference's.
#pragma once

#include "fl_numerics.hpp"

namespace fl {
namespace fast {

// Newton steps refining v_rcp_f64: one (about an ulp; tuning builds: 2).
#ifndef FL_FAST_NEWTON
#define FL_FAST_NEWTON 1
#endif

__device__ __forceinline__ double rcp(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
#if FL_FAST_NEWTON > 1
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
#endif
  return r;
}
__device__ __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }

// Square root.  FP64 with FL_FAST_SQRT (default): v_rsq_f64, one Goldschmidt
// step and one residual correction (the IEEE sequence's second correction and
// its half-inverse-root refinement left out); +-0 and +inf return x.
// FL_FAST_SQRT=0: the correctly rounded fl::sq.  FP32: v_sqrt_f32.
#ifndef FL_FAST_SQRT
#define FL_FAST_SQRT 1
#endif

__device__ __forceinline__ double sqt(double x) {
#if FL_FAST_SQRT
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  const double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  const double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return __builtin_amdgcn_class(x, 0x260) ? x : g;
#else
  return sq(x);
#endif
}
__device__ __forceinline__ float sqt(float x) { return sq(x); }

template <typename T> __device__ __forceinline__ T div(T a, T b) { return a * rcp(b); }

// Minimum and maximum as single v_min/v_max instructions.  They differ from
// the comparison-based fl::smin/smax only in which operand a NaN or a signed
// zero leaves: a NaN operand yields the other one, as std::min(max, NaN) and
// std::max(0, NaN) do where the reference uses them to clamp.
template <typename T> __device__ __forceinline__ T mn(T a, T b) { return __builtin_fmin(a, b); }
template <typename T> __device__ __forceinline__ T mx(T a, T b) { return __builtin_fmax(a, b); }

// Primitive quantities and 1/rho of one conservative state.
template <typename T, int ND> struct PrimR {
  T vel[ND];
  T p;
  T rinv;
};

template <typename T, int ND>
__device__ __forceinline__ PrimR<T, ND> prim_of(const T* U, const Material<T>& m) {
#pragma clang fp contract(fast)
  PrimR<T, ND> w;
  w.rinv = fast::rcp(U[0]);
  T sum = T(0);
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    w.vel[i] = U[2 + i] * w.rinv;
    sum += w.vel[i] * w.vel[i];
  }
  w.p = m.gm1 * (U[1] - T(0.5) * U[0] * sum);
  if (m.corr) w.p = w.p - m.gpinf;
  return w;
}

template <typename T, int ND>
__device__ __forceinline__ Prim<T, ND> as_prim(const PrimR<T, ND>& w) {
  Prim<T, ND> o;
#pragma unroll
  for (int i = 0; i < ND; ++i) o.vel[i] = w.vel[i];
  o.p = w.p;
  return o;
}

template <typename T>
__device__ __forceinline__ T sound_speed(T p, T rinv, const Material<T>& m) {
#pragma clang fp contract(fast)
  if (m.stiff) return fast::sqt(m.gamma * (p + m.p_inf) * rinv);
  return fast::sqt(m.gamma * p * rinv);
}

template <typename T, int ND, int D>
__device__ __forceinline__ void euler_flux(const T* U, const PrimR<T, ND>& w, T* F) {
#pragma clang fp contract(fast)
  const T v = w.vel[D];
  F[0] = U[2 + D];
  F[2 + D] = U[2 + D] * v + w.p;
  F[1] = v * (U[1] + w.p);
#pragma unroll
  for (int o = 0; o < ND; ++o)
    if (o != D) F[2 + o] = U[2 + o] * v;
}

template <typename T, int ND>
__device__ __forceinline__ T max_wavespeed(const PrimR<T, ND>& w, const Material<T>& m) {
  T mv = fabs(w.vel[0]);
#pragma unroll
  for (int i = 1; i < ND; ++i) mv = fast::mx(mv, T(fabs(w.vel[i])));
  return mv + fast::sound_speed(w.p, w.rinv, m);
}

// Superbee's limit (superbee_limiter.hpp:72-89) of the ratio a / b.
template <typename T>
__device__ __forceinline__ T superbee(T a, T b) {
#pragma clang fp contract(fast)
  const T r = fast::div(a, b);
  if (r <= T(0) || b == T(0)) return T(0);
  if (r <= T(0.5)) return T(2) * r;
  if (r <= T(1)) return T(1);
  return fast::mn(fast::div(T(2), T(1) + r), fast::mn(r, T(2)));
}

// The limited slope (back + fwrd)/2 * limit(back, fwrd) of one component,
// conservative form (limiting/limiter.hpp:79-97).  van Leer
// (van_leer_limiter.hpp:65-74) in closed form: minmod(back, fwrd).
template <typename T>
__device__ __forceinline__ T slope_cons(int limiter, T back, T fwrd) {
#pragma clang fp contract(fast)
  if (limiter == kLimSuperbee) return (back + fwrd) * T(0.5) * fast::superbee(back, fwrd);
  if (!(back * fwrd > T(0))) return T(0);
  return (fabs(back) <= fabs(fwrd)) ? back : fwrd;
}

// Primitive form: the limit of the primitive differences (rb, rf) scales the
// average of the conservative ones (limiting/limiter.hpp:79-97).
template <typename T>
__device__ __forceinline__ T slope_prim(int limiter, T back, T fwrd, T rb, T rf) {
#pragma clang fp contract(fast)
  const T avg = (back + fwrd) * T(0.5);
  if (limiter == kLimSuperbee) return avg * fast::superbee(rb, rf);
  if (!(rb * rf > T(0))) return T(0);
  // 2 min(r, 1) / (1 + r) with r = rb / rf: 2 rb / (rf + rb) or 2 rf / (rf + rb)
  const T m = (fabs(rb) <= fabs(rf)) ? rb : rf;
  return avg * fast::div(T(2) * m, rf + rb);
}

template <typename T>
__device__ __forceinline__ T chi_factor(T p_pprev, T p_prev, T p_next, T p_nnext, T v_prev,
                                        T v_next, const Material<T>& m) {
#pragma clang fp contract(fast)
  const T u_factor = v_prev - v_next;
  const T dp = fabs(p_next - p_prev);
  T pmin = fast::mn(p_next, p_prev);
  if (m.corr) pmin = pmin + m.p_inf;
  // dp / pmin > 0.33 without the quotient: for pmin > 0 it is dp > 0.33 pmin;
  // pmin = 0 gives +inf (dp > 0) or NaN (dp = 0); pmin < 0 a quotient <= 0;
  // a NaN operand fails every comparison.
  const bool flatten = u_factor > T(0) && pmin >= T(0) && dp > T(0.33) * pmin;
  if (!flatten) return T(1);
  const T z = fast::div(dp, T(fabs(p_nnext - p_pprev)));
  const T temp = T(1) - (z - T(0.75)) * T(10);
  return fast::mn(T(1), fast::mx(T(0), temp));
}

// Both evolved face states of one cell (reconstruction/muscl_reconstructor.hpp:
// 99-172), as fl::evolve_faces.
template <typename T, int ND, int D>
__device__ __forceinline__ void evolve_faces(const T* um, const T* u0, const T* up, T chi,
                                             T half_dtdh, int limiter, int form,
                                             const Material<T>& m, T* R, T* L) {
#pragma clang fp contract(fast)
  constexpr int S = 2 + ND;
  T rb[S], rf[S];
  if (form == kFormPrim) {
    const PrimR<T, ND> wm = fast::prim_of<T, ND>(um, m);
    const PrimR<T, ND> w0 = fast::prim_of<T, ND>(u0, m);
    const PrimR<T, ND> wp = fast::prim_of<T, ND>(up, m);
    rb[0] = u0[0] - um[0];
    rf[0] = up[0] - u0[0];
    rb[1] = w0.p - wm.p;
    rf[1] = wp.p - w0.p;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      rb[2 + i] = w0.vel[i] - wm.vel[i];
      rf[2 + i] = wp.vel[i] - w0.vel[i];
    }
  }
  const T cscale = chi * T(0.5);
  T plus[S], minus[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const T back = u0[k] - um[k];
    const T fwrd = up[k] - u0[k];
    const T lim = (form == kFormPrim) ? fast::slope_prim(limiter, back, fwrd, rb[k], rf[k])
                                      : fast::slope_cons(limiter, back, fwrd);
    const T delta = lim * cscale;
    plus[k] = u0[k] + delta;
    minus[k] = u0[k] - delta;
  }
  T fm[S], fp[S];
  const PrimR<T, ND> wmi = fast::prim_of<T, ND>(minus, m);
  const PrimR<T, ND> wpl = fast::prim_of<T, ND>(plus, m);
  fast::euler_flux<T, ND, D>(minus, wmi, fm);
  fast::euler_flux<T, ND, D>(plus, wpl, fp);
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const T dflux = (fm[k] - fp[k]) * half_dtdh;
    R[k] = plus[k] + dflux;
    L[k] = minus[k] + dflux;
  }
}

template <typename T>
__device__ __forceinline__ T hllc_wavespeed(T u, T p, T pstar, T sound, const Material<T>& m) {
#pragma clang fp contract(fast)
  if (pstar <= p) return u + sound;
  if (m.corr) return u + sound * fast::sqt(T(1) + m.q * (fast::div(pstar + m.p_inf, p + m.p_inf) - T(1)));
  return u + sound * fast::sqt(T(1) + m.q * (fast::div(pstar, p) - T(1)));
}

// HLLC (flux_method/flux_hllc.hpp:82-245), as fl::hllc_flux.
template <typename T, int ND, int D>
__device__ __forceinline__ void hllc_flux(const T* ul, const T* ur, const Material<T>& m,
                                          T* F) {
#pragma clang fp contract(fast)
  constexpr int S = 2 + ND;
  const PrimR<T, ND> wl = fast::prim_of<T, ND>(ul, m);
  const PrimR<T, ND> wr = fast::prim_of<T, ND>(ur, m);
  const T al = fast::sound_speed(wl.p, wl.rinv, m);
  const T ar = fast::sound_speed(wr.p, wr.rinv, m);
  const T vl = wl.vel[D], vr = wr.vel[D];
  const T p_pvrs = T(0.5) * ((wl.p + wr.p) - (T(0.25) * (vr - vl) * (ur[0] + ul[0]) * (al + ar)));
  const T p_star = m.corr ? fast::mx(-m.p_inf, p_pvrs) : fast::mx(T(0), p_pvrs);

  const T wsl = fast::hllc_wavespeed(vl, wl.p, p_star, -al, m);
  if (T(0) <= wsl) {
    fast::euler_flux<T, ND, D>(ul, wl, F);
    return;
  }
  const T wsr = fast::hllc_wavespeed(vr, wr.p, p_star, ar, m);
  if (T(0) >= wsr) {
    fast::euler_flux<T, ND, D>(ur, wr, F);
    return;
  }
  const T factorl = ul[0] * (wsl - vl);
  const T factorr = ur[0] * (wsr - vr);
  const T ws_star = fast::div(wr.p - wl.p + vl * factorl - vr * factorr, factorl - factorr);

  const bool left = T(0) <= ws_star;
  const T* st = left ? ul : ur;
  const PrimR<T, ND>& w = left ? wl : wr;
  const T ws = left ? wsl : wsr;
  T f[S];
  fast::euler_flux<T, ND, D>(st, w, f);
  const T state_factor = ws - w.vel[D];
  // 1/(ws - ws_star) and 1/state_factor from one reciprocal of their product
  const T rr = fast::rcp((ws - ws_star) * state_factor);
  const T scale_factor = state_factor * state_factor * rr;
  const T inv_sf = (ws - ws_star) * rr;
  T us[S];
#pragma unroll
  for (int k = 0; k < S; ++k) us[k] = st[k] * scale_factor;
  us[2 + D] = ws_star * st[0] * scale_factor;
  us[1] = scale_factor *
          (st[1] + st[0] * (ws_star - w.vel[D]) * (ws_star + w.p * w.rinv * inv_sf));
#pragma unroll
  for (int k = 0; k < S; ++k) F[k] = f[k] + (us[k] - st[k]) * ws;
}

template <typename T, int ND, int D>
__device__ __forceinline__ void force_flux(const T* ul, const T* ur, T dtdh, T dhdt,
                                           const Material<T>& m, T* F) {
#pragma clang fp contract(fast)
  constexpr int S = 2 + ND;
  const PrimR<T, ND> wl = fast::prim_of<T, ND>(ul, m);
  const PrimR<T, ND> wr = fast::prim_of<T, ND>(ur, m);
  T fl[S], fr[S], uri[S];
  fast::euler_flux<T, ND, D>(ul, wl, fl);
  fast::euler_flux<T, ND, D>(ur, wr, fr);
#pragma unroll
  for (int k = 0; k < S; ++k) uri[k] = (ul[k] + ur[k] + (fl[k] - fr[k]) * dtdh) * T(0.5);
  const PrimR<T, ND> wri = fast::prim_of<T, ND>(uri, m);
  T fri[S];
  fast::euler_flux<T, ND, D>(uri, wri, fri);
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const T lf = (fl[k] + fr[k] + (ul[k] - ur[k]) * dhdt) * T(0.5);
    F[k] = (lf + fri[k]) * T(0.5);
  }
}

template <typename T, int ND, int D>
__device__ __forceinline__ void lax_friedrichs_flux(const T* ul, const T* ur, T dhdt,
                                                    const Material<T>& m, T* F) {
#pragma clang fp contract(fast)
  constexpr int S = 2 + ND;
  const PrimR<T, ND> wl = fast::prim_of<T, ND>(ul, m);
  const PrimR<T, ND> wr = fast::prim_of<T, ND>(ur, m);
  T fl[S], fr[S];
  fast::euler_flux<T, ND, D>(ul, wl, fl);
  fast::euler_flux<T, ND, D>(ur, wr, fr);
#pragma unroll
  for (int k = 0; k < S; ++k) F[k] = (fl[k] + fr[k] + (ul[k] - ur[k]) * dhdt) * T(0.5);
}

template <typename T, int ND, int D>
__device__ __forceinline__ void richtmyer_flux(const T* ul, const T* ur, T dtdh,
                                               const Material<T>& m, T* F) {
#pragma clang fp contract(fast)
  constexpr int S = 2 + ND;
  const PrimR<T, ND> wl = fast::prim_of<T, ND>(ul, m);
  const PrimR<T, ND> wr = fast::prim_of<T, ND>(ur, m);
  T fl[S], fr[S], uri[S];
  fast::euler_flux<T, ND, D>(ul, wl, fl);
  fast::euler_flux<T, ND, D>(ur, wr, fr);
#pragma unroll
  for (int k = 0; k < S; ++k) uri[k] = (ul[k] + ur[k] + (fl[k] - fr[k]) * dtdh) * T(0.5);
  const PrimR<T, ND> wri = fast::prim_of<T, ND>(uri, m);
  fast::euler_flux<T, ND, D>(uri, wri, F);
}

template <typename T, int ND, int D, int FLUX>
__device__ __forceinline__ void numerical_flux(const T* ul, const T* ur, T dtdh, T dhdt,
                                               const Material<T>& m, T* F) {
  if constexpr (FLUX == kFluxHllc) {
    fast::hllc_flux<T, ND, D>(ul, ur, m, F);
  } else if constexpr (FLUX == kFluxLaxFriedrichs) {
    fast::lax_friedrichs_flux<T, ND, D>(ul, ur, dhdt, m, F);
  } else if constexpr (FLUX == kFluxRichtmyer) {
    fast::richtmyer_flux<T, ND, D>(ul, ur, dtdh, m, F);
  } else {
    fast::force_flux<T, ND, D>(ul, ur, dtdh, dhdt, m, F);
  }
}

}  // namespace fast
}  // namespace fl
