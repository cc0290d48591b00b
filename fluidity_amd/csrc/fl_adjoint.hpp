// fl_adjoint.hpp -- hand-written reverse-mode (adjoint) versions of the FP64
// per-cell numerics of fl_numerics.hpp: primitive conversion, sound speed,
// Euler flux, van Leer / Superbee limiting, the MUSCL-Hancock evolved face
// states (conservative-form limiting), the HLLC flux and the FORCE flux.
//
// Each *_adj function recomputes nothing that the caller already holds: it
// takes the primal inputs (and, where cheaper, the primal intermediates),
// follows exactly the branches the primal took, and ADDS the adjoint of its
// inputs into the g_* arguments, given the adjoint of its outputs.  The
// adjoint of the material enters through d/dgamma (gm1 = gamma - 1 and
// q = (gamma+1)/(2 gamma) depend on it; p_inf is a constant).
// The evolved face states also return the adjoint of their flattening
// factor chi; the caller carries it to the pressures chi was formed from.
//
// Primal quantities recomputed here (which decide branches) use the same
// correctly rounded divisions as the forward step.  Derivative terms that
// divide by one denominator share one reciprocal, recip(): they may differ
// from a division in the last bits, which the adjoint's tolerance admits.
#pragma once

#include "fl_numerics.hpp"

namespace fl {
namespace adj {

// 1/x for the derivative terms: v_rcp_f64 refined by two Newton steps
// (r' = r + r (1 - x r), each step squaring the relative error), within a
// few ulps of 1/x for finite nonzero x, which every caller guarantees.
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
  r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
  return r;
}

// Adjoint accumulators of one material's parameters.
struct GMat {
  double gamma;
};

// w = prim(U): vel_i = m_i / rho; p = gm1 (E - 0.5 rho sum vel_i^2).
template <int ND>
__device__ __forceinline__ void prim_adj(const double* U, const Prim<double, ND>& w,
                                         const Material<double>& m, const double* g_vel,
                                         double g_p, double* gU, GMat& gm) {
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < ND; ++i) sum += w.vel[i] * w.vel[i];
  gm.gamma += g_p * (U[1] - 0.5 * U[0] * sum);   // d gm1 / d gamma = 1
  gU[1] += g_p * m.gm1;
  const double t = -0.5 * g_p * m.gm1;
  gU[0] += t * sum;
  const double g_sum = t * U[0];
  const double inv_rho = recip(U[0]);
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const double gv = g_vel[i] + 2.0 * g_sum * w.vel[i];
    gU[2 + i] += gv * inv_rho;
    gU[0] -= gv * w.vel[i] * inv_rho;
  }
}

// a = sqrt(gamma (p [+ p_inf]) / rho).
__device__ __forceinline__ void sound_adj(double p, double rho, double a,
                                          const Material<double>& m, double g_a, double& g_p,
                                          double& g_rho, GMat& gm) {
  const double pp = m.stiff ? p + m.p_inf : p;
  const double g_s = g_a * recip(2.0 * a);
  const double inv_rho = recip(rho);
  gm.gamma += g_s * pp * inv_rho;
  g_p += g_s * m.gamma * inv_rho;
  g_rho -= g_s * (m.gamma * pp * inv_rho) * inv_rho;
}

// Euler flux in dimension D from U and its primitives; adds to gU and to the
// primitive adjoints (g_vel[D], g_p).
template <int ND, int D>
__device__ __forceinline__ void flux_adj(const double* U, const Prim<double, ND>& w,
                                         const double* gF, double* gU, double* g_vel,
                                         double& g_p) {
  const double v = w.vel[D];
  gU[2 + D] += gF[0] + gF[2 + D] * v;
  double gv = gF[2 + D] * U[2 + D] + gF[1] * (U[1] + w.p);
  g_p += gF[2 + D] + gF[1] * v;
  gU[1] += gF[1] * v;
#pragma unroll
  for (int o = 0; o < ND; ++o)
    if (o != D) {
      gU[2 + o] += gF[2 + o] * v;
      gv += gF[2 + o] * U[2 + o];
    }
  g_vel[D] += gv;
}

// Euler flux adjoint all the way to U (through the primitive conversion).
template <int ND, int D>
__device__ __forceinline__ void flux_full_adj(const double* U, const Prim<double, ND>& w,
                                              const Material<double>& m, const double* gF,
                                              double* gU, GMat& gm) {
  double g_vel[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) g_vel[i] = 0.0;
  double g_p = 0.0;
  flux_adj<ND, D>(U, w, gF, gU, g_vel, g_p);
  prim_adj<ND>(U, w, m, g_vel, g_p, gU, gm);
}

// phi = limit_single(limiter, L, R); adds dphi/dL * g, dphi/dR * g.
__device__ __forceinline__ void limit_adj(int limiter, double L, double R, double g,
                                          double& gL, double& gR) {
  const double r = L / R;
  double dphi_dr = 0.0;
  if (limiter == kLimSuperbee) {
    if (r <= 0.0 || R == 0.0) return;
    if (r <= 0.5) {
      dphi_dr = 2.0;
    } else if (r <= 1.0) {
      return;
    } else {
      const double Er = 2.0 / (1.0 + r);
      const double inner = smin(r, 2.0);
      if (inner < Er) {
        dphi_dr = (2.0 < r) ? 0.0 : 1.0;
      } else {
        dphi_dr = -2.0 * recip((1.0 + r) * (1.0 + r));
      }
    }
  } else {
    if (r <= 0.0 || R == 0.0) return;
    if (1.0 < r) {
      dphi_dr = -2.0 * recip((1.0 + r) * (1.0 + r));   // 2 / (1 + r)
    } else {
      dphi_dr = 2.0 * recip((1.0 + r) * (1.0 + r));    // 2 r / (1 + r)
    }
  }
  const double gr = g * dphi_dr;
  const double inv_R = recip(R);
  gL += gr * inv_R;
  gR -= gr * r * inv_R;
}

// Adjoint of evolve_faces<double, ND, D> with conservative-form limiting.
// Given the adjoints of the right and left evolved face states (gR, gL; either
// may be all zeros), adds to the adjoints of the stencil cells (gum, gu0,
// gup), of half_dtdh, of the flattening factor chi and of gamma.
template <int ND, int D>
__device__ __forceinline__ void evolve_adj(const double* um, const double* u0, const double* up,
                                           double chi, double half_dtdh, int limiter, int form,
                                           const Material<double>& m, const double* gR,
                                           const double* gL, double* gum, double* gu0,
                                           double* gup, double& g_half_dtdh, double& g_chi,
                                           GMat& gm) {
  constexpr int S = 2 + ND;
  const double cscale = chi * 0.5;
  const bool pf = form == kFormPrim;
  double plus[S], minus[S], back[S], fwrd[S], lim0[S], phi[S];
  // primitive form: the limiter reads differences of (rho, p, vel)
  // (limiting/limiting_impl.hpp:77-81, 142-146), as evolve_faces does
  Prim<double, ND> wm{}, w0{}, wp{};
  double bd[S], fd[S];
  if (pf) {
    wm = prim_of<double, ND>(um, m);
    w0 = prim_of<double, ND>(u0, m);
    wp = prim_of<double, ND>(up, m);
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
#pragma unroll
  for (int k = 0; k < S; ++k) {
    back[k] = u0[k] - um[k];
    fwrd[k] = up[k] - u0[k];
    lim0[k] = (back[k] + fwrd[k]) * 0.5;
    phi[k] = pf ? limit_single(limiter, bd[k], fd[k]) : limit_single(limiter, back[k], fwrd[k]);
    const double delta = (lim0[k] * phi[k]) * cscale;
    plus[k] = u0[k] + delta;
    minus[k] = u0[k] - delta;
  }
  const Prim<double, ND> wmi = prim_of<double, ND>(minus, m);
  const Prim<double, ND> wpl = prim_of<double, ND>(plus, m);
  double fm[S], fp[S];
  euler_flux<double, ND, D>(minus, wmi, fm);
  euler_flux<double, ND, D>(plus, wpl, fp);
  double g_plus[S], g_minus[S], g_fm[S], g_fp[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const double g_dflux = gR[k] + gL[k];
    g_plus[k] = gR[k];
    g_minus[k] = gL[k];
    g_fm[k] = g_dflux * half_dtdh;
    g_fp[k] = -g_dflux * half_dtdh;
    g_half_dtdh += g_dflux * (fm[k] - fp[k]);
  }
  flux_full_adj<ND, D>(minus, wmi, m, g_fm, g_minus, gm);
  flux_full_adj<ND, D>(plus, wpl, m, g_fp, g_plus, gm);
  double g_bd[S], g_fd[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    gu0[k] += g_plus[k] + g_minus[k];
    const double g_delta = g_plus[k] - g_minus[k];
    // delta = (lim0 phi) (chi / 2)
    g_chi += g_delta * (lim0[k] * phi[k]) * 0.5;
    const double g_lim = g_delta * cscale;
    const double g_lim0 = g_lim * phi[k];
    const double g_phi = g_lim * lim0[k];
    double g_back = 0.5 * g_lim0, g_fwrd = 0.5 * g_lim0;
    g_bd[k] = 0.0;
    g_fd[k] = 0.0;
    if (pf) limit_adj(limiter, bd[k], fd[k], g_phi, g_bd[k], g_fd[k]);
    else limit_adj(limiter, back[k], fwrd[k], g_phi, g_back, g_fwrd);
    gu0[k] += g_back - g_fwrd;
    gum[k] -= g_back;
    gup[k] += g_fwrd;
  }
  if (pf) {
    // density differences are conservative ones; p and vel through prim()
    gu0[0] += g_bd[0] - g_fd[0];
    gum[0] -= g_bd[0];
    gup[0] += g_fd[0];
    double gvm[ND], gv0[ND], gvp[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      gv0[i] = g_bd[2 + i] - g_fd[2 + i];
      gvm[i] = -g_bd[2 + i];
      gvp[i] = g_fd[2 + i];
    }
    prim_adj<ND>(um, wm, m, gvm, -g_bd[1], gum, gm);
    prim_adj<ND>(u0, w0, m, gv0, g_bd[1] - g_fd[1], gu0, gm);
    prim_adj<ND>(up, wp, m, gvp, g_fd[1], gup, gm);
  }
}

// Adjoint of hllc_wavespeed(u, p, pstar, sound, q) -> adds to g_u, g_p,
// g_pstar, g_sound, g_q.
__device__ __forceinline__ void hllc_wavespeed_adj(double u, double p, double pstar, double sound,
                                                   double q, double g, double& g_u, double& g_p,
                                                   double& g_pstar, double& g_sound,
                                                   double& g_q) {
  g_u += g;
  if (pstar <= p) {
    g_sound += g;
    return;
  }
  const double t = pstar / p;
  const double s = 1.0 + q * (t - 1.0);
  const double sr = sqrt(s);
  g_sound += g * sr;
  const double g_s = g * sound * recip(2.0 * sr);
  g_q += g_s * (t - 1.0);
  const double g_t = g_s * q;
  const double inv_p = recip(p);
  g_pstar += g_t * inv_p;
  g_p -= g_t * t * inv_p;
}

// Adjoint of hllc_flux<double, ND, D>(ul, ur, m, F) given gF; adds to gl, gr
// and gamma.
template <int ND, int D>
__device__ __forceinline__ void hllc_adj(const double* ul, const double* ur,
                                         const Material<double>& m, const double* gF,
                                         double* gl, double* gr, GMat& gm) {
  constexpr int S = 2 + ND;
  const Prim<double, ND> wl = prim_of<double, ND>(ul, m);
  const Prim<double, ND> wr = prim_of<double, ND>(ur, m);
  const double al = sound_speed(wl.p, ul[0], m);
  const double ar = sound_speed(wr.p, ur[0], m);
  const double vl = wl.vel[D], vr = wr.vel[D];
  const double p_pvrs =
      0.5 * ((wl.p + wr.p) - (0.25 * (vr - vl) * (ur[0] + ul[0]) * (al + ar)));
  const double p_star = smax(0.0, p_pvrs);
  const double wsl = hllc_wavespeed(vl, wl.p, p_star, -al, m);
  if (0.0 <= wsl) {
    flux_full_adj<ND, D>(ul, wl, m, gF, gl, gm);
    return;
  }
  const double wsr = hllc_wavespeed(vr, wr.p, p_star, ar, m);
  if (0.0 >= wsr) {
    flux_full_adj<ND, D>(ur, wr, m, gF, gr, gm);
    return;
  }
  const double factorl = ul[0] * (wsl - vl);
  const double factorr = ur[0] * (wsr - vr);
  const double num = wr.p - wl.p + vl * factorl - vr * factorr;
  const double den = factorl - factorr;
  const double ws_star = num / den;
  const bool left = 0.0 <= ws_star;
  const double* st = left ? ul : ur;
  const Prim<double, ND>& w = left ? wl : wr;
  const double ws = left ? wsl : wsr;
  double f[S];
  euler_flux<double, ND, D>(st, w, f);
  const double sf = ws - w.vel[D];
  const double dd = ws - ws_star;
  const double scale = sf / dd;
  double us[S];
#pragma unroll
  for (int k = 0; k < S; ++k) us[k] = st[k] * scale;
  us[2 + D] = ws_star * st[0] * scale;
  const double A = ws_star - w.vel[D];
  const double den2 = st[0] * sf;
  const double Bq = w.p / den2;
  const double Bt = ws_star + Bq;
  const double inner = st[1] + st[0] * A * Bt;
  us[1] = scale * inner;

  // F[k] = f[k] + (us[k] - st[k]) * ws
  double g_st[S], g_vel[ND];
#pragma unroll
  for (int k = 0; k < S; ++k) g_st[k] = 0.0;
#pragma unroll
  for (int i = 0; i < ND; ++i) g_vel[i] = 0.0;
  double g_p = 0.0, g_ws = 0.0, g_wsstar = 0.0, g_scale = 0.0, g_sf = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    g_ws += gF[k] * (us[k] - st[k]);
    g_st[k] -= gF[k] * ws;
  }
  // us
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const double g_us = gF[k] * ws;
    if (k == 1) {
      g_scale += g_us * inner;
      const double g_inner = g_us * scale;
      g_st[1] += g_inner;
      g_st[0] += g_inner * A * Bt;
      const double g_A = g_inner * st[0] * Bt;
      const double g_Bt = g_inner * st[0] * A;
      g_wsstar += g_A + g_Bt;
      g_vel[D] -= g_A;
      const double inv_den2 = recip(den2);
      g_p += g_Bt * inv_den2;
      const double g_den2 = -g_Bt * Bq * inv_den2;
      g_st[0] += g_den2 * sf;
      g_sf += g_den2 * st[0];
    } else if (k == 2 + D) {
      g_wsstar += g_us * st[0] * scale;
      g_st[0] += g_us * ws_star * scale;
      g_scale += g_us * ws_star * st[0];
    } else {
      g_st[k] += g_us * scale;
      g_scale += g_us * st[k];
    }
  }
  // scale = sf / (ws - ws_star)
  const double inv_dd = recip(dd);
  const double g_dd = g_scale * scale * inv_dd;
  g_sf += g_scale * inv_dd;
  g_ws -= g_dd;
  g_wsstar += g_dd;
  // sf = ws - v
  g_ws += g_sf;
  g_vel[D] -= g_sf;
  // f = F(st)
  flux_adj<ND, D>(st, w, gF, g_st, g_vel, g_p);

  // per-side primitive adjoints
  double g_vl[ND], g_vr[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    g_vl[i] = left ? g_vel[i] : 0.0;
    g_vr[i] = left ? 0.0 : g_vel[i];
  }
  double g_pl = left ? g_p : 0.0, g_pr = left ? 0.0 : g_p;
  double g_rl = 0.0, g_rr = 0.0;
  if (left) {
#pragma unroll
    for (int k = 0; k < S; ++k) gl[k] += g_st[k];
  } else {
#pragma unroll
    for (int k = 0; k < S; ++k) gr[k] += g_st[k];
  }
  double g_wsl = left ? g_ws : 0.0, g_wsr = left ? 0.0 : g_ws;
  // ws_star = num / den
  const double inv_den = recip(den);
  const double g_num = g_wsstar * inv_den;
  const double g_den = -g_wsstar * ws_star * inv_den;
  g_pr += g_num;
  g_pl -= g_num;
  g_vl[D] += g_num * factorl;
  g_vr[D] -= g_num * factorr;
  const double g_fl = g_num * vl + g_den;
  const double g_fr = -g_num * vr - g_den;
  // factorl = rho_l (wsl - vl), factorr = rho_r (wsr - vr)
  g_rl += g_fl * (wsl - vl);
  g_wsl += g_fl * ul[0];
  g_vl[D] -= g_fl * ul[0];
  g_rr += g_fr * (wsr - vr);
  g_wsr += g_fr * ur[0];
  g_vr[D] -= g_fr * ur[0];
  // wavespeeds
  double g_pstar = 0.0, g_q = 0.0, g_al = 0.0, g_ar = 0.0, g_msound = 0.0;
  hllc_wavespeed_adj(vl, wl.p, p_star, -al, m.q, g_wsl, g_vl[D], g_pl, g_pstar, g_msound, g_q);
  g_al -= g_msound;
  hllc_wavespeed_adj(vr, wr.p, p_star, ar, m.q, g_wsr, g_vr[D], g_pr, g_pstar, g_ar, g_q);
  // q = (gamma + 1) / (2 gamma)
  gm.gamma += g_q * (-0.5 / (m.gamma * m.gamma));
  // p_star = max(0, p_pvrs)
  if (0.0 < p_pvrs) {
    const double g_pv = g_pstar;
    g_pl += 0.5 * g_pv;
    g_pr += 0.5 * g_pv;
    const double c = -0.125 * g_pv;
    const double dv_ = vr - vl, rs = ur[0] + ul[0], as = al + ar;
    const double g_dv = c * rs * as;
    g_vr[D] += g_dv;
    g_vl[D] -= g_dv;
    const double g_rs = c * dv_ * as;
    g_rl += g_rs;
    g_rr += g_rs;
    const double g_as = c * dv_ * rs;
    g_al += g_as;
    g_ar += g_as;
  }
  // sound speeds
  sound_adj(wl.p, ul[0], al, m, g_al, g_pl, g_rl, gm);
  sound_adj(wr.p, ur[0], ar, m, g_ar, g_pr, g_rr, gm);
  gl[0] += g_rl;
  gr[0] += g_rr;
  prim_adj<ND>(ul, wl, m, g_vl, g_pl, gl, gm);
  prim_adj<ND>(ur, wr, m, g_vr, g_pr, gr, gm);
}

// Adjoint of the cell wavespeed max_d |u_d| + a (state/state_impl.hpp:167-178).
// FORCE flux (force_flux, flux_method/flux_force.hpp:73-84):
// F = (LF + RI) / 2 with LF = (fl + fr + (ul - ur) dhdt) / 2, RI = flux(uri),
// uri = (ul + ur + (fl - fr) dtdh) / 2 and dhdt = 1 / dtdh.  Adds the adjoints
// of ul, ur, dtdh (directly and through dhdt) and gamma, given gF.
template <int ND, int D>
__device__ __forceinline__ void force_adj(const double* ul, const double* ur, double dtdh,
                                          const Material<double>& m, const double* gF,
                                          double* gL, double* gR, double& g_dtdh, GMat& gm) {
  constexpr int S = 2 + ND;
  const double dhdt = 1.0 / dtdh;
  const Prim<double, ND> wl = prim_of<double, ND>(ul, m);
  const Prim<double, ND> wr = prim_of<double, ND>(ur, m);
  double fl[S], fr[S], uri[S];
  euler_flux<double, ND, D>(ul, wl, fl);
  euler_flux<double, ND, D>(ur, wr, fr);
#pragma unroll
  for (int k = 0; k < S; ++k) uri[k] = (ul[k] + ur[k] + (fl[k] - fr[k]) * dtdh) * 0.5;
  const Prim<double, ND> wri = prim_of<double, ND>(uri, m);
  // adjoints of LF and of RI (each half of F), then of uri through RI's flux
  double g_half[S], g_uri[S], g_fl[S], g_fr[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    g_half[k] = 0.5 * gF[k];
    g_uri[k] = 0.0;
  }
  flux_full_adj<ND, D>(uri, wri, m, g_half, g_uri, gm);
  double g_dhdt = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const double gl = 0.5 * g_half[k];  // LF's terms
    const double gu = 0.5 * g_uri[k];   // uri's terms
    gL[k] += gu + gl * dhdt;
    gR[k] += gu - gl * dhdt;
    g_fl[k] = gl + gu * dtdh;
    g_fr[k] = gl - gu * dtdh;
    g_dtdh += gu * (fl[k] - fr[k]);
    g_dhdt += gl * (ul[k] - ur[k]);
  }
  flux_full_adj<ND, D>(ul, wl, m, g_fl, gL, gm);
  flux_full_adj<ND, D>(ur, wr, m, g_fr, gR, gm);
  g_dtdh -= g_dhdt * dhdt * dhdt;  // d(1/x)/dx = -1/x^2
}

template <int ND>
__device__ __forceinline__ void wavespeed_adj(const double* U, const Material<double>& m,
                                              double g, double* gU, GMat& gm) {
  const Prim<double, ND> w = prim_of<double, ND>(U, m);
  double mv = fabs(w.vel[0]);
  int dmax = 0;
#pragma unroll
  for (int i = 1; i < ND; ++i) {
    const double a = fabs(w.vel[i]);
    if (mv < a) {
      mv = a;
      dmax = i;
    }
  }
  double g_vel[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) g_vel[i] = 0.0;
  const double v = w.vel[dmax];
  g_vel[dmax] = g * ((0.0 < v) - (v < 0.0));
  const double a = sound_speed(w.p, U[0], m);
  double g_p = 0.0, g_rho = 0.0;
  sound_adj(w.p, U[0], a, m, g, g_p, g_rho, gm);
  gU[0] += g_rho;
  prim_adj<ND>(U, w, m, g_vel, g_p, gU, gm);
}

}  // namespace adj
}  // namespace fl
