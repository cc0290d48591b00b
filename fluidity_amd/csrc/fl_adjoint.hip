// fl_adjoint.hip -- gfx950 kernels of the hand-written adjoint of the split
// MUSCL-Hancock + HLLC step (FP64).
//
// Forward: out_i = U_i + dtdh (G_{i-1/2} - G_{i+1/2}), G_{i+1/2} =
// HLLC(R_i, L_{i+1}), R_i / L_i the evolved face states of cell i from cells
// i-1..i+1 (solver/split_solver.hpp:138-162,
// reconstruction/muscl_reconstructor.hpp:99-172, flux_method/flux_hllc.hpp).
// Adjoint of one sweep, given lambda = dLoss/d out:
//  * face kernel, one thread per face i+1/2: recompute R_i, L_{i+1} and G,
//    mu = dtdh (lambda_{i+1} - lambda_i) (ghost cells carry no lambda), and
//    back-propagate mu through HLLC and both evolutions into the adjoints of
//    the four stencil cells i-1..i+2 (stored per face and slot), of dt/dh and
//    of gamma (stored per face);
//  * gather kernel, one thread per cell: lambda_in = lambda + the slot
//    contributions of every face whose stencil holds the cell, also through
//    ghost cells (periodic wrap, mirror, reflective flip), in a fixed order;
//  * the per-face dt/dh and gamma adjoints are summed in a fixed-order
//    two-level reduction, so the whole adjoint is deterministic.
// dt = cfl dh / max wavespeed couples the step to the state through the
// wavespeed of the cell holding the maximum (lowest index among ties); the
// step's finalize kernel adds that term.
#include "fl_adjoint.hpp"
#include "fl_adjoint_kernels.hpp"

namespace fl {
namespace adj {

__device__ __forceinline__ int64_t ghost_index(int64_t i, int64_t n, int bc, bool& flip) {
  i = i < -4 ? -4 : (i > n + 3 ? n + 3 : i);
  flip = false;
  if (i >= 0 && i < n) return i;
  if (bc == kBcPeriodic) return i < 0 ? i + n : i - n;
  flip = (bc == kBcReflective);
  return i < 0 ? -i - 1 : 2 * n - 1 - i;
}

// State at (x, y) with the ghost rule of each dimension (reflective ghosts
// get the velocity flip of the forward sweep).
template <int ND>
__device__ __forceinline__ void load_state(const AGrid& g, int64_t x, int64_t y, double* u) {
  bool fx = false, fy = false;
  const int64_t xi = ghost_index(x, g.nx, g.bcx, fx);
  const int64_t yi = ND > 1 ? ghost_index(y, g.ny, g.bcy, fy) : 0;
  const double* p = g.U + (yi * g.nx + xi) * g.es;
#pragma unroll
  for (int k = 0; k < 2 + ND; ++k) u[k] = p[k * g.cs];
  if (fx) reflect<double, ND>(u, 0);
  if (ND > 1 && fy) reflect<double, ND>(u, 1);
}

// Pressure and velocities of every position within 4 ghost cells of the grid
// (ghost rule and reflective flip applied), on the padded grid P (x fastest,
// pad 4 in each dimension of the grid).
template <int ND>
__global__ void __launch_bounds__(256) prim_pad_kernel(AGrid g, Material<double> m,
                                                       double* __restrict__ P,
                                                       double* __restrict__ V0,
                                                       double* __restrict__ V1) {
  const int64_t px = g.nx + 8, py = ND > 1 ? g.ny + 8 : 1;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= px * py) return;
  const int64_t x = q % px - 4, y = ND > 1 ? q / px - 4 : 0;
  double u[2 + ND];
  load_state<ND>(g, x, y, u);
  const Prim<double, ND> w = prim_of<double, ND>(u, m);
  P[q] = w.p;
  V0[q] = w.vel[0];
  if (ND > 1) V1[q] = w.vel[ND - 1];
}

// chi_v of MHReconstructor::evolve (reconstruction/muscl_reconstructor.hpp:59-87,
// 106-118) -- the minimum over dimensions of chi about the cell and about its
// neighbour on the higher-pressure side -- for every position within one
// ghost cell of the grid, from the padded primitives.
template <int ND>
__global__ void __launch_bounds__(256) chi_pad_kernel(AGrid g, const double* __restrict__ P,
                                                      const double* __restrict__ V0,
                                                      const double* __restrict__ V1,
                                                      double* __restrict__ chi,
                                                      int* __restrict__ any) {
  const int64_t px = g.nx + 8;
  const int64_t cx = g.nx + 2, cy = ND > 1 ? g.ny + 2 : 1;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0) *any = 0;  // chi_adj_kernel's record flag, read by chi_gather_kernel
  if (q >= cx * cy) return;
  const int64_t x = q % cx - 1, y = ND > 1 ? q / cx - 1 : 0;
  auto at = [&](const double* A, int64_t xx, int64_t yy) {
    return A[(ND > 1 ? (yy + 4) * px : 0) + xx + 4];
  };
  const Material<double> no_shift{};  // the adjoint runs the reference's equations of state
  auto chi_d = [&](int64_t xx, int64_t yy, int d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    const double* V = d == 0 ? V0 : V1;
    return chi_factor<double>(at(P, xx - 2 * dx, yy - 2 * dy), at(P, xx - dx, yy - dy),
                              at(P, xx + dx, yy + dy), at(P, xx + 2 * dx, yy + 2 * dy),
                              at(V, xx - dx, yy - dy), at(V, xx + dx, yy + dy), no_shift);
  };
  double c = DBL_MAX;
  for (int d = 0; d < ND; ++d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    const int sign = signum(at(P, x + dx, y + dy) - at(P, x - dx, y - dy));
    c = smin(c, smin(chi_d(x, y, d), chi_d(x + sign * dx, y + sign * dy, d)));
  }
  chi[q] = c;
}

// PRIM: primitive-form limiting (a template parameter, so that the
// conservative-form kernel carries none of its code).
template <int ND, int D, int FLUX, bool PRIM>
__global__ void __launch_bounds__(256) face_adj_kernel(AGrid g, const double* __restrict__ lam,
                                                       Material<double> m, int limiter,
                                                       const StepRec* __restrict__ rec,
                                                       const double* __restrict__ chi_pad,
                                                       double* __restrict__ contrib,
                                                       double* __restrict__ face_g,
                                                       double* __restrict__ chi_g) {
  constexpr int S = 2 + ND;
  constexpr int form = PRIM ? kFormPrim : kFormCons;
  const int64_t n = D == 0 ? g.nx : g.ny;          // cells along the sweep
  const int64_t lines = D == 0 ? g.ny : g.nx;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (n + 1) * lines) return;
  int64_t i, line;
  if (D == 0) {
    i = f % (n + 1) - 1;
    line = f / (n + 1);
  } else {
    i = f / lines - 1;
    line = f % lines;
  }
  auto pos = [&](int64_t k, int64_t& x, int64_t& y) {
    if (D == 0) {
      x = k;
      y = line;
    } else {
      x = line;
      y = k;
    }
  };
  const double dtdh = rec->dtdh;
  const double half_dtdh = 0.5 * dtdh;
  double u[4][S];
  for (int s = 0; s < 4; ++s) {
    int64_t x, y;
    pos(i - 1 + s, x, y);
    load_state<ND>(g, x, y, u[s]);
  }
  int64_t xa, ya, xb, yb;
  pos(i, xa, ya);
  pos(i + 1, xb, yb);
  const int64_t cxw = g.nx + 2;
  const double chi_a = chi_pad[(ND > 1 ? (ya + 1) * cxw : 0) + xa + 1];
  const double chi_b = chi_pad[(ND > 1 ? (yb + 1) * cxw : 0) + xb + 1];
  double R[S], L[S], tmp[S], G[S];
  evolve_faces<double, ND, D>(u[0], u[1], u[2], chi_a, half_dtdh, limiter, form, m, R, tmp);
  evolve_faces<double, ND, D>(u[1], u[2], u[3], chi_b, half_dtdh, limiter, form, m, tmp, L);
  if constexpr (FLUX == kFluxHllc) hllc_flux<double, ND, D>(R, L, m, G);
  else force_flux<double, ND, D>(R, L, dtdh, 1.0 / dtdh, m, G);
  double la[S], lb[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    la[k] = 0.0;
    lb[k] = 0.0;
  }
  if (i >= 0) {
    const double* p = lam + (ya * g.nx + xa) * g.es;
#pragma unroll
    for (int k = 0; k < S; ++k) la[k] = p[k * g.cs];
  }
  if (i + 1 < n) {
    const double* p = lam + (yb * g.nx + xb) * g.es;
#pragma unroll
    for (int k = 0; k < S; ++k) lb[k] = p[k * g.cs];
  }
  double mu[S], g_dtdh = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const double dl = lb[k] - la[k];
    mu[k] = dtdh * dl;
    g_dtdh += G[k] * dl;
  }
  GMat gm{0.0};
  double gR[S], gL[S], zero[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    gR[k] = 0.0;
    gL[k] = 0.0;
    zero[k] = 0.0;
  }
  if constexpr (FLUX == kFluxHllc) hllc_adj<ND, D>(R, L, m, mu, gR, gL, gm);
  else force_adj<ND, D>(R, L, dtdh, m, mu, gR, gL, g_dtdh, gm);
  double c[4][S];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int k = 0; k < S; ++k) c[s][k] = 0.0;
  double g_half = 0.0, g_chi_a = 0.0, g_chi_b = 0.0;
  evolve_adj<ND, D>(u[0], u[1], u[2], chi_a, half_dtdh, limiter, form, m, gR, zero, c[0], c[1],
                    c[2], g_half, g_chi_a, gm);
  evolve_adj<ND, D>(u[1], u[2], u[3], chi_b, half_dtdh, limiter, form, m, zero, gL, c[1], c[2],
                    c[3], g_half, g_chi_b, gm);
  chi_g[2 * f] = g_chi_a;      // of cell i (its right face state)
  chi_g[2 * f + 1] = g_chi_b;  // of cell i + 1 (its left face state)
  // contributions as planes (slot, component) x faces: coalesced both here and
  // in the gather
  const int64_t nf = (n + 1) * lines;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int k = 0; k < S; ++k) contrib[(s * S + k) * nf + f] = c[s][k];
  face_g[2 * f] = g_dtdh + 0.5 * g_half;
  face_g[2 * f + 1] = gm.gamma;
}

// lambda_in = lambda_out + contributions of every face whose stencil holds
// the cell (directly or as a ghost alias).
template <int ND, int D>
__global__ void __launch_bounds__(256) gather_adj_kernel(AGrid g, const double* __restrict__ lam_out,
                                                         double* __restrict__ lam_in,
                                                         const double* __restrict__ contrib) {
  constexpr int S = 2 + ND;
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.nx * g.ny) return;
  const int64_t x = cell % g.nx, y = cell / g.nx;
  const int64_t n = D == 0 ? g.nx : g.ny;
  const int64_t lines = D == 0 ? g.ny : g.nx;
  const int64_t j = D == 0 ? x : y;
  const int64_t line = D == 0 ? y : x;
  const int bc = D == 0 ? g.bcx : g.bcy;
  double acc[S];
  const double* lp = lam_out + cell * g.es;
#pragma unroll
  for (int k = 0; k < S; ++k) acc[k] = lp[k * g.cs];
  // virtual positions holding this cell: itself, then ghost aliases
  int64_t vs[3];
  int flips[3];
  int nv = 0;
  vs[nv] = j;
  flips[nv++] = 0;
  if (bc == kBcPeriodic) {
    if (j >= n - 2) { vs[nv] = j - n; flips[nv++] = 0; }
    if (j <= 1) { vs[nv] = j + n; flips[nv++] = 0; }
  } else {
    const int fl = bc == kBcReflective ? 1 : 0;
    if (j <= 1) { vs[nv] = -j - 1; flips[nv++] = fl; }
    if (j >= n - 2) { vs[nv] = 2 * n - 1 - j; flips[nv++] = fl; }
  }
  for (int a = 0; a < nv; ++a) {
    const int64_t v = vs[a];
    for (int s = 0; s < 4; ++s) {
      const int64_t i = v + 1 - s;
      if (i < -1 || i > n - 1) continue;
      const int64_t f = D == 0 ? line * (n + 1) + (i + 1) : (i + 1) * lines + line;
      const int64_t nf = (n + 1) * lines;
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const double ck = contrib[(s * S + k) * nf + f];
        acc[k] += (flips[a] && k == 2 + D) ? -ck : ck;
      }
    }
  }
  double* o = lam_in + cell * g.es;
#pragma unroll
  for (int k = 0; k < S; ++k) o[k * g.cs] = acc[k];
}

// ---- the flattening factor's adjoint ----------------------------------------
// chi_v of a cell (chi_pad_kernel) is the smallest of the flattening factors
// K_d(j) of the cell and of its neighbour on the higher-pressure side in each
// dimension d; where the selected factor is flattened and not clamped,
// K = 1 - (z - 0.75) / 0.1 with z = |p(j+1) - p(j-1)| / |p(j+2) - p(j-2)|
// along d carries the gradient (muscl_reconstructor.hpp:59-87); the
// flattening test, the neighbour and the minimum are branches.

// At each chi position (the grid padded by one cell) whose chi lies strictly
// between 0 and 1 (elsewhere the selected factor is clamped or not
// flattened and carries no gradient): the adjoint of chi, the sum of what the
// position's right face state (face i+1/2) and left face state (face i-1/2)
// passed back in sweep dimension D; which factor the minimum selected,
// rec_k = 4 d + (j - c along d) + 1 for K_d(j) (-1: none); and the adjoints
// of the pressures at j-2, j-1, j+1, j+2 along d (rec_v, 4 per position).
// *any is set when some position has a record.
template <int ND, int D>
__global__ void __launch_bounds__(256) chi_adj_kernel(AGrid g, const double* __restrict__ P,
                                                      const double* __restrict__ V0,
                                                      const double* __restrict__ V1,
                                                      const double* __restrict__ chi,
                                                      const double* __restrict__ chi_g,
                                                      int* __restrict__ rec_k,
                                                      double* __restrict__ rec_v,
                                                      int* __restrict__ any) {
  const int64_t px = g.nx + 8;
  const int64_t cx = g.nx + 2, cy = ND > 1 ? g.ny + 2 : 1;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cx * cy) return;
  const int64_t x = q % cx - 1, y = ND > 1 ? q / cx - 1 : 0;
  const double cv = chi[q];
  double G = 0.0;
  if (cv > 0.0 && cv < 1.0) {
    const int64_t n = D == 0 ? g.nx : g.ny;
    const int64_t lines = D == 0 ? g.ny : g.nx;
    const int64_t i = D == 0 ? x : y;     // along the sweep: -1 .. n
    const int64_t line = D == 0 ? y : x;  // faces exist on interior lines only
    if (line >= 0 && line < lines) {
      auto face = [&](int64_t ii) {
        return D == 0 ? line * (n + 1) + (ii + 1) : (ii + 1) * lines + line;
      };
      if (i <= n - 1) G += chi_g[2 * face(i)];
      if (i >= 0) G += chi_g[2 * face(i - 1) + 1];
    }
  }
  auto at = [&](const double* A, int64_t xx, int64_t yy) {
    return A[(ND > 1 ? (yy + 4) * px : 0) + xx + 4];
  };
  const Material<double> no_shift{};
  auto chi_d = [&](int64_t xx, int64_t yy, int d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    const double* V = d == 0 ? V0 : V1;
    return chi_factor<double>(at(P, xx - 2 * dx, yy - 2 * dy), at(P, xx - dx, yy - dy),
                              at(P, xx + dx, yy + dy), at(P, xx + 2 * dx, yy + 2 * dy),
                              at(V, xx - dx, yy - dy), at(V, xx + dx, yy + dy), no_shift);
  };
  int best = -1;
  if (G != 0.0) {
    // the selection of chi_pad_kernel: smin(chi, smin(K(c), K(c + s)))
    double c = DBL_MAX;
    for (int d = 0; d < ND; ++d) {
      const int64_t dx = (d == 0), dy = (d == 1);
      const int sign = signum(at(P, x + dx, y + dy) - at(P, x - dx, y - dy));
      const double k0 = chi_d(x, y, d);
      const double k1 = chi_d(x + sign * dx, y + sign * dy, d);
      double mv = k0;
      int sel = 4 * d + 1;
      if (k1 < k0) {
        mv = k1;
        sel = 4 * d + sign + 1;
      }
      if (mv < c) {
        c = mv;
        best = sel;
      }
    }
  }
  if (best >= 0) {
    const int d = best / 4, jo = best % 4 - 1;
    const int64_t dx = (d == 0), dy = (d == 1);
    const int64_t jx = x + jo * dx, jy = y + jo * dy;
    const double* V = d == 0 ? V0 : V1;
    const double p_pp = at(P, jx - 2 * dx, jy - 2 * dy), p_p = at(P, jx - dx, jy - dy);
    const double p_n = at(P, jx + dx, jy + dy), p_nn = at(P, jx + 2 * dx, jy + 2 * dy);
    const double u_factor = at(V, jx - dx, jy - dy) - at(V, jx + dx, jy + dy);
    const double dp = fabs(p_n - p_p);
    const double pmin = smin(p_n, p_p);
    const bool flatten = u_factor > 0.0 && dv(dp, pmin) > 0.33;
    const double den = fabs(p_nn - p_pp);
    if (flatten && den > 0.0) {
      const double z = dv(dp, den);
      const double temp = 1.0 - dv(z - 0.75, 0.85 - 0.75);
      if (temp > 0.0 && temp < 1.0) {
        const double g_z = -G / (0.85 - 0.75);
        const double g_dp = g_z / den;
        const double g_den = -g_z * z / den;
        const double sn = (double)signum(p_n - p_p), snn = (double)signum(p_nn - p_pp);
        rec_v[4 * q + 0] = -g_den * snn;  // p(j-2)
        rec_v[4 * q + 1] = -g_dp * sn;    // p(j-1)
        rec_v[4 * q + 2] = g_dp * sn;     // p(j+1)
        rec_v[4 * q + 3] = g_den * snn;   // p(j+2)
        rec_k[q] = best;
        *any = 1;
        return;
      }
    }
  }
  rec_k[q] = -1;
}

// Positions within 4 ghost cells of a line of n cells whose ghost rule maps
// to cell x (x itself first), at most 3.
__device__ __forceinline__ int alias_positions(int64_t x, int64_t n, int bc, int64_t* pos) {
  int k = 0;
  pos[k++] = x;
  int64_t cand[2];
  if (bc == kBcPeriodic) {
    cand[0] = x - n;
    cand[1] = x + n;
  } else {
    cand[0] = -x - 1;
    cand[1] = 2 * n - 1 - x;
  }
  for (int a = 0; a < 2; ++a) {
    const int64_t c = cand[a];
    if (c < -4 || c > n + 3 || (c >= 0 && c < n)) continue;
    bool flip;
    if (ghost_index(c, n, bc, flip) == x) pos[k++] = c;
  }
  return k;
}

// lambda += (dp/dU) times the pressure adjoints that the chi positions'
// selected factors pass to each cell (at the cell and at its ghost aliases,
// in a fixed order), and the gamma adjoint of the same terms (dp/dgamma)
// summed per workgroup into partials[2 b + 1] (partials[2 b] = 0).
// kPartialBlocks workgroups of 256 threads stride over the cells, so the sums
// have a fixed order; with no record at all (*any == 0) they only write
// zeros.
template <int ND>
__global__ void __launch_bounds__(256) chi_gather_kernel(AGrid g, Material<double> m,
                                                         const int* __restrict__ rec_k,
                                                         const double* __restrict__ rec_v,
                                                         const int* __restrict__ any,
                                                         double* __restrict__ lam,
                                                         double* __restrict__ partials) {
  constexpr int S = 2 + ND;
  __shared__ double red[256];
  double acc = 0.0;
  const int64_t cx = g.nx + 2;
  if (*any) {
    for (int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x; cell < g.nx * g.ny;
         cell += (int64_t)kPartialBlocks * 256) {
      const int64_t x = cell % g.nx, y = cell / g.nx;
      int64_t ax[3], ay[3] = {0, 0, 0};
      const int nax = alias_positions(x, g.nx, g.bcx, ax);
      const int nay = ND > 1 ? alias_positions(y, g.ny, g.bcy, ay) : 1;
      double gp = 0.0;
      for (int a = 0; a < nax; ++a)
        for (int b = 0; b < nay; ++b)
          for (int d = 0; d < ND; ++d) {
            const int64_t dx = (d == 0), dy = (d == 1);
            // a record at position c selecting K_d(c + jo) reaches the
            // pressure at c + jo + k, k in {-2, -1, 1, 2}: here t = jo + k
            for (int t = -3; t <= 3; ++t) {
              const int64_t qx = ax[a] - t * dx, qy = ay[b] - t * dy;
              if (qx < -1 || qx > g.nx || (ND > 1 && (qy < -1 || qy > g.ny))) continue;
              const int64_t qc = (ND > 1 ? (qy + 1) * cx : 0) + qx + 1;
              const int r = rec_k[qc];
              if (r < 0 || r / 4 != d) continue;
              const int k = t - (r % 4 - 1);
              if (k < -2 || k > 2 || k == 0) continue;
              gp += rec_v[4 * qc + (k < 0 ? k + 2 : k + 1)];
            }
          }
      if (gp != 0.0) {
        double u[S];
        const double* p = g.U + cell * g.es;
#pragma unroll
        for (int k = 0; k < S; ++k) u[k] = p[k * g.cs];
        // p = (gamma - 1) (E - rho |v|^2 / 2), v = m / rho
        const Prim<double, ND> w = prim_of<double, ND>(u, m);
        double v2 = 0.0;
#pragma unroll
        for (int i = 0; i < ND; ++i) v2 += w.vel[i] * w.vel[i];
        double* l = lam + cell * g.es;
        l[0] += gp * (m.gm1 * 0.5 * v2);
        l[1 * g.cs] += gp * m.gm1;
#pragma unroll
        for (int i = 0; i < ND; ++i) l[(2 + i) * g.cs] -= gp * (m.gm1 * w.vel[i]);
        acc += gp * (u[1] - 0.5 * u[0] * v2);
      }
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = 0.0;
    partials[2 * blockIdx.x + 1] = red[0];
  }
}

// Per-block sums of the per-face (dt/dh, gamma) adjoints; kPartialBlocks
// blocks of 256 threads, fixed order.
__global__ void __launch_bounds__(256) partial_sum_kernel(const double* __restrict__ face_g,
                                                          int64_t nfaces,
                                                          double* __restrict__ partials) {
  __shared__ double s0[256], s1[256];
  double a0 = 0.0, a1 = 0.0;
  for (int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x; f < nfaces;
       f += (int64_t)kPartialBlocks * 256) {
    a0 += face_g[2 * f];
    a1 += face_g[2 * f + 1];
  }
  s0[threadIdx.x] = a0;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s0[0];
    partials[2 * blockIdx.x + 1] = s1[0];
  }
}

// Records this step's max wavespeed and dt/dh from the control block and
// clears the argmax slot.
__global__ void record_kernel(const StepCtl* ctl, int parity, StepRec* rec) {
  rec->maxws_bits = ctl->ws_bits[parity];
  rec->dtdh = ctl->dtdh;
  rec->argmax = ~0ull;
}

// Lowest index of a cell whose wavespeed bits equal the recorded maximum.
template <int ND>
__global__ void __launch_bounds__(256) argmax_kernel(AGrid g, Material<double> m,
                                                     StepRec* rec) {
  const unsigned long long target = rec->maxws_bits;
  unsigned long long best = ~0ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.nx * g.ny;
       i += (int64_t)gridDim.x * blockDim.x) {
    double u[2 + ND];
    const double* p = g.U + i * g.es;
#pragma unroll
    for (int k = 0; k < 2 + ND; ++k) u[k] = p[k * g.cs];
    const Prim<double, ND> w = prim_of<double, ND>(u, m);
    double ws = max_wavespeed<double, ND>(u, w, m);
    ws = ws == 0.0 ? 0.0 : ws;
    const unsigned long long b = (unsigned long long)__double_as_longlong(ws);
    if (b == target && (unsigned long long)i < best) best = (unsigned long long)i;
  }
  if (best != ~0ull) atomicMin(&rec->argmax, best);
}

// Sums the partials of the step's sweeps (x then y), turns the dt/dh adjoint
// into the max-wavespeed adjoint (d(dt/dh)/d maxws = -(dt/dh)/maxws), adds
// that cell's wavespeed adjoint into lambda_n and accumulates the gamma
// adjoint.  One workgroup of 256 threads.
template <int ND>
__global__ void __launch_bounds__(256) finalize_kernel(AGrid g, Material<double> m,
                                                       const StepRec* __restrict__ rec,
                                                       const double* __restrict__ partials,
                                                       int nsweeps, double* __restrict__ lam_n,
                                                       double* __restrict__ acc) {
  __shared__ double s0[256], s1[256];
  double a0 = 0.0, a1 = 0.0;
  for (int sw = 0; sw < nsweeps; ++sw)
    for (int b = threadIdx.x; b < kPartialBlocks; b += 256) {
      a0 += partials[sw * 2 * kPartialBlocks + 2 * b];
      a1 += partials[sw * 2 * kPartialBlocks + 2 * b + 1];
    }
  s0[threadIdx.x] = a0;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double g_dtdh = s0[0];
  const double maxws = __longlong_as_double((long long)rec->maxws_bits);
  const double g_maxws = -g_dtdh * rec->dtdh / maxws;
  GMat gm{0.0};
  const unsigned long long c = rec->argmax;
  if (c < (unsigned long long)(g.nx * g.ny)) {
    double u[2 + ND], gu[2 + ND];
    const double* p = g.U + (int64_t)c * g.es;
#pragma unroll
    for (int k = 0; k < 2 + ND; ++k) {
      u[k] = p[k * g.cs];
      gu[k] = 0.0;
    }
    wavespeed_adj<ND>(u, m, g_maxws, gu, gm);
    double* l = lam_n + (int64_t)c * g.es;
#pragma unroll
    for (int k = 0; k < 2 + ND; ++k) l[k * g.cs] += gu[k];
  }
  acc[0] += s1[0] + gm.gamma;
  acc[1] = g_dtdh;
}

__global__ void acc_clear_kernel(double* acc) {
  acc[0] = 0.0;
  acc[1] = 0.0;
}

#define FL_FACE_INST(ND, D, FLUX, PRIM)                                                     \
  template __global__ void face_adj_kernel<ND, D, FLUX, PRIM>(                               \
      AGrid, const double*, Material<double>, int, const StepRec*, const double*, double*,     \
      double*, double*);
#define FL_ADJ_INST(ND, D)                                                                  \
  FL_FACE_INST(ND, D, kFluxHllc, false)                                                      \
  FL_FACE_INST(ND, D, kFluxForce, false)                                                     \
  FL_FACE_INST(ND, D, kFluxHllc, true)                                                       \
  FL_FACE_INST(ND, D, kFluxForce, true)                                                      \
  template __global__ void gather_adj_kernel<ND, D>(AGrid, const double*, double*,           \
                                                    const double*);                          \
  template __global__ void chi_adj_kernel<ND, D>(AGrid, const double*, const double*,         \
                                                 const double*, const double*, const double*, \
                                                 int*, double*, int*);

FL_ADJ_INST(1, 0)
FL_ADJ_INST(2, 0)
FL_ADJ_INST(2, 1)
template __global__ void prim_pad_kernel<1>(AGrid, Material<double>, double*, double*, double*);
template __global__ void prim_pad_kernel<2>(AGrid, Material<double>, double*, double*, double*);
template __global__ void chi_pad_kernel<1>(AGrid, const double*, const double*, const double*,
                                           double*, int*);
template __global__ void chi_pad_kernel<2>(AGrid, const double*, const double*, const double*,
                                           double*, int*);
template __global__ void chi_gather_kernel<1>(AGrid, Material<double>, const int*, const double*,
                                              const int*, double*, double*);
template __global__ void chi_gather_kernel<2>(AGrid, Material<double>, const int*, const double*,
                                              const int*, double*, double*);
template __global__ void argmax_kernel<1>(AGrid, Material<double>, StepRec*);
template __global__ void argmax_kernel<2>(AGrid, Material<double>, StepRec*);
template __global__ void finalize_kernel<1>(AGrid, Material<double>, const StepRec*,
                                            const double*, int, double*, double*);
template __global__ void finalize_kernel<2>(AGrid, Material<double>, const StepRec*,
                                            const double*, int, double*, double*);

}  // namespace adj
}  // namespace fl
