// fl_levelset.hip -- gfx950 kernels of the ghost-fluid / level-set step (FP64).
//
// Replaces the reference's level-set and ghost-fluid kernels: the interface
// predicates (levelset/levelset.hpp:166-217), set_velocities
// (levelset/cuda/velocity_initialization.cuh:108-143), the WENO-HJ5 + Godunov
// + RK3 evolve kernel (scheme/cuda/evolve.cuh:51-157 with
// scheme/stencils/weno_hj_5.hpp, scheme/schemes/godunov_upwind.hpp,
// scheme/evaluators/levelset_hamiltonian.hpp, scheme/updaters/runge_kutta_3.hpp),
// the projection (levelset/cuda/levelset_projection.cuh:54-66), the fast
// iterative eikonal solver (scheme/cuda/fast_iterative_method.cuh) and the
// Riemann ghost-fluid star states and extrapolation
// (ghost_fluid/riemann_ghost_fluid.hpp, ghost_fluid/ghost_fluid_extrapolator.hpp).
//
// One thread per cell; level-set fields are small next to the conserved
// state and read through L2.  Where the reference is undefined the kernels
// follow these rules: RK3 stages are global; level-set ghosts mirror the
// interior and a zero gradient gives a zero normal; level set m moves with
// u . n_m; the eikonal solve is a Jacobi iteration to the exact fixed point
// inside a band of cells, from sources carrying |phi|, with neighbours across
// the interface entering with negated distance; ghost-fluid states are interpolated bilinearly,
// rotated through the primitive state, stored at the interfacial cell and
// extrapolated by Jacobi sweeps over 0 < phi <= 6 dh / sqrt(2).  Every
// kernel is a deterministic function of its inputs, so results do not depend
// on launch geometry or scheduling.
#include "fl_levelset.hpp"
#include "fl_numerics_fast.hpp"

namespace fl {
namespace ls {

// Empty statement that uses a loaded value: the compiler cannot move the
// load below it, so loads issued before it are waited for together instead
// of one at a time inside the branches that use them.
__device__ __forceinline__ void keep(double v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void keep(int v) { asm volatile("" ::"v"(v)); }

// Mirrored ghost index; clamped so that no input can address outside the grid.
__device__ __forceinline__ int64_t mirror(int64_t i, int64_t n) {
  if (i < 0) i = -i - 1;
  if (i >= n) i = 2 * n - 1 - i;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

__device__ __forceinline__ double phi_at(const double* __restrict__ phi, const Grid& g,
                                         int64_t x, int64_t y) {
  const int64_t xi = mirror(x, g.nx);
  const int64_t yi = g.nd > 1 ? mirror(y, g.ny) : 0;
  return phi[yi * g.nx + xi];
}

// phi at (x, y) for x, y at most one cell outside the grid: the mirror of
// -1 is 0 and of n is n - 1, so the mirror of such an index is a clamp.
__device__ __forceinline__ int64_t clamp1(int64_t i, int64_t n) {
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

template <int ND>
__device__ __forceinline__ double phi_at1(const double* __restrict__ phi, const Grid& g,
                                          int64_t x, int64_t y) {
  const int64_t yi = ND > 1 ? clamp1(y, g.ny) : 0;
  return phi[yi * g.nx + clamp1(x, g.nx)];
}

// levelset.hpp:166-217 as mask bits.
__device__ __forceinline__ uint8_t mask_value(double p, double dh) {
  const int inside = (p <= 0.0) || (p <= -0.0);
  const int outside = p > 0.0;
  const int on_boundary = (p == 0.0) || (p == -0.0);
  const int in_ifc = (inside || on_boundary) && (p >= -dh);
  const int out_ifc = outside && (p < fabs(dh));
  return (uint8_t)(inside | (outside << 1) | (on_boundary << 2) | (in_ifc << 3) |
                   (out_ifc << 4));
}

// MultidimIterator::norm (iterator/multidim_iterator.hpp:291-306); ND = g.nd.
// Unrolled so that the 2 ND neighbour loads are issued together.
template <int ND>
__device__ __forceinline__ void norm_at(const double* __restrict__ phi, const Grid& g,
                                        int64_t x, int64_t y, double dh, double* nrm) {
  const double fact = 0.5 / dh;
  double mag = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    nrm[d] = fact * (phi_at1<ND>(phi, g, x + dx, y + dy) - phi_at1<ND>(phi, g, x - dx, y - dy));
    mag += nrm[d] * nrm[d];
  }
  const double s = sqrt(mag);
#pragma unroll
  for (int d = 0; d < ND; ++d) nrm[d] = (mag == 0.0) ? 0.0 : nrm[d] / s;
}

// The state of cell i of the grid (no ghost mapping), in FP64.
template <int S>
__device__ __forceinline__ void load_state_i(const StateView& v, int64_t i, double* u) {
  if (v.f32) {
    const float* p = (const float*)v.data + i * v.es;
#pragma unroll
    for (int k = 0; k < S; ++k) u[k] = (double)p[k * v.cs];
  } else {
    const double* p = v.data + i * v.es;
#pragma unroll
    for (int k = 0; k < S; ++k) u[k] = p[k * v.cs];
  }
}

template <int S>
__device__ __forceinline__ void load_state(const StateView& v, const Grid& g, int64_t x,
                                           int64_t y, double* u) {
  const int64_t xi = mirror(x, g.nx);
  const int64_t yi = g.nd > 1 ? mirror(y, g.ny) : 0;
  load_state_i<S>(v, yi * g.nx + xi, u);
}

// Stores an FP64 state into cell i (rounded to float on an FP32 grid).
template <int S>
__device__ __forceinline__ void store_state(const StateView& v, int64_t i, const double* u) {
  if (v.f32) {
    float* p = (float*)v.data + i * v.es;
#pragma unroll
    for (int k = 0; k < S; ++k) p[k * v.cs] = (float)u[k];
  } else {
    double* p = v.data + i * v.es;
#pragma unroll
    for (int k = 0; k < S; ++k) p[k * v.cs] = u[k];
  }
}

__device__ __forceinline__ int64_t gid() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------------------
// predicates, velocity, projection
// ---------------------------------------------------------------------------

__global__ void masks_kernel(const double* __restrict__ phi, Grid g, double dh,
                             uint8_t* __restrict__ masks) {
  const int64_t i = gid();
  if (i < g.nx * g.ny) masks[i] = mask_value(phi[i], dh);
}

// set_velocities: V = u . n of the first material whose phi <= 0.
// target < 0: the normal of the present material's level set (the
// reference's field); target = m: the normal of level set m.
template <int ND>
__global__ void velocity_kernel(const double* __restrict__ phi0, const double* __restrict__ phi1,
                                StateView u0, StateView u1, int nmat, Grid g, double dh,
                                int target, double* __restrict__ V) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  const int64_t x = i % g.nx, y = i / g.nx;
  double v = 0.0;
  for (int m = 0; m < nmat; ++m) {
    const double* phi = m == 0 ? phi0 : phi1;
    const double p = phi[i];
    if ((p <= 0.0) || (p <= -0.0)) {
      double nrm[2];
      norm_at<ND>(target < 0 ? phi : (target == 0 ? phi0 : phi1), g, x, y, dh, nrm);
      double u[2 + ND];
      load_state_i<2 + ND>(m == 0 ? u0 : u1, i, u);
      double sum = 0.0;
#pragma unroll
      for (int d = 0; d < ND; ++d) sum += (u[2 + d] / u[0]) * nrm[d];
      v = sum;
      break;
    }
  }
  V[i] = v;
}

// Both level sets' speeds in one pass (the two-material step's
// velocity_kernel with target 0 and target 1): V_m = u.n_m of the material
// present, with the same expressions per target.
template <int ND>
__global__ void velocity2_kernel(const double* __restrict__ phi0, const double* __restrict__ phi1,
                                 StateView u0, StateView u1, Grid g, double dh,
                                 double* __restrict__ V0, double* __restrict__ V1) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  const int64_t x = i % g.nx, y = i / g.nx;
  // the first material whose phi <= 0 (both values loaded up front)
  // (bitwise ors: no short circuit that would move the second load after
  // the first test)
  const double p0 = phi0[i], p1 = phi1[i];
  keep(p0);
  keep(p1);
  const bool in0 = ((int)(p0 <= 0.0) | (int)(p0 <= -0.0)) != 0;
  const bool in1 = ((int)(p1 <= 0.0) | (int)(p1 <= -0.0)) != 0;
  const int m = in0 ? 0 : (in1 ? 1 : -1);
  double v0 = 0.0, v1 = 0.0;
  if (m >= 0) {
    double n0[2], n1[2];
    norm_at<ND>(phi0, g, x, y, dh, n0);
    norm_at<ND>(phi1, g, x, y, dh, n1);
    double u[2 + ND];
    load_state_i<2 + ND>(m == 0 ? u0 : u1, i, u);
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      s0 += (u[2 + d] / u[0]) * n0[d];
      s1 += (u[2 + d] / u[0]) * n1[d];
    }
    v0 = s0;
    v1 = s1;
  }
  V0[i] = v0;
  V1[i] = v1;
}

__global__ void project2_kernel(double* __restrict__ phi0, double* __restrict__ phi1,
                                int64_t count) {
  const int64_t i = gid();
  if (i >= count) return;
  const double avg = 0.5 * (phi0[i] + phi1[i]);
  const double a = phi0[i] - avg;
  phi0[i] = a;
  phi1[i] = -a;
}

// ---------------------------------------------------------------------------
// WENO-HJ5 + Godunov + RK3
// ---------------------------------------------------------------------------

__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// HJWeno5::diff (scheme/stencils/weno_hj_5.hpp:97-143), std::pow(x, 2) as x*x.
// FAST: the fast FP64 numerics (four refined reciprocals instead of the
// correctly rounded quotients).
template <bool FAST>
__device__ __forceinline__ double weno_diff(double d1, double d2, double d3, double d4,
                                            double d5) {
  const double phi_1 = 2.0 * d1 - 7.0 * d2 + 11.0 * d3;
  const double phi_2 = -d2 + 5.0 * d3 + 2.0 * d4;
  const double phi_3 = 2.0 * d3 + 5.0 * d4 - d5;
  const double f1 = 13.0 / 12.0;
  const double f2 = 1.0 / 4.0;
  const double s11 = d1 - 2.0 * d2 + d3, s12 = d1 - 4.0 * d2 + 3.0 * d3;
  const double s21 = d2 - 2.0 * d3 + d4, s22 = d2 - d4;
  const double s31 = d3 - 2.0 * d4 + d5, s32 = 3.0 * d3 - 4.0 * d4 + d5;
  double alpha_1 = f1 * (s11 * s11) + f2 * (s12 * s12);
  double alpha_2 = f1 * (s21 * s21) + f2 * (s22 * s22);
  double alpha_3 = f1 * (s31 * s31) + f2 * (s32 * s32);
  const double e =
      10.0e-6 * smax(smax(smax(smax(d1 * d1, d2 * d2), d3 * d3), d4 * d4), d5 * d5) + 10.0e-99;
  const double a1e = alpha_1 + e, a2e = alpha_2 + e, a3e = alpha_3 + e;
  if (FAST) {
    // the quotients as products with refined reciprocals; a*e*^2 >= 1e-196
    // and the weight sum <= 6e196 keep every reciprocal in the normal range
    alpha_1 = 0.1 * fast::rcp(a1e * a1e);
    alpha_2 = 0.6 * fast::rcp(a2e * a2e);
    alpha_3 = 0.3 * fast::rcp(a3e * a3e);
    const double w_denom = (alpha_1 + alpha_2 + alpha_3) * 6.0;
    return (alpha_1 * phi_1 + alpha_2 * phi_2 + alpha_3 * phi_3) * fast::rcp(w_denom);
  }
  alpha_1 = 0.1 / (a1e * a1e);
  alpha_2 = 0.6 / (a2e * a2e);
  alpha_3 = 0.3 / (a3e * a3e);
  const double w_denom = (alpha_1 + alpha_2 + alpha_3) * 6.0;
  return (alpha_1 * phi_1 + alpha_2 * phi_2 + alpha_3 * phi_3) / w_denom;
}

// Mirrored index within a line of n cells (32-bit; n < 2^30).
__device__ __forceinline__ int mirror32(int i, int n) {
  if (i < 0) i = -i - 1;
  if (i >= n) i = 2 * n - 1 - i;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// L = V * H(phi) (levelset_hamiltonian.hpp:63-98, godunov_upwind.hpp:94-166).
// h1 = 1 / dh, computed once by the caller (the same IEEE quotient).  ND =
// g.nd; the 7-cell stencils of all dimensions are loaded before any is used.
// xin / yin: every cell of the stencil in x / y lies inside the grid (the
// caller's workgroup-uniform test), so the mirror index math is skipped.
template <int ND, bool FAST>
__device__ double ls_rhs(const double* __restrict__ phi, const Grid& g, int x, int y, double v,
                         double h1, bool xin, bool yin) {
  double result = 0.0;
  const int nx = (int)g.nx, ny = (int)g.ny;
  double fs[ND][7];
  {
    const double* row = phi + (int64_t)y * nx;
    if (xin) {
#pragma unroll
      for (int k = -3; k <= 3; ++k) fs[0][k + 3] = row[x + k];
    } else {
#pragma unroll
      for (int k = -3; k <= 3; ++k) fs[0][k + 3] = row[mirror32(x + k, nx)];
    }
  }
  if (ND > 1) {
    if (yin) {
#pragma unroll
      for (int k = -3; k <= 3; ++k) fs[ND - 1][k + 3] = phi[(int64_t)(y + k) * nx + x];
    } else {
#pragma unroll
      for (int k = -3; k <= 3; ++k) fs[ND - 1][k + 3] = phi[(int64_t)mirror32(y + k, ny) * nx + x];
    }
  }
  // A stencil whose cells all hold the centre's bit pattern (the clamped
  // values outside the reinitialised band) has every difference +0: each
  // one-sided WENO derivative is then 0 / (6 * 10^195) = +0 and the Godunov
  // sum is +0, so L = v * sqrt(+0) = v * 0 exactly.  A wave whose cells are
  // all such skips the WENO arithmetic; the others take it.
  bool flat = isfinite(fs[0][3]);
#pragma unroll
  for (int d = 0; d < ND; ++d)
#pragma unroll
    for (int k = 0; k < 7; ++k)
      flat = flat && __double_as_longlong(fs[d][k]) == __double_as_longlong(fs[0][3]);
  if (__all(flat)) return v * 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double* f = fs[d];
    const double bd = weno_diff<FAST>(h1 * (f[1] - f[0]), h1 * (f[2] - f[1]), h1 * (f[3] - f[2]),
                                h1 * (f[4] - f[3]), h1 * (f[5] - f[4]));
    const double fd = weno_diff<FAST>(h1 * (f[6] - f[5]), h1 * (f[5] - f[4]), h1 * (f[4] - f[3]),
                                h1 * (f[3] - f[2]), h1 * (f[2] - f[1]));
    double back, fwrd;
    if (v <= 0.0) {
      back = smax(bd, 0.0);
      fwrd = smin(fd, 0.0);
    } else {
      back = smin(bd, 0.0);
      fwrd = smax(fd, 0.0);
    }
    result += smax(back * back, fwrd * fwrd);
  }
  return v * (FAST ? fast::sqt(result) : sqrt(result));
}

// One RK3 stage (scheme/updaters/runge_kutta_3.hpp:117-148), ND = g.nd:
//  stage 0: out = phi_n - dt L(phi_n)
//  stage 1: out = 0.75 phi_n + 0.25 (s1 - dt L(s1))
//  stage 2: out = (1/3) phi_n + (2 (1/3)) (s2 - dt L(s2))
// Launched on a (ceil(nx / blockDim.x), min(ny, 65535)) grid; rows stride by
// gridDim.y.
template <int ND, bool FAST>
__device__ __forceinline__ void rk3_stage_body(int stage, const double* __restrict__ phi_n,
                                               const double* __restrict__ src,
                                               const double* __restrict__ V, const Grid& g,
                                               double dt, double h1, double* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t x0 = (int64_t)blockIdx.x * blockDim.x;
  const bool xin = x0 >= 3 && x0 + blockDim.x + 3 <= g.nx;  // workgroup-uniform
  if (x >= g.nx) return;
  for (int y = blockIdx.y; y < g.ny; y += gridDim.y) {
    const int64_t i = (int64_t)y * g.nx + x;
    const bool yin = y >= 3 && y + 3 < g.ny;
    const double l = ls_rhs<ND, FAST>(src, g, x, y, V[i], h1, xin, yin);
    if (stage == 0) {
      out[i] = phi_n[i] - dt * l;
    } else if (stage == 1) {
      out[i] = 0.75 * phi_n[i] + 0.25 * (src[i] - dt * l);
    } else {
      const double fact_13 = 1.0 / 3.0;
      const double fact_23 = 2.0 * fact_13;
      const double s3 = src[i] - dt * l;
      out[i] = (fact_13 * phi_n[i]) + (fact_23 * s3);
    }
  }
}

// dt_dev (nullable): dt in device memory (the step's control block), read
// instead of dt.
template <int ND>
__global__ void rk3_stage_kernel(int stage, const double* __restrict__ phi_n,
                                 const double* __restrict__ src, const double* __restrict__ V,
                                 Grid g, double dt, const double* dt_dev, double h1,
                                 double* __restrict__ out) {
  if (dt_dev) dt = *dt_dev;
  rk3_stage_body<ND, false>(stage, phi_n, src, V, g, dt, h1, out);
}

// The same stage of two level sets' advection in one launch (blockIdx.z
// selects the job), dt in device memory; FAST: the fast FP64 numerics.
template <int ND, bool FAST>
__global__ void rk3_stage2_kernel(int stage, Rk3Job j0, Rk3Job j1, Grid g,
                                  const double* dt_dev, double h1) {
  const Rk3Job j = blockIdx.z ? j1 : j0;
  rk3_stage_body<ND, FAST>(stage, j.phi_n, j.src, j.V, g, *dt_dev, h1, j.out);
}

// ---------------------------------------------------------------------------
// eikonal reinitialisation
// ---------------------------------------------------------------------------

__device__ __forceinline__ int signum(double x) { return (0.0 < x) - (x < 0.0); }

// The job of this workgroup (selected by value so that the kernel argument
// block is never indexed dynamically).
__device__ __forceinline__ EikJob eik_job(const EikJobs& js) {
  return blockIdx.z ? js.j[1] : js.j[0];
}

// Sources (fast_iterative_init, scheme/cuda/fast_iterative_method.cuh:48-85):
// a cell beside a sign change with the smaller |phi|; distance |phi| there.
// info = source | (signum(phi) + 1) << 1.  Both Jacobi buffers start equal.
__global__ void eik_init_kernel(EikJobs jobs, Grid g) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  const EikJob J = eik_job(jobs);
  const double* __restrict__ phi = J.phi;
  double* __restrict__ d = J.d[0];
  double* __restrict__ d1 = J.d[1];
  uint8_t* __restrict__ info = J.info;
  const int64_t x = i % g.nx, y = i / g.nx;
  const double c = phi[i];
  // the four axis neighbours (those past the edge, or in y on a 1D grid,
  // read the cell itself and do not count), loaded together
  const bool ok[4] = {x + 1 < g.nx, x > 0, g.nd > 1 && y + 1 < g.ny, g.nd > 1 && y > 0};
  const int64_t nb[4] = {i + 1, i - 1, i + g.nx, i - g.nx};
  double o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = phi[ok[k] ? nb[k] : i];
  int source = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (ok[k] && signum(c) != signum(o[k]) && fabs(c) < fabs(o[k])) source = 1;
  info[i] = (uint8_t)(source | ((signum(c) + 1) << 1));
  const double v = source ? fabs(c) : DBL_MAX;
  d[i] = v;
  d1[i] = v;
}

// Tiles of the eikonal sweeps: kEikTileW cells x kEikTileH rows, one per
// workgroup of kEikTileW * kEikTileH threads.  src[t] = the tile holds a
// source.  Launched on (tiles per row, min(tile rows, 65535)) workgroups; tile
// rows stride by gridDim.y.
__device__ __forceinline__ void eik_cell_of(int64_t tr, int64_t& x, int64_t& y) {
  x = (int64_t)blockIdx.x * kEikTileW + (threadIdx.x % kEikTileW);
  y = tr * kEikTileH + (threadIdx.x / kEikTileW);
}

__global__ void __launch_bounds__(kEikTileW * kEikTileH) eik_tile_src_kernel(
    EikJobs jobs, Grid g, int64_t trows) {
  const EikJob J = eik_job(jobs);
  const uint8_t* __restrict__ info = J.info;
  uint8_t* __restrict__ src = J.src;
  const int64_t segs = gridDim.x;
  for (int64_t tr = blockIdx.y; tr < trows; tr += gridDim.y) {
    int64_t x, y;
    eik_cell_of(tr, x, y);
    const int has = (x < g.nx && y < g.ny) ? (info[y * g.nx + x] & 1) : 0;
    const int any = __syncthreads_or(has);
    if (threadIdx.x == 0) src[tr * segs + blockIdx.x] = (uint8_t)(any != 0);
  }
}

// Tile t is active when some tile within `tr_r` tile rows and `segs_r`
// tiles of it holds a source; the active tiles are appended to the job's
// list (one atomicAdd per wave; ctl->ntiles starts at zero).  A cell's
// distance grows by at least dh/sqrt(2) per axis step from the nearest
// source (the Godunov update q >= min(a, b) + dh/sqrt(2)), so a cell farther
// than band * sqrt(2) steps from every source never enters the band and
// tiles without such cells can be skipped by every sweep.
__global__ void eik_tile_act_kernel(EikJobs jobs, int64_t trows, int64_t segs, int tr_r,
                                    int segs_r) {
  const int64_t t = gid();
  if (t >= trows * segs) return;
  const EikJob J = eik_job(jobs);
  const uint8_t* __restrict__ src = J.src;
  const int64_t y = t / segs, sx = t % segs;
  const int64_t y0 = y - tr_r < 0 ? 0 : y - tr_r, y1 = y + tr_r >= trows ? trows - 1 : y + tr_r;
  const int64_t s0 = sx - segs_r < 0 ? 0 : sx - segs_r;
  const int64_t s1 = sx + segs_r >= segs ? segs - 1 : sx + segs_r;
  uint8_t a = 0;
  for (int64_t yy = y0; yy <= y1 && !a; ++yy)
    for (int64_t ss = s0; ss <= s1; ++ss) a |= src[yy * segs + ss];
  const unsigned long long m = __ballot(a != 0);
  if (!m) return;
  const int lane = (int)(threadIdx.x % 64);
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(&J.ctl->ntiles, __popcll(m));
  base = __shfl(base, leader, 64);
  if (a) J.list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int)t;
}

// Neighbour distance v (info byte inf) seen from a cell with sign bits sc:
// negated across the interface unless not yet reached.
__device__ __forceinline__ double eik_nbr_v(double v, uint8_t inf, uint8_t sc) {
  return ((inf & 6) == sc || v == DBL_MAX) ? v : -v;
}

// One thread per job.
__global__ void eik_ctl_init_kernel(EikJobs jobs, int njobs) {
  if ((int)threadIdx.x >= njobs) return;
  EikCtl* c = threadIdx.x ? jobs.j[1].ctl : jobs.j[0].ctl;
  c->flags[0] = c->flags[1] = c->flags[2] = 0;
  c->done = 0;
  c->iters = 0;
  c->ntiles = 0;
  c->pad = 0;
}

// Sweep k of the Godunov eikonal update
// (scheme/schemes/godunov_upwind_scheme.hpp:37-53, f = 1) restricted to the
// band: a cell takes q when q < d and q <= limit.  A sweep after one that
// changed nothing marks the loop done; later sweeps return at once, so a
// batch of sweeps can be launched ahead of the convergence test.  Every
// active sweep writes every cell of the active tiles (the others hold their
// initial value in both buffers), so once converged both buffers agree.
// Sweep k reads the job's buffer d[k & 1] and writes d[(k + 1) & 1].  Each
// job has its own control words, so one job's sweeps return at once when it
// has converged while the other's go on.  Launched on (kEikSweepGroups, 1,
// jobs) workgroups of kEikTileW * kEikTileH threads that stride over the
// job's active tiles, one tile per workgroup at a time.
__global__ void __launch_bounds__(kEikTileW * kEikTileH) eik_sweep_kernel(
    EikJobs jobs, Grid g, int64_t segs, double dh, double limit, int k) {
  const EikJob J = eik_job(jobs);
  const double* __restrict__ d = (k & 1) ? J.d[1] : J.d[0];
  double* __restrict__ e = (k & 1) ? J.d[0] : J.d[1];
  const uint8_t* __restrict__ info = J.info;
  const int* __restrict__ list = J.list;
  EikCtl* __restrict__ ctl = J.ctl;
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  // the control words, loaded together ((k + 2) % 3 is sweep k-1's flag)
  const int done = ctl->done, prev_changed = ctl->flags[(k + 2) % 3];
  const int ntiles = ctl->ntiles;
  keep(done);
  keep(prev_changed);
  keep(ntiles);
  if (done) return;
  if (k > 0 && prev_changed == 0) {
    if (first) ctl->done = 1;
    return;
  }
  if (first) {
    ctl->flags[(k + 1) % 3] = 0;
    ctl->iters = k + 1;
  }
  bool changed = false;
  for (int li = blockIdx.x; li < ntiles; li += gridDim.x) {
    const int64_t t = list[li];
    const int64_t x = (t % segs) * kEikTileW + (threadIdx.x % kEikTileW);
    const int64_t y = (t / segs) * kEikTileH + (threadIdx.x / kEikTileW);
    if (x >= g.nx || y >= g.ny) continue;
    const int64_t i = y * g.nx + x;
    const double cur = d[i];
    double nv = cur;
    const uint8_t me = info[i];
    // neighbours loaded together, before the branch (those past the edge
    // read the cell itself and are replaced by DBL_MAX)
    const bool okxm = x > 0, okxp = x < g.nx - 1;
    const bool okym = g.nd > 1 && y > 0, okyp = g.nd > 1 && y < g.ny - 1;
    const int64_t jxm = okxm ? i - 1 : i, jxp = okxp ? i + 1 : i;
    const int64_t jym = okym ? i - g.nx : i, jyp = okyp ? i + g.nx : i;
    const double dxm = d[jxm], dxp = d[jxp], dym = d[jym], dyp = d[jyp];
    const uint8_t ixm = info[jxm], ixp = info[jxp], iym = info[jym], iyp = info[jyp];
    keep(dxm);
    keep(dxp);
    keep(dym);
    keep(dyp);
    keep((int)(ixm | (ixp << 8) | (iym << 16) | (iyp << 24)));
    if (!(me & 1)) {
      const uint8_t sc = me & 6;
      const double fh = 1.0 * dh;
      const double xm = okxm ? eik_nbr_v(dxm, ixm, sc) : DBL_MAX;
      const double xp = okxp ? eik_nbr_v(dxp, ixp, sc) : DBL_MAX;
      const double a = smin(xm, xp);
      double q;
      if (g.nd == 1) {
        q = a + fh;
      } else {
        const double ym = okym ? eik_nbr_v(dym, iym, sc) : DBL_MAX;
        const double yp = okyp ? eik_nbr_v(dyp, iyp, sc) : DBL_MAX;
        const double b = smin(ym, yp);
        const double amb = a - b;
        q = fabs(amb) >= fh ? smin(a, b) + fh
                            : 0.5 * (a + b + sqrt(2.0 * fh * fh - (amb * amb)));
      }
      if (q < cur && q <= limit) {
        nv = q;
        changed = true;
      }
    }
    e[i] = nv;
  }
  // one store of the change flag per wave
  const unsigned long long m = __ballot(changed);
  if (m && (int)(threadIdx.x % 64) == __ffsll((long long)m) - 1) ctl->flags[k % 3] = 1;
}

// d = the converged buffer d[iters & 1] of the job; both hold the same
// values once the loop has converged.
__global__ void eik_finish_kernel(EikJobs jobs, Grid g, double limit, double clamp) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  const EikJob J = eik_job(jobs);
  const double* __restrict__ phi = J.phi;
  const double* __restrict__ d = (J.ctl->iters & 1) ? J.d[1] : J.d[0];
  double* __restrict__ out = J.out;
  uint8_t* __restrict__ band_mask = J.band_mask;
  const int in_band = d[i] <= limit;
  const double dist = in_band ? d[i] : clamp;
  out[i] = (double)signum(phi[i]) * dist;
  if (band_mask) band_mask[i] = (uint8_t)in_band;
}

// ---------------------------------------------------------------------------
// Riemann ghost fluid
// ---------------------------------------------------------------------------

// Conservative <-> primitive with the reference's expressions
// (state/state_impl.hpp:117-142, 340-406); stiff: pressure ignores p_inf,
// energy uses it (material/stiff_fluid.hpp:77-98).
template <int ND>
__device__ __forceinline__ void cons_to_prim(const double* u, const Eos& e, double* w) {
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const double v = u[2 + i] / u[0];
    sum += v * v;
  }
  w[0] = u[0];
  w[1] = (e.gamma - 1.0) * (u[1] - 0.5 * u[0] * sum);
  if (e.corr) w[1] = w[1] - e.gamma * e.p_inf;
#pragma unroll
  for (int i = 0; i < ND; ++i) w[2 + i] = u[2 + i] / u[0];
}

template <int ND>
__device__ __forceinline__ void prim_to_cons(const double* w, const Eos& e, double* u) {
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < ND; ++i) sum += w[2 + i] * w[2 + i];
  const double eos = e.stiff ? (w[1] + e.gamma * e.p_inf) / ((e.gamma - 1.0) * w[0])
                             : w[1] / ((e.gamma - 1.0) * w[0]);
  u[0] = w[0];
  u[1] = w[0] * (0.5 * sum + eos);
#pragma unroll
  for (int i = 0; i < ND; ++i) u[2 + i] = w[2 + i] * w[0];
}

template <int ND>
__device__ __forceinline__ double pressure(const double* u, const Eos& e) {
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const double v = u[2 + i] / u[0];
    sum += v * v;
  }
  const double p = (e.gamma - 1.0) * (u[1] - 0.5 * u[0] * sum);
  return e.corr ? p - e.gamma * e.p_inf : p;
}

template <int ND>
__device__ __forceinline__ double sound(const double* u, const Eos& e) {
  const double p = pressure<ND>(u, e);
  return e.stiff ? sqrt(e.gamma * (p + e.p_inf) / u[0]) : sqrt(e.gamma * p / u[0]);
}

// Hllc::solve_star_conservative_left (flux_method/flux_hllc.hpp:301-335) with
// wavespeed_slr, wavespeed_star and star_state_cons (:420-571), x direction.
template <int ND>
__device__ void hllc_star_left(const double* ul, const double* ur, const Eos& el,
                               const Eos& er, double* out) {
  const double al = sound<ND>(ul, el), ar = sound<ND>(ur, er);
  const double pl = pressure<ND>(ul, el), pr = pressure<ND>(ur, er);
  const double vl = ul[2] / ul[0], vr = ur[2] / ur[0];
  const double p_pvrs = 0.5 * ((pl + pr) - (0.25 * (vr - vl) * (ur[0] + ul[0]) * (al + ar)));
  // a corrected stiffened gas on both sides bounds p* by -min(p_inf)
  const double p_star = (el.corr && er.corr) ? smax(-smin(el.p_inf, er.p_inf), p_pvrs)
                                             : smax(0.0, p_pvrs);
  const double ql = (el.gamma + 1.0) / (2.0 * el.gamma);
  const double qr = (er.gamma + 1.0) / (2.0 * er.gamma);
  // the corrected stiffened gas takes pressure ratios of p + p_inf
  const double rl = el.corr ? (p_star + el.p_inf) / (pl + el.p_inf) : p_star / pl;
  const double rr = er.corr ? (p_star + er.p_inf) / (pr + er.p_inf) : p_star / pr;
  const double wsl = (p_star <= pl) ? vl + -al : vl + -al * sqrt(1.0 + ql * (rl - 1.0));
  const double wsr = (p_star <= pr) ? vr + ar : vr + ar * sqrt(1.0 + qr * (rr - 1.0));
  const double factorl = ul[0] * (wsl - vl);
  const double factorr = ur[0] * (wsr - vr);
  const double ws_star = (pr - pl + vl * factorl - vr * factorr) / (factorl - factorr);
  const double state_factor = wsl - vl;
  const double scale_factor = state_factor / (wsl - ws_star);
#pragma unroll
  for (int k = 0; k < 2 + ND; ++k) out[k] = ul[k] * scale_factor;
  out[2] = ws_star * ul[0] * scale_factor;
  out[1] = scale_factor * (ul[1] + ul[0] * (ws_star - vl) * (ws_star + pl / (ul[0] * state_factor)));
}

// math::interp (math/interp.hpp:56-165) as a proper (bi)linear interpolation.
template <int ND>
__device__ void interp_state(const StateView& v, const Grid& g, int64_t x, int64_t y,
                             const double* w, double* out) {
  constexpr int S = 2 + ND;
  int64_t o0[2] = {0, 0}, o1[2] = {0, 0};
  double fr[2] = {0.0, 0.0};
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const int sign = signum(w[d]);
    const double abs_w = fabs(w[d]);
    const double abs_off = floor(abs_w);
    o0[d] = (int64_t)(sign * abs_off);
    o1[d] = o0[d] + sign;
    fr[d] = abs_w - abs_off;
  }
  double a[S], b[S];
  if constexpr (ND == 1) {
    load_state<S>(v, g, x + o0[0], y, a);
    load_state<S>(v, g, x + o1[0], y, b);
#pragma unroll
    for (int k = 0; k < S; ++k) out[k] = a[k] * (1.0 - fr[0]) + b[k] * fr[0];
  } else {
    double c[S], e[S];
    load_state<S>(v, g, x + o0[0], y + o0[1], a);
    load_state<S>(v, g, x + o1[0], y + o0[1], b);
    load_state<S>(v, g, x + o0[0], y + o1[1], c);
    load_state<S>(v, g, x + o1[0], y + o1[1], e);
    const double factor = 1.0 - fr[0];
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const double close = a[k] * factor + b[k] * fr[0];
      const double far = c[k] * factor + e[k] * fr[0];
      out[k] = close * (1.0 - fr[1]) + far * fr[1];
    }
  }
}

// RiemannGFM<3>::invoke (ghost_fluid/riemann_ghost_fluid.hpp:50-195) at an
// inside-interfacial cell i = (x, y) of material m: p = phi_m[i], po =
// phi_o[i].  Writes star[i] and sets has_star[i] when the cell takes a star
// state.
template <int ND>
__device__ __forceinline__ void star_cell(const double* __restrict__ phi_m, double po,
                                          const StateView& um, const StateView& uo,
                                          const Eos& em, const Eos& eo, const Grid& g, double dh,
                                          int64_t x, int64_t y, int64_t i, double p,
                                          double* __restrict__ star,
                                          uint8_t* __restrict__ has_star) {
  constexpr int S = 2 + ND;
  for (int d = 0; d < ND; ++d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    if (phi_at1<ND>(phi_m, g, x - dx, y - dy) == 0.0 || phi_at1<ND>(phi_m, g, x + dx, y + dy) == 0.0)
      return;
  }
  const uint8_t mo = mask_value(po, dh);
  if (!((mo & 16) || (mo & 4))) return;
  double nrm[2] = {0.0, 0.0};
  norm_at<ND>(phi_m, g, x, y, dh, nrm);
  const double abs_phi = fabs(p);
  const double dh_three_two = 1.5 * dh;
  double w[2] = {0.0, 0.0}, wl[S], wr[S];
#pragma unroll
  for (int d = 0; d < ND; ++d) w[d] = nrm[d] * (abs_phi - dh_three_two) / dh;
  interp_state<ND>(um, g, x, y, w, wl);
#pragma unroll
  for (int d = 0; d < ND; ++d) w[d] = nrm[d] * (abs_phi + dh_three_two) / dh;
  interp_state<ND>(uo, g, x, y, w, wr);
  double Wl[S], Wr[S], vl[2] = {0.0, 0.0};
  cons_to_prim<ND>(wl, em, Wl);
  cons_to_prim<ND>(wr, eo, Wr);
  double vl_n = 0.0, vr_n = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    vl[d] = Wl[2 + d];
    vl_n += Wl[2 + d] * nrm[d];
    vr_n += Wr[2 + d] * nrm[d];
  }
  Wl[2] = vl_n;
  Wr[2] = vr_n;
#pragma unroll
  for (int d = 1; d < ND; ++d) {
    Wl[2 + d] = 0.0;
    Wr[2 + d] = 0.0;
  }
  prim_to_cons<ND>(Wl, em, wl);
  prim_to_cons<ND>(Wr, eo, wr);
  double st[S], Ws[S];
  hllc_star_left<ND>(wl, wr, em, eo, st);
  cons_to_prim<ND>(st, em, Ws);
  const double vs_n = Ws[2];
#pragma unroll
  for (int d = 0; d < ND; ++d) Ws[2 + d] = vl[d] + (vs_n - vl_n) * nrm[d];
  double out[S];
  prim_to_cons<ND>(Ws, em, out);
#pragma unroll
  for (int k = 0; k < S; ++k) star[i * S + k] = out[k];
  has_star[i] = 1;
}

// Star states of one material (one thread per cell); count_reset (nullable)
// is zeroed for the extrapolation that follows.
template <int ND>
__global__ void gfm_star_kernel(const double* __restrict__ phi_m, const double* __restrict__ phi_o,
                                StateView um, StateView uo, Eos em, Eos eo, Grid g, double dh,
                                double* __restrict__ star, uint8_t* __restrict__ has_star,
                                int* __restrict__ count_reset) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  if (count_reset && i == 0) *count_reset = 0;
  const int64_t x = i % g.nx, y = i / g.nx;
  has_star[i] = 0;
  const double p = phi_m[i];
  if (!(mask_value(p, dh) & 8)) return;
  star_cell<ND>(phi_m, phi_o[i], um, uo, em, eo, g, dh, x, y, i, p, star, has_star);
}

// Star states of both materials of a two-material step in one pass: thread i
// tests cell i for both, from one load of each level set; both band-list
// counts are zeroed.
template <int ND>
__global__ void gfm_star2_kernel(const double* __restrict__ phi0, const double* __restrict__ phi1,
                                 StateView u0, StateView u1, Eos e0, Eos e1, Grid g, double dh,
                                 double* __restrict__ star0, double* __restrict__ star1,
                                 uint8_t* __restrict__ has0, uint8_t* __restrict__ has1,
                                 int* __restrict__ count0, int* __restrict__ count1) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  if (i == 0) {
    *count0 = 0;
    *count1 = 0;
  }
  const int64_t x = i % g.nx, y = i / g.nx;
  const double p0 = phi0[i], p1 = phi1[i];
  has0[i] = 0;
  has1[i] = 0;
  if (mask_value(p0, dh) & 8)
    star_cell<ND>(phi0, p1, u0, u1, e0, e1, g, dh, x, y, i, p0, star0, has0);
  if (mask_value(p1, dh) & 8)
    star_cell<ND>(phi1, p0, u1, u0, e1, e0, g, dh, x, y, i, p1, star1, has1);
}

// Star states of material m among ms.nmat materials (one thread per cell):
// at an inside-interfacial cell of m, every other material o in order whose
// level set is outside-interfacial or on the boundary there gives a star
// state, a later o overwriting an earlier one, as the reference's loop over
// the other materials does (ghost_fluid/riemann_ghost_fluid.hpp:56-190).
template <int ND>
__global__ void gfm_star_n_kernel(MatSet ms, int m, Grid g, double dh,
                                  double* __restrict__ star, uint8_t* __restrict__ has_star,
                                  int* __restrict__ count_reset) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  if (count_reset && i == 0) *count_reset = 0;
  const int64_t x = i % g.nx, y = i / g.nx;
  has_star[i] = 0;
  const double* phi_m = ms.phi[m];
  const double p = phi_m[i];
  if (!(mask_value(p, dh) & 8)) return;
  for (int o = 0; o < ms.nmat; ++o) {
    if (o == m) continue;
    star_cell<ND>(phi_m, ms.phi[o][i], ms.u[m], ms.u[o], ms.e[m], ms.e[o], g, dh, x, y, i, p,
                  star, has_star);
  }
}

// set_velocities for level set `target` among ms.nmat materials: V = u . n
// with u the velocity of the first material whose phi <= 0 and n the normal
// of level set `target` (velocity_kernel with n materials).
template <int ND>
__global__ void velocity_n_kernel(MatSet ms, int target, Grid g, double dh,
                                  double* __restrict__ V) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  const int64_t x = i % g.nx, y = i / g.nx;
  double v = 0.0;
  for (int m = 0; m < ms.nmat; ++m) {
    const double p = ms.phi[m][i];
    if ((p <= 0.0) || (p <= -0.0)) {
      double nrm[2];
      norm_at<ND>(ms.phi[target], g, x, y, dh, nrm);
      double u[2 + ND];
      load_state_i<2 + ND>(ms.u[m], i, u);
      double sum = 0.0;
#pragma unroll
      for (int d = 0; d < ND; ++d) sum += (u[2 + d] / u[0]) * nrm[d];
      v = sum;
      break;
    }
  }
  V[i] = v;
}

// project_impl for n > 2 level sets (levelset/cuda/levelset_projection.cuh:
// 87-118): the average of the two smallest values at a cell, from the
// running pair (max - 1, max), is subtracted from every level set.
__global__ void project_n_kernel(PhiSet ps, int64_t count) {
  const int64_t i = gid();
  if (i >= count) return;
  double smallest = DBL_MAX - 1.0, next_smallest = DBL_MAX;
  for (int m = 0; m < ps.nmat; ++m) {
    const double v = ps.phi[m][i];
    if (v < smallest) {
      next_smallest = smallest;
      smallest = v;
    } else if (v < next_smallest) {
      next_smallest = v;
    }
  }
  const double avg = 0.5 * (smallest + next_smallest);
  for (int m = 0; m < ps.nmat; ++m) ps.phi[m][i] -= avg;
}

// 0 < p <= upper, as an int so that several can be or-ed without a short
// circuit.
__device__ __forceinline__ int in_band(double p, double upper) {
  return (int)(p > 0.0) & (int)(p <= upper);
}

// Extrapolation source field, in both Jacobi buffers, and the band list, in
// one pass: the material's own state (AoS scratch), with the star state at
// interfacial cells.  Only the cells an extrapolation sweep can read are
// written: the band cells and their neighbours (a band cell reads itself and
// its upwind neighbours).  Band cells 0 < phi <= upper are appended to `list`
// in no particular order: each band cell's update is independent of the
// others within a sweep.
template <int ND>
__device__ __forceinline__ void band_flags(const double* __restrict__ phi_m, double upper,
                                           const Grid& g, int64_t x, int64_t y, int64_t i,
                                           bool& band, bool& needed) {
  // neighbours past the edge read the cell itself, which adds nothing to
  // `needed`; the loads are unconditional so that they are issued together
  const double pc = phi_m[i];
  const double pxm = phi_m[x > 0 ? i - 1 : i];
  const double pxp = phi_m[x < g.nx - 1 ? i + 1 : i];
  const int b = in_band(pc, upper);
  int n = b | in_band(pxm, upper) | in_band(pxp, upper);
  if (ND > 1) {
    const double pym = phi_m[y > 0 ? i - g.nx : i];
    const double pyp = phi_m[y < g.ny - 1 ? i + g.nx : i];
    n |= in_band(pym, upper) | in_band(pyp, upper);
  }
  band = b != 0;
  needed = n != 0;
}

// Appends cell i to `list` where `take` holds: one atomicAdd per wave.  Every
// lane of the wave that is still running must call it.
__device__ __forceinline__ void list_append(bool take, int64_t i, int* __restrict__ list,
                                            int* __restrict__ count) {
  const unsigned long long m = __ballot(take);
  if (!m) return;
  const int lane = (int)(threadIdx.x % 64);
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader, 64);
  if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int)i;
}

template <int ND>
__device__ __forceinline__ void band_seed_cell(const StateView& um,
                                               const double* __restrict__ star,
                                               const uint8_t* __restrict__ has_star,
                                               const Grid& g, int64_t x, int64_t y, int64_t i,
                                               double* __restrict__ A, double* __restrict__ B) {
  constexpr int S = 2 + ND;
  double u[S];
  if (has_star[i]) {
#pragma unroll
    for (int k = 0; k < S; ++k) u[k] = star[i * S + k];
  } else {
    load_state<S>(um, g, x, y, u);
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    A[i * S + k] = u[k];
    B[i * S + k] = u[k];
  }
}

template <int ND>
__global__ void gfm_band_seed_kernel(const double* __restrict__ phi_m, double upper,
                                     StateView um, const double* __restrict__ star,
                                     const uint8_t* __restrict__ has_star, Grid g,
                                     double* __restrict__ A, double* __restrict__ B,
                                     int* __restrict__ list, int* __restrict__ count) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  const int64_t x = i % g.nx, y = i / g.nx;
  bool band, needed;
  band_flags<ND>(phi_m, upper, g, x, y, i, band, needed);
  list_append(band, i, list, count);
  if (needed) band_seed_cell<ND>(um, star, has_star, g, x, y, i, A, B);
}

// Both materials' band seeds in one pass (thread i takes cell i of both).
template <int ND>
__global__ void gfm_band_seed2_kernel(ExtrapJob j0, ExtrapJob j1, double upper,
                                      const double* __restrict__ star0,
                                      const double* __restrict__ star1,
                                      const uint8_t* __restrict__ has0,
                                      const uint8_t* __restrict__ has1, Grid g) {
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  const int64_t x = i % g.nx, y = i / g.nx;
  bool band0, needed0, band1, needed1;
  band_flags<ND>(j0.phi, upper, g, x, y, i, band0, needed0);
  band_flags<ND>(j1.phi, upper, g, x, y, i, band1, needed1);
  list_append(band0, i, j0.list, j0.count);
  list_append(band1, i, j1.list, j1.count);
  if (needed0) band_seed_cell<ND>(j0.um, star0, has0, g, x, y, i, j0.A, j0.B);
  if (needed1) band_seed_cell<ND>(j1.um, star1, has1, g, x, y, i, j1.A, j1.B);
}

__global__ void count_clear_kernel(int* count) { *count = 0; }

// GhostCellExtrapolator::invoke (ghost_fluid/ghost_fluid_extrapolator.hpp:54-202):
// the Jacobi update of band cell i from buffer A into buffer B.
template <int ND>
__device__ __forceinline__ void extrap_cell(const double* __restrict__ phi_m,
                                            const double* __restrict__ A, Grid g, double dh,
                                            int64_t i, double* __restrict__ B) {
  constexpr int S = 2 + ND;
  const double p = phi_m[i];
  const int64_t x = i % g.nx, y = i / g.nx;
  double nrm[2] = {0.0, 0.0};
  norm_at<ND>(phi_m, g, x, y, dh, nrm);
  const int sg = signum(p);
  double l1 = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    nrm[d] = -sg * nrm[d];
    l1 += fabs(nrm[d]);
  }
  double acc[S];
#pragma unroll
  for (int k = 0; k < S; ++k) acc[k] = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const int64_t dx = (d == 0), dy = (d == 1);
    // both neighbours' states are loaded with the level-set values; the
    // upwind one is taken
    const int64_t ylo = ND > 1 ? clamp1(y - dy, g.ny) : 0, yhi = ND > 1 ? clamp1(y + dy, g.ny) : 0;
    const double* lo = A + (ylo * g.nx + clamp1(x - dx, g.nx)) * S;
    const double* hi = A + (yhi * g.nx + clamp1(x + dx, g.nx)) * S;
    double ul[S], uh[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
      ul[k] = lo[k];
      uh[k] = hi[k];
    }
    const bool down = phi_at1<ND>(phi_m, g, x - dx, y - dy) < phi_at1<ND>(phi_m, g, x + dx, y + dy);
    const double wgt = l1 > 1e-2 ? fabs(nrm[d]) : 0.5;
#pragma unroll
    for (int k = 0; k < S; ++k) acc[k] += (down ? ul[k] : uh[k]) * wgt;
  }
#pragma unroll
  for (int k = 0; k < S; ++k) B[i * S + k] = (l1 > 1e-2) ? acc[k] / l1 : acc[k];
}

// One extrapolation sweep of up to two materials in one launch: the first
// kExtrapBlocks workgroups walk job 0's band list, the next kExtrapBlocks job
// 1's.  Sweep k reads buffer A (k even) or B (k odd) and writes the other; the
// last sweep also writes the band cells into the material's state (the
// write-back).  With sweep < 0 the seeded buffer A is written back as it is
// (no sweeps).
template <int ND>
__global__ void __launch_bounds__(256) gfm_extrap_kernel(ExtrapJob j0, ExtrapJob j1, Grid g,
                                                         double dh, int sweep, int last) {
  constexpr int S = 2 + ND;
  const ExtrapJob j = blockIdx.x < kExtrapBlocks ? j0 : j1;
  const int n = *j.count;
  const int first = (int)(blockIdx.x % kExtrapBlocks) * blockDim.x + threadIdx.x;
  const int stride = kExtrapBlocks * blockDim.x;
  if (sweep < 0) {
    for (int k = first; k < n; k += stride) {
      const int64_t i = j.list[k];
      store_state<S>(j.um, i, j.A + i * S);
    }
    return;
  }
  const double* src = (sweep & 1) ? j.B : j.A;
  double* dst = (sweep & 1) ? j.A : j.B;
  for (int k = first; k < n; k += stride) {
    const int64_t i = j.list[k];
    extrap_cell<ND>(j.phi, src, g, dh, i, dst);
    if (last) store_state<S>(j.um, i, dst + i * S);
  }
}

// combine_into_first (material/cuda/combine_materials.cuh:39-63): the first
// material whose phi <= 0 provides the output state.
template <int ND>
__global__ void combine_kernel(const double* __restrict__ phi0, StateView u0, StateView u1,
                               Grid g, StateView out) {
  constexpr int S = 2 + ND;
  const int64_t i = gid();
  if (i >= g.nx * g.ny) return;
  double s[S];
  const double p = phi0[i];
  load_state<S>(((p <= 0.0) || (p <= -0.0)) ? u0 : u1, g, i % g.nx, i / g.nx, s);
  store_state<S>(out, i, s);
}

#define FL_LS_INST(ND)                                                                       \
  template __global__ void velocity_kernel<ND>(const double*, const double*, StateView,      \
                                               StateView, int, Grid, double, int, double*); \
  template __global__ void velocity2_kernel<ND>(const double*, const double*, StateView,     \
                                                StateView, Grid, double, double*, double*); \
  template __global__ void rk3_stage2_kernel<ND, false>(int, Rk3Job, Rk3Job, Grid,            \
                                                        const double*, double);              \
  template __global__ void rk3_stage2_kernel<ND, true>(int, Rk3Job, Rk3Job, Grid,             \
                                                       const double*, double);               \
  template __global__ void rk3_stage_kernel<ND>(int, const double*, const double*,            \
                                                const double*, Grid, double, const double*,  \
                                                double, double*);                            \
  template __global__ void gfm_star_n_kernel<ND>(MatSet, int, Grid, double, double*,         \
                                                 uint8_t*, int*);                             \
  template __global__ void velocity_n_kernel<ND>(MatSet, int, Grid, double, double*);        \
  template __global__ void gfm_star_kernel<ND>(const double*, const double*, StateView,      \
                                               StateView, Eos, Eos, Grid, double, double*,   \
                                               uint8_t*, int*);                             \
  template __global__ void gfm_band_seed_kernel<ND>(const double*, double, StateView,       \
                                                    const double*, const uint8_t*, Grid,      \
                                                    double*, double*, int*, int*);            \
  template __global__ void gfm_star2_kernel<ND>(const double*, const double*, StateView,     \
                                                StateView, Eos, Eos, Grid, double, double*,  \
                                                double*, uint8_t*, uint8_t*, int*, int*);    \
  template __global__ void gfm_band_seed2_kernel<ND>(ExtrapJob, ExtrapJob, double,          \
                                                     const double*, const double*,           \
                                                     const uint8_t*, const uint8_t*, Grid);  \
  template __global__ void gfm_extrap_kernel<ND>(ExtrapJob, ExtrapJob, Grid, double, int,   \
                                                 int);                                        \
  template __global__ void combine_kernel<ND>(const double*, StateView, StateView, Grid,     \
                                              StateView);

FL_LS_INST(1)
FL_LS_INST(2)

}  // namespace ls
}  // namespace fl
