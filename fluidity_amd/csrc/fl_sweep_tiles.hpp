// fl_sweep_tiles.hpp -- tile bodies of the gfx950 Euler sweeps, shared by the
// single-grid kernels (fl_sweep.hip) and the batched persistent kernels
// (fl_batch.hip).
//
// Replaces the reference's solve_dim<S> / SplitSolver::invoke sweep kernel
// (solver/cuda/split_solver.cuh:39-80, solver/split_solver.hpp:138-162), its
// set_wavespeeds_impl + reduce_impl CFL passes
// (simulator/cuda/wavespeed_initialization.cuh:41-97, algorithm/reduce.cuh:44-131)
// and the host-side dt update (simulator/parameters.hpp:60-99).
//
// Work decomposition (both 2D sweeps): a workgroup owns a band of columns and
// marches down a chunk of rows, so each input row is read from HBM once per
// sweep and the cross-dimension part of the MUSCL-Hancock flattening factor
// (a +-3 cell stencil in the non-sweep dimension) comes from per-thread
// register windows instead of re-reads.
//  * x sweep: lanes run along x.  Face states, face fluxes and the x-halo are
//    exchanged through LDS (4 barriers per row).  A workgroup of B lanes
//    evaluates B cells' face states and writes B-2 updated cells.
//  * y sweep: lanes run along x, each lane marches its own column, so the
//    +-4 cell sweep stencil, the face states and the face fluxes stay in
//    registers; only the x-direction flattening factor uses LDS.
// The fused CFL epilogue takes the max wavespeed of every updated cell,
// reduces it per workgroup and issues one 64-bit atomicMax on its bit pattern
// (non-negative doubles order like their bit patterns; NaN sorts above +inf,
// so a NaN anywhere surfaces as a NaN max).
#pragma once

#include <type_traits>

#include "fl_kernels.hpp"
#include "fl_numerics_fast.hpp"

// Waves per SIMD the sweep kernels are compiled for (register budget
// 512 / waves VGPRs), per precision.
#ifndef FL_SWEEP_WAVES
#define FL_SWEEP_WAVES 2
#endif
#ifndef FL_SWEEP_WAVES_F32
#define FL_SWEEP_WAVES_F32 4
#endif

// Rows each lane keeps in flight ahead of the row it computes, per
// precision (the FP32 kernels run twice the waves in half the registers).
// The FP64 x sweep runs faster with one row (433 against 449 us at 4096^2),
// the FP64 y sweep with two (410 against 428 us).
#ifndef FL_PREFETCH
#define FL_PREFETCH 2
#endif
#ifndef FL_PREFETCH_X
#define FL_PREFETCH_X 1
#endif
#ifndef FL_PREFETCH_F32
#define FL_PREFETCH_F32 2
#endif

// Waves per SIMD of the FP64 sweeps in the fast numerics (fewer live
// registers than the exact ones).
#ifndef FL_SWEEP_WAVES_FAST
#define FL_SWEEP_WAVES_FAST FL_SWEEP_WAVES
#endif

namespace fl {
template <typename T> struct SweepWaves { static constexpr int value = FL_SWEEP_WAVES; };
template <> struct SweepWaves<float> { static constexpr int value = FL_SWEEP_WAVES_F32; };
template <typename T> struct SweepPrefetch { static constexpr int value = FL_PREFETCH; };
template <> struct SweepPrefetch<float> { static constexpr int value = FL_PREFETCH_F32; };
template <typename T> struct SweepPrefetchX { static constexpr int value = FL_PREFETCH_X; };
template <> struct SweepPrefetchX<float> { static constexpr int value = FL_PREFETCH_F32; };
// Rows in flight of the FP64 sweeps in the fast numerics (x, y).
#ifndef FL_PREFETCH_X_FAST
#define FL_PREFETCH_X_FAST FL_PREFETCH_X
#endif
#ifndef FL_PREFETCH_FAST
#define FL_PREFETCH_FAST FL_PREFETCH
#endif
template <typename T, bool F> struct PrefetchXOf { static constexpr int value = SweepPrefetchX<T>::value; };
template <> struct PrefetchXOf<double, true> { static constexpr int value = FL_PREFETCH_X_FAST; };
template <typename T, bool F> struct PrefetchYOf { static constexpr int value = SweepPrefetch<T>::value; };
template <> struct PrefetchYOf<double, true> { static constexpr int value = FL_PREFETCH_FAST; };
// Waves per SIMD of a sweep kernel in precision T with fast (F) or exact numerics.
template <typename T, bool F> struct SweepWavesOf { static constexpr int value = SweepWaves<T>::value; };
template <> struct SweepWavesOf<double, true> { static constexpr int value = FL_SWEEP_WAVES_FAST; };
}  // namespace fl


namespace fl {

// ---------------------------------------------------------------------------
// memory access
// ---------------------------------------------------------------------------

template <typename T, int S>
__device__ __forceinline__ void load_cell(const T* __restrict__ base, int64_t idx,
                                          int64_t es, int64_t cs, T* u) {
  const T* p = base + idx * es;
#pragma unroll
  for (int k = 0; k < S; ++k) u[k] = p[k * cs];
}

template <typename T, int S>
__device__ __forceinline__ void store_cell(T* __restrict__ base, int64_t idx, int64_t es,
                                           int64_t cs, const T* u) {
  T* p = base + idx * es;
#pragma unroll
  for (int k = 0; k < S; ++k) p[k * cs] = u[k];
}

// Ghost index map (solver/boundary_loader.hpp:175-199): transmissive and
// reflective ghosts mirror the interior, ghost(-k-1) = cell(k) and
// ghost(N+k) = cell(N-1-k); periodic wraps.  i is clamped to the 4-cell pad
// first so that lanes past the end of a partial band stay in bounds.
__device__ __forceinline__ int64_t map_ghost(int64_t i, int64_t n, int kind, bool& flip) {
  i = i < -4 ? -4 : (i > n + 3 ? n + 3 : i);
  flip = false;
  if (i >= 0 && i < n) return i;
  if (kind == kBcPeriodic) return i < 0 ? i + n : i - n;
  flip = (kind == kBcReflective);
  return i < 0 ? -i - 1 : 2 * n - 1 - i;
}

// A cell load in flight.  The reflective-ghost velocity flip is applied when
// the cell is taken, not when the load is issued, so that no instruction
// touches the destination registers while the load is pending.
template <typename T, int S>
struct Pending {
  T u[S];
  int flip;  // bit 0: reflect in x, bit 1: reflect in y
};

template <typename T, int ND>
__device__ __forceinline__ void issue_load(const SweepArgs<T>& a, int64_t x, int64_t y,
                                           Pending<T, 2 + ND>& q) {
  bool fx = false, fy = false;
  const int64_t xi = map_ghost(x, a.nx, a.bcx, fx);
  int64_t yi = 0;
  if (ND > 1) yi = map_ghost(y, a.ny, a.bcy, fy);
  load_cell<T, 2 + ND>(a.in, yi * a.nx + xi, a.es, a.cs, q.u);
  q.flip = (fx ? 1 : 0) | (fy ? 2 : 0);
}

// Loads and stores of the 2D sweeps.  The column a lane reads stays fixed
// along its march and the row is workgroup-uniform, so a cell's address is a
// scalar row base (plus the component offset) and a per-lane 32-bit byte
// offset computed once: the loads and stores take the scalar-base address
// form and spend no vector instructions on addresses.  Rows of 2D grids are
// below 2^32 bytes (checked on the host).
struct LaneCol {
  uint32_t off;  // byte offset of the lane's ghost-mapped column within a row
  int flip;      // 1: a reflective ghost column (x momentum flips)
};

template <typename T>
__device__ __forceinline__ LaneCol lane_col(const SweepArgs<T>& a, int64_t x) {
  bool fx = false;
  const int64_t xi = map_ghost(x, a.nx, a.bcx, fx);
  LaneCol l;
  l.off = (uint32_t)(xi * a.es * (int64_t)sizeof(T));
  l.flip = fx ? 1 : 0;
  return l;
}

// Byte offset of column x (in the grid) for stores.
template <typename T>
__device__ __forceinline__ uint32_t col_off(const SweepArgs<T>& a, int64_t x) {
  return (uint32_t)(x * a.es * (int64_t)sizeof(T));
}

// Row index of a 2D grid (ny < 2^31) through the ghost map, in 32 bits so
// that the workgroup-uniform row arithmetic stays on the scalar unit.
__device__ __forceinline__ int map_row(int y, int ny, int kind, bool& flip) {
  y = y < -4 ? -4 : (y > ny + 3 ? ny + 3 : y);
  flip = false;
  if (y >= 0 && y < ny) return y;
  if (kind == kBcPeriodic) return y < 0 ? y + ny : y - ny;
  flip = (kind == kBcReflective);
  return y < 0 ? -y - 1 : 2 * ny - 1 - y;
}

template <typename T, int S>
__device__ __forceinline__ void issue_row(const SweepArgs<T>& a, int64_t y, const LaneCol& lc,
                                          Pending<T, S>& q) {
  bool fy = false;
  const int yi = map_row((int)y, (int)a.ny, a.bcy, fy);
  const char* row = (const char*)a.in + (int64_t)yi * (a.nx * a.es * (int64_t)sizeof(T));
  // The component row bases stay opaque scalars and the lane offset an
  // opaque 32-bit value, so that the compiler cannot fold them into hoisted
  // 64-bit per-lane pointers (vector address arithmetic on every row): each
  // load is a scalar base plus a 32-bit lane offset.
  uint32_t off = lc.off;
  asm volatile("" : "+v"(off));
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const char* base = row + k * a.cs * (int64_t)sizeof(T);
    asm volatile("" : "+s"(base));
    q.u[k] = *(const __attribute__((address_space(1))) T*)(
        (const __attribute__((address_space(1))) char*)base + off);
  }
  q.flip = lc.flip | (fy ? 2 : 0);
}

template <typename T, int S>
__device__ __forceinline__ void store_row(const SweepArgs<T>& a, int64_t y, uint32_t off,
                                          const T* u) {
  char* row = (char*)a.out + y * a.nx * a.es * (int64_t)sizeof(T);
  asm volatile("" : "+v"(off));
#pragma unroll
  for (int k = 0; k < S; ++k) {
    char* base = row + k * a.cs * (int64_t)sizeof(T);
    asm volatile("" : "+s"(base));
    *(__attribute__((address_space(1))) T*)((__attribute__((address_space(1))) char*)base + off) =
        u[k];
  }
}

template <typename T, int ND>
__device__ __forceinline__ void take(const Pending<T, 2 + ND>& q, T* u) {
#pragma unroll
  for (int k = 0; k < 2 + ND; ++k) u[k] = q.u[k];
  if (q.flip & 1) reflect<T, ND>(u, 0);
  if (ND > 1 && (q.flip & 2)) reflect<T, ND>(u, 1);
}

// ---------------------------------------------------------------------------
// fused CFL epilogue
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long ws_bits(double w) {
  w = (w == 0.0) ? 0.0 : w;  // -0 -> +0 so that the bit order is the value order
  return (unsigned long long)__double_as_longlong(w);
}

// A 64-bit key of a double whose unsigned order is the order of the values
// (-0 and +0 share one key); every NaN maps to the largest key, so a NaN
// among the values surfaces as the maximum.  Key 0 is below every value.
__device__ __forceinline__ unsigned long long ordered_key(double v) {
  if (v != v) return ~0ull;
  v = (v == 0.0) ? 0.0 : v;
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffull), off, 64);
    const unsigned hi = __shfl_xor((unsigned)(v >> 32), off, 64);
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    v = o > v ? o : v;
  }
  return v;
}

// Block-wide max of v; thread 0 issues one atomicMax.  Must be reached by every
// thread of the block.
__device__ __forceinline__ void block_atomic_max(unsigned long long v, unsigned long long* dst) {
  __shared__ unsigned long long red[kMaxBlock / 64];
  v = wave_max_u64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) v = red[i] > v ? red[i] : v;
    atomicMax(dst, v);
  }
}

template <typename T, int ND>
__device__ __forceinline__ double cell_wavespeed64(const T* u, const Material<double>& m) {
  double ud[2 + ND];
#pragma unroll
  for (int k = 0; k < 2 + ND; ++k) ud[k] = (double)u[k];
  const Prim<double, ND> w = prim_of<double, ND>(ud, m);
  return max_wavespeed<double, ND>(ud, w, m);
}

// The numerics a sweep runs: the reference's operation order (F = false,
// FP64 bit parity) or the fast forms of fl_numerics_fast.hpp (FP64 on
// request, FP32 always).
template <typename T, bool FAST>
struct UseFast { static constexpr bool value = FAST || std::is_same<T, float>::value; };

// Halo lanes of a tile of B lanes: the first kHalo lanes stage the left halo
// cells (index 0..2, LDS slots 0..2), the last kHalo lanes the right ones
// (index 3..5, slots B+3..B+5), so that the first and the last wave of a
// workgroup share the halo work.  The other lanes load one of the six halo
// cells as well (unused), which keeps the halo loads unconditional.
__device__ __forceinline__ int halo_index(int t, int B) {
  return t < kHalo ? t : (t >= B - kHalo ? t - (B - 2 * kHalo) : t % (2 * kHalo));
}
__device__ __forceinline__ bool halo_lane(int t, int B) { return t < kHalo || t >= B - kHalo; }
// the lanes that compute the flattening factor just outside the lane range
__device__ __forceinline__ bool halo_chi_lane(int t, int B) { return t == 0 || t == B - 1; }

// min of the flattening factors: the reference's comparison form, or one
// v_min in the fast numerics
template <bool F, typename T>
__device__ __forceinline__ T min_sel(T a, T b) {
  if constexpr (F) return fast::mn(a, b);
  else return smin(a, b);
}

template <bool F, typename T, int ND>
__device__ __forceinline__ Prim<T, ND> prim_sel(const T* U, const Material<T>& m) {
  if constexpr (F) return fast::as_prim(fast::prim_of<T, ND>(U, m));
  else return prim_of<T, ND>(U, m);
}

template <bool F, typename T>
__device__ __forceinline__ T chi_sel(T a, T b, T c, T d, T e, T f, const Material<T>& m) {
  if constexpr (F) return fast::chi_factor(a, b, c, d, e, f, m);
  else return chi_factor(a, b, c, d, e, f, m);
}

template <bool F, typename T, int ND, int D>
__device__ __forceinline__ void evolve_sel(const T* um, const T* u0, const T* up, T chi,
                                           T half_dtdh, int limiter, int form,
                                           const Material<T>& m, T* R, T* L) {
  if constexpr (F) fast::evolve_faces<T, ND, D>(um, u0, up, chi, half_dtdh, limiter, form, m, R, L);
  else evolve_faces<T, ND, D>(um, u0, up, chi, half_dtdh, limiter, form, m, R, L);
}

template <bool F, typename T, int ND, int D, int FLUX>
__device__ __forceinline__ void flux_sel(const T* ul, const T* ur, T dtdh, T dhdt,
                                         const Material<T>& m, T* G) {
  if constexpr (F) fast::numerical_flux<T, ND, D, FLUX>(ul, ur, dtdh, dhdt, m, G);
  else numerical_flux<T, ND, D, FLUX>(ul, ur, dtdh, dhdt, m, G);
}

// Wavespeed of an updated cell for the fused CFL epilogue, as an FP64 value
// for the FP64 reduction: exact FP64, fast FP64, or computed in FP32 from an
// FP32 cell.
template <bool F, typename T, int ND>
__device__ __forceinline__ double cell_ws(const T* u, const SweepArgs<T>& a) {
  if constexpr (!F) {
    return cell_wavespeed64<T, ND>(u, a.mat64);
  } else {
    return (double)fast::max_wavespeed<T, ND>(fast::prim_of<T, ND>(u, a.mat), a.mat);
  }
}

// Folds an updated cell's wavespeed into the lane's running maximum (bit
// pattern).  Fast FP64 forms: with c^2 = gamma p / rho and t = max - |v|max,
// a cell whose c^2 <= t^2 (1 + 1e-14) cannot raise the maximum by more than
// about 5e-15 of it (well inside the fast numerics' bound), so a wave whose
// every lane holds such a cell skips the square roots.  The slack admits the
// cells of a uniform region, whose t rounds to just below their own c.  A
// NaN or infinite value fails the test and is folded in, so it surfaces.
template <bool F, typename T, int ND>
__device__ __forceinline__ void ws_fold(const T* u, const SweepArgs<T>& a,
                                        unsigned long long& wsmax) {
  if constexpr (F && std::is_same<T, double>::value) {
    const fast::PrimR<double, ND> w = fast::prim_of<double, ND>(u, a.mat);
    double mv = fabs(w.vel[0]);
#pragma unroll
    for (int i = 1; i < ND; ++i) mv = fast::mx(mv, fabs(w.vel[i]));
    const double c2 = a.mat.stiff ? a.mat.gamma * (w.p + a.mat.p_inf) * w.rinv
                                  : a.mat.gamma * w.p * w.rinv;
    const double t = __longlong_as_double((long long)wsmax) - mv;
    if (__all(t > 0.0 && c2 <= t * t * (1.0 + 1e-14))) return;
    const unsigned long long b = ws_bits(mv + fast::sqt(c2));
    wsmax = b > wsmax ? b : wsmax;
  } else {
    const unsigned long long b = ws_bits(cell_ws<F, T, ND>(u, a));
    wsmax = b > wsmax ? b : wsmax;
  }
}

template <typename T>
__device__ __forceinline__ bool sweep_dtdh(const SweepArgs<T>& a, T& dtdh) {
  if (a.ctl) {
    if (!a.ctl->active) return false;
    dtdh = (T)a.ctl->dtdh;
  } else {
    dtdh = a.dtdh_host;
  }
  return true;
}

// ---------------------------------------------------------------------------
// x sweep (also the 1D sweep)
// ---------------------------------------------------------------------------

// One x-sweep tile: columns [bx (B-2), (bx+1) (B-2)) of rows
// [by rows_per_block, (by+1) rows_per_block), B = blockDim.x.  In 1D a tile
// is rows_per_block consecutive segments of B-2 cells, starting at segment
// bx rows_per_block, marched like rows.  Every thread of the workgroup must
// call it with the same arguments.
template <typename T, int ND, int FLUX, bool FAST = false, bool VLC = false>
__device__ __forceinline__ void sweep_x_tile(const SweepArgs<T>& a, int64_t bx, int64_t by) {
  constexpr int S = 2 + ND;
  constexpr bool F = UseFast<T, FAST>::value;
  // VLC: the scheme is known to limit with van Leer in conservative form, so
  // the per-component limiter dispatch folds away at compile time.
  const int limiter = VLC ? int(kLimVanLeer) : a.limiter;
  const int form = VLC ? int(kFormCons) : a.form;
  // LDS sized for this launch's block width (dynamic shared memory of
  // x_tile_lds_bytes): states, pressures, x velocities and flattening
  // factors of the B + 6 staged columns, then the left face states and
  // fluxes of the B lanes.
  extern __shared__ __align__(16) unsigned char fl_x_lds[];
  const int W = (int)blockDim.x + 2 * kHalo;
  T* const sU = reinterpret_cast<T*>(fl_x_lds);  // [S][W]
  T* const sP = sU + S * W;
  T* const sV = sP + W;
  T* const sK = sV + W;
  T* const sL = sK + W;                          // [S][B]
  T* const sG = sL + S * (int)blockDim.x;        // [S][B]

  T dtdh;
  if (!sweep_dtdh(a, dtdh)) return;
  const T half_dtdh = T(0.5) * dtdh;
  const T dhdt = T(1) / dtdh;

  const int B = blockDim.x;
  const int t = threadIdx.x;
  const int s = t + kHalo;                       // LDS slot of this lane's column
  const int64_t seg0 = (ND == 2) ? bx : bx * a.rows_per_block;   // first column band
  const int64_t x0 = seg0 * (B - 2);
  const int64_t c = x0 - 1 + t;                  // column whose face states this lane builds
  const int64_t y0 = (ND == 2) ? by * a.rows_per_block : 0;
  const int64_t yend = (ND == 2) ? min(y0 + (int64_t)a.rows_per_block, a.ny) : 1;
  // 1D: segments of this tile, at most rows_per_block and not past the grid
  const int64_t nseg1 = (a.nx + B - 3) / (B - 2) - seg0;
  const int64_t segs = (ND == 2) ? 1 : min((int64_t)a.rows_per_block, nseg1);

  // Halo lanes (halo_index) stage columns x0-4..x0-2 (slots 0..2) and
  // x0+B-1..x0+B+1 (slots B+3..B+5) of the output row.
  const int hl = halo_index(t, B);
  const int hslot = (hl < 3) ? hl : B + hl;
  const int64_t hcol = x0 - 1 - kHalo + hslot;
  // column of this lane and its halo cell for row (1D: segment) r
  const int64_t cstep = (ND == 2) ? 0 : B - 2;

  // Register windows along y (2D): rows r-3..r of U, u_x; rows r-4..r of p;
  // rows r-3..r of u_y; K_y of rows r-4..r-2.
  T Uw[4][S];
  T pw[5], vxw[4], vyw[4];
  // K_y window as named scalars: an array indexed by the sign of the pressure
  // difference would be demoted to scratch memory.
  T ky0 = T(1), ky1 = T(1), ky2 = T(1);
#pragma unroll
  for (int i = 0; i < 5; ++i) pw[i] = T(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) { vxw[i] = T(0); vyw[i] = T(0); }

  unsigned long long wsmax = 0ull;
  const int64_t rbeg = (ND == 2) ? y0 - 3 : 0;
  const int64_t rlast = (ND == 2) ? yend + 2 : segs - 1;

  // Software pipelining with PF row buffers in fixed registers: the loads of
  // row r+PF and of the halo cells of the output row PF steps ahead are issued
  // while row r is computed, without ever copying a register whose load is
  // still in flight.
  // All loads are unconditional (rows past the end are re-loaded, clamped), so
  // the compiler can count the loads issued after each one and wait only for
  // the one it needs.
  auto hrow = [&](int64_t r) { return (ND == 2) ? r - 3 : int64_t(0); };
  // 2D: the lane's column and halo column, fixed along the march
  const LaneCol lc = lane_col(a, c), lh = lane_col(a, hcol);
  const uint32_t out_off = col_off(a, c);
  auto issue = [&](int64_t x, int64_t r, const LaneCol& l, Pending<T, S>& q) {
    if constexpr (ND == 2) issue_row<T, S>(a, r, l, q);
    else issue_load<T, ND>(a, x, r, q);
  };
  constexpr int PF = PrefetchXOf<T, FAST>::value;
  Pending<T, S> Qs[PF], Hs[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i) {
    const int64_t ri = min(rbeg + i, rlast);
    issue(c + ri * cstep, ri, lc, Qs[i]);
    issue(hcol + ri * cstep, hrow(ri), lh, Hs[i]);
  }

  // Step i of each group of PF rows owns buffer i.  The inner loop is fully
  // unrolled so that each copy names its buffers statically.
  for (int64_t r0 = rbeg; r0 <= rlast; r0 += PF) {
#pragma unroll
  for (int step = 0; step < PF; ++step) {
    const int64_t r = r0 + step;
    if (r > rlast) break;
    Pending<T, S>& Q = Qs[step];
    Pending<T, S>& H = Hs[step];
    // ---- the newest row of this lane's column and its primitives
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int k = 0; k < S; ++k) Uw[i][k] = Uw[i + 1][k];
      vxw[i] = vxw[i + 1];
      vyw[i] = vyw[i + 1];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) pw[i] = pw[i + 1];
    take<T, ND>(Q, Uw[3]);
    const int64_t rn = min(r + PF, rlast);
    issue(c + rn * cstep, rn, lc, Q);
    T uh[S];  // halo cell of output row hrow(r)
    take<T, ND>(H, uh);
    issue(hcol + rn * cstep, hrow(rn), lh, H);
    {
      const Prim<T, ND> w = prim_sel<F, T, ND>(Uw[3], a.mat);
      pw[4] = w.p;
      vxw[3] = w.vel[0];
      vyw[3] = w.vel[ND - 1];
    }
    if (ND == 2) {
      if (r >= y0 + 1) {  // K_y of row r-2 (needs rows r-4..r)
        ky0 = ky1;
        ky1 = ky2;
        ky2 = chi_sel<F>(pw[0], pw[1], pw[3], pw[4], vyw[0], vyw[2], a.mat);
      }
      if (r < y0 + 3) continue;  // block-uniform: warm-up rows
    }
    const int64_t jo = (ND == 2) ? r - 3 : 0;    // output row
    const T* U0 = (ND == 2) ? Uw[0] : Uw[3];
    const T p0 = (ND == 2) ? pw[1] : pw[4];
    const T vx0 = (ND == 2) ? vxw[0] : vxw[3];

    // ---- phase A: stage the row's states, pressures and x velocities
#pragma unroll
    for (int k = 0; k < S; ++k) sU[k * W + s] = U0[k];
    sP[s] = p0;
    sV[s] = vx0;
    if (halo_lane(t, B)) {
      const Prim<T, ND> wh = prim_sel<F, T, ND>(uh, a.mat);
#pragma unroll
      for (int k = 0; k < S; ++k) sU[k * W + hslot] = uh[k];
      sP[hslot] = wh.p;
      sV[hslot] = wh.vel[0];
    }
    __syncthreads();

    // ---- phase B: x flattening factors K_x (own column, plus the two
    // columns just outside the lane range)
    sK[s] = chi_sel<F>(sP[s - 2], sP[s - 1], sP[s + 1], sP[s + 2], sV[s - 1], sV[s + 1], a.mat);
    if (halo_chi_lane(t, B)) {
      const int hs = (t == 0) ? kHalo - 1 : B + kHalo;
      sK[hs] = chi_sel<F>(sP[hs - 2], sP[hs - 1], sP[hs + 1], sP[hs + 2], sV[hs - 1],
                         sV[hs + 1], a.mat);
    }
    const int sx = signum(sP[s + 1] - sP[s - 1]);
    T um[S], up[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
      um[k] = sU[k * W + s - 1];
      up[k] = sU[k * W + s + 1];
    }
    __syncthreads();

    // ---- phase C: chi (muscl_reconstructor.hpp:106-114) and face states
    T chi = min_sel<F>(Limits<T>::max(), min_sel<F>(sK[s], sK[s + sx]));
    if (ND == 2) {
      const int sy = signum(pw[2] - pw[0]);
      const T ky_nb = sy > 0 ? ky2 : (sy < 0 ? ky0 : ky1);
      chi = min_sel<F>(chi, min_sel<F>(ky1, ky_nb));
    }
    T R[S], L[S];
    evolve_sel<F, T, ND, 0>(um, U0, up, chi, half_dtdh, limiter, form, a.mat, R, L);
#pragma unroll
    for (int k = 0; k < S; ++k) sL[k * B + t] = L[k];
    __syncthreads();

    // ---- phase D: face flux between column c and c+1
    T G[S];
    if (t < B - 1) {
      T Lr[S];
#pragma unroll
      for (int k = 0; k < S; ++k) Lr[k] = sL[k * B + t + 1];
      flux_sel<F, T, ND, 0, FLUX>(R, Lr, dtdh, dhdt, a.mat, G);
#pragma unroll
      for (int k = 0; k < S; ++k) sG[k * B + t] = G[k];
    }
    __syncthreads();

    // ---- phase E: conservative update U + dtdh (F_{c-1/2} - F_{c+1/2})
    const int64_t cr = c + r * cstep;
    if (t >= 1 && t <= B - 2 && cr < a.nx) {
      T o[S];
#pragma unroll
      for (int k = 0; k < S; ++k) o[k] = U0[k] + (sG[k * B + t - 1] - G[k]) * dtdh;
      if constexpr (ND == 2) store_row<T, S>(a, jo, out_off, o);
      else store_cell<T, S>(a.out, jo * a.nx + cr, a.es, a.cs, o);
      if (a.ws_out) {
        ws_fold<F, T, ND>(o, a, wsmax);
      }
    }
  }
  }
  if (a.ws_out) block_atomic_max(wsmax, a.ws_out);
}

// ---------------------------------------------------------------------------
// y sweep
// ---------------------------------------------------------------------------

// One y-sweep tile: columns [bx B, (bx+1) B) of rows [by rows_per_block,
// (by+1) rows_per_block), B = blockDim.x.  Workgroup-uniform arguments.
template <typename T, int FLUX, bool FAST = false, bool VLC = false>
__device__ __forceinline__ void sweep_y_tile(const SweepArgs<T>& a, int64_t bx, int64_t by) {
  constexpr int ND = 2, S = 4;
  constexpr bool F = UseFast<T, FAST>::value;
  const int limiter = VLC ? int(kLimVanLeer) : a.limiter;  // as in the x sweep
  const int form = VLC ? int(kFormCons) : a.form;
  const int B = blockDim.x;
  const int W = B + 2 * kHalo;
  // LDS sized for this launch's block width (dynamic shared memory of
  // y_tile_lds_bytes): pressures, x velocities and flattening factors of the
  // B + 6 columns of the newest row, then a ring of the last 5 rows of this
  // workgroup's states, per wave [slot][component][lane of 64] (so that the
  // component offsets are immediates).  A lane reads back only its
  // own column of the ring, so the ring needs no barriers; it holds the rows
  // the face states of row r - 3 need while rows r - 1 and r finish their
  // flattening factors, which would otherwise stay in registers.
  extern __shared__ __align__(16) unsigned char fl_y_lds[];
  T* const sP = reinterpret_cast<T*>(fl_y_lds);
  T* const sV = sP + W;
  T* const sK = sV + W;
  T* const ring = sK + W;                        // [B / 64][kYRing][S][64]

  T dtdh;
  if (!sweep_dtdh(a, dtdh)) return;
  const T half_dtdh = T(0.5) * dtdh;
  const T dhdt = T(1) / dtdh;

  const int t = threadIdx.x;
  const int s = t + kHalo;
  const int64_t x0 = bx * B;
  const int64_t c = x0 + t;
  const int64_t y0 = by * a.rows_per_block;
  const int64_t yend = min(y0 + (int64_t)a.rows_per_block, a.ny);

  // Halo lanes (halo_index) stage columns x0-3..x0-1 (slots 0..2) and
  // x0+B..x0+B+2 (slots B+3..B+5).
  const int hl = halo_index(t, B);
  const int hslot = (hl < 3) ? hl : B + hl;
  const int64_t hcol = x0 - kHalo + hslot;

  // Windows along y: p rows r-4..r, u_y rows r-3..r, K_y rows r-4..r-2,
  // x flattening chi_x rows r-3..r; the states of rows r-4..r in the ring.
  T pw[5], vyw[4], cxw[4];
  T ky0 = T(1), ky1 = T(1), ky2 = T(1);  // named scalars (see the x sweep)
#pragma unroll
  for (int i = 0; i < 5; ++i) pw[i] = T(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) { vyw[i] = T(0); cxw[i] = T(1); }
  T Rprev[S], Gprev[S];
#pragma unroll
  for (int k = 0; k < S; ++k) { Rprev[k] = T(0); Gprev[k] = T(0); }

  unsigned long long wsmax = 0ull;
  // Software pipelining with PF row buffers in fixed registers (see the x
  // sweep): rows r+PF (and their halo cells) load while row r is computed.
  const int64_t rbeg = y0 - 4, rlast = yend + 3;
  // All loads are unconditional (see the x sweep); the halo cells of row r are
  // loaded on the same schedule as the row itself.
  constexpr int PF = PrefetchYOf<T, FAST>::value;
  const LaneCol lc = lane_col(a, c), lh = lane_col(a, hcol);
  const uint32_t out_off = col_off(a, c);
  Pending<T, S> Qs[PF], Hs[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i) {
    issue_row<T, S>(a, min(rbeg + i, rlast), lc, Qs[i]);
    issue_row<T, S>(a, min(rbeg + i, rlast), lh, Hs[i]);
  }
  // ring slot of row r: (r - rbeg) mod kYRing, advanced per row
  int slot = 0;
  T* const ring_t = ring + (t >> 6) * (kYRing * S * 64) + (t & 63);
  // Step i of each group of PF rows owns buffer i (see the x sweep).
  for (int64_t r0 = rbeg; r0 <= rlast; r0 += PF) {
#pragma unroll
  for (int step = 0; step < PF; ++step) {
    const int64_t r = r0 + step;
    if (r > rlast) break;
    Pending<T, S>& Q = Qs[step];
    Pending<T, S>& H = Hs[step];
#pragma unroll
    for (int i = 0; i < 4; ++i) pw[i] = pw[i + 1];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      vyw[i] = vyw[i + 1];
      cxw[i] = cxw[i + 1];
    }
    T un[S];
    take<T, ND>(Q, un);
    issue_row<T, S>(a, min(r + PF, rlast), lc, Q);
    T uh[S];
    take<T, ND>(H, uh);
    issue_row<T, S>(a, min(r + PF, rlast), lh, H);
#pragma unroll
    for (int k = 0; k < S; ++k) ring_t[(slot * S + k) * 64] = un[k];
    const Prim<T, ND> w = prim_sel<F, T, ND>(un, a.mat);
    pw[4] = w.p;
    vyw[3] = w.vel[1];

    // ---- x flattening of row r: min(K_x(c), K_x(c + sign_x))
    if (r >= y0 - 1) {  // block-uniform
      sP[s] = w.p;
      sV[s] = w.vel[0];
      if (halo_lane(t, B)) {
        const Prim<T, ND> wh = prim_sel<F, T, ND>(uh, a.mat);
        sP[hslot] = wh.p;
        sV[hslot] = wh.vel[0];
      }
      __syncthreads();
      sK[s] = chi_sel<F>(sP[s - 2], sP[s - 1], sP[s + 1], sP[s + 2], sV[s - 1], sV[s + 1], a.mat);
      if (halo_chi_lane(t, B)) {
        const int hs = (t == 0) ? kHalo - 1 : B + kHalo;
        sK[hs] = chi_sel<F>(sP[hs - 2], sP[hs - 1], sP[hs + 1], sP[hs + 2], sV[hs - 1],
                         sV[hs + 1], a.mat);
      }
      const int sx = signum(sP[s + 1] - sP[s - 1]);
      __syncthreads();
      cxw[3] = min_sel<F>(sK[s], sK[s + sx]);
    }

    // ---- y flattening K_y of row r-2
    if (r >= y0) {
      ky0 = ky1;
      ky1 = ky2;
      ky2 = chi_sel<F>(pw[0], pw[1], pw[3], pw[4], vyw[0], vyw[2], a.mat);
    }

    // ---- face states of row jr = r-3 (rows r-4..r-2 from the ring)
    if (r >= y0 + 2) {
      const int s4 = slot == kYRing - 1 ? 0 : slot + 1;  // row r-4
      const int s3 = s4 == kYRing - 1 ? 0 : s4 + 1;      // row r-3
      const int s2 = s3 == kYRing - 1 ? 0 : s3 + 1;      // row r-2
      T um[S], u0[S], up[S];
#pragma unroll
      for (int k = 0; k < S; ++k) {
        um[k] = ring_t[(s4 * S + k) * 64];
        u0[k] = ring_t[(s3 * S + k) * 64];
        up[k] = ring_t[(s2 * S + k) * 64];
      }
      const int sy = signum(pw[2] - pw[0]);
      T chi = min_sel<F>(Limits<T>::max(), cxw[0]);
      const T ky_nb = sy > 0 ? ky2 : (sy < 0 ? ky0 : ky1);
      chi = min_sel<F>(chi, min_sel<F>(ky1, ky_nb));
      T R[S], L[S];
      evolve_sel<F, T, ND, 1>(um, u0, up, chi, half_dtdh, limiter, form, a.mat, R, L);
      if (r >= y0 + 3) {
        // face between rows jr-1 and jr
        T G[S];
        flux_sel<F, T, ND, 1, FLUX>(Rprev, L, dtdh, dhdt, a.mat, G);
        const int64_t jo = r - 4;
        if (jo >= y0 && c < a.nx) {
          T o[S];
#pragma unroll
          for (int k = 0; k < S; ++k) o[k] = um[k] + (Gprev[k] - G[k]) * dtdh;
          store_row<T, S>(a, jo, out_off, o);
          if (a.ws_out) {
            ws_fold<F, T, ND>(o, a, wsmax);
          }
        }
#pragma unroll
        for (int k = 0; k < S; ++k) Gprev[k] = G[k];
      }
#pragma unroll
      for (int k = 0; k < S; ++k) Rprev[k] = R[k];
    }
    slot = slot == kYRing - 1 ? 0 : slot + 1;
  }
  }
  if (a.ws_out) block_atomic_max(wsmax, a.ws_out);
}

}  // namespace fl
