// fl_batch.hip -- persistent kernels of the batched Euler step.
//
// A step of a batch is three launches, captured once in a hipGraph and
// replayed: the control kernel (one thread per problem: dt from the problem's
// fused maximum wavespeed, its run time and iteration count, as
// Parameters::update_time_delta / dt_dh / update_simulation_info,
// simulator/parameters.hpp:60-99), then one persistent launch per sweep
// phase.  Phase 0 runs the x sweep of every 2D problem and the whole step of
// every 1D problem; phase 1 runs the y sweep of every 2D problem.  The
// workgroups of a phase take tiles from a shared counter (largest tiles
// first), so problems of very different sizes share one launch and the
// machine stays busy to the end of the phase.  Tile bodies are those of the
// single-grid sweeps (fl_sweep_tiles.hpp), so every problem's result is the
// single-grid result bit for bit.
#include "fl_batch.hpp"
#include "fl_sweep_tiles.hpp"

namespace fl {
namespace batch {

__global__ void init_kernel(StepCtl* ctls, int nprob) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nprob) return;
  StepCtl& c = ctls[i];
  c.ws_bits[0] = 0ull;
  c.ws_bits[1] = 0ull;
  c.dt = 0.0;
  c.dtdh = 0.0;
  c.t = 0.0;
  c.iters = 0;
  c.active = 1;
  c.pad = 0;
}

// The step's parity comes from each problem's own iteration count: the
// previous step's fused maximum sits in ws_bits[iters & 1] and this step's
// sweeps write ws_bits[(iters + 1) & 1].
__global__ void control_kernel(StepCtl* ctls, const Params* params, int nprob, Counters* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    cnt->next[0] = 0;
    cnt->next[1] = 0;
  }
  if (i >= nprob) return;
  StepCtl& c = ctls[i];
  const Params p = params[i];
  const bool active = c.t < p.sim_time && c.iters < p.max_iters;
  if (active) {
    const int parity = (int)(c.iters & 1);
    const double maxws = __longlong_as_double((long long)c.ws_bits[parity]);
    const double dt = p.cfl * p.dh / maxws;
    c.dt = dt;
    c.dtdh = dt / p.dh;
    c.t += dt;
    c.iters += 1;
    c.ws_bits[parity ^ 1] = 0ull;
  }
  c.active = active ? 1 : 0;
}

__global__ void reset_counters_kernel(Counters* cnt) {
  cnt->next[0] = 0;
  cnt->next[1] = 0;
}

// Persistent sweep phase.  timing != 0 runs every tile state -> work with a
// fixed dt/dh and no wavespeed output (launch-shape measurement; the states
// are not modified).
// FAST: the fast FP64 numerics (FP32 has its own forms); VLC: every problem
// limits with van Leer in conservative form (fixed at compile time, as in
// the single-grid sweeps).
template <typename T, int FLUX, int PHASE, bool FAST, bool VLC>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(SweepWaves<T>::value)))
phase_kernel(const Prob<T>* __restrict__ probs, const Tile* __restrict__ tiles, int ntiles,
             Counters* __restrict__ cnt, int timing, T timing_dtdh) {
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(&cnt->next[PHASE], 1);
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (tile >= ntiles) break;
    const Tile td = tiles[tile];
    const Prob<T>& P = probs[td.prob];
    SweepArgs<T> a = P.x;
    if (timing) {
      a.ctl = nullptr;
      a.dtdh_host = timing_dtdh;
      a.ws_out = nullptr;
    }
    if (PHASE == 0) {
      if (P.nd == 1) {
        if (!timing) {
          const long long it = P.ctl->iters;  // already advanced by this step's control
          if ((it - 1) & 1) {                 // odd steps run work -> state
            a.in = P.x.out;
            a.out = const_cast<T*>(P.x.in);
          }
          a.ws_out = &P.ctl->ws_bits[it & 1];
        }
        sweep_x_tile<T, 1, FLUX, FAST, VLC>(a, td.bx, 0);
      } else {
        a.ws_out = nullptr;
        sweep_x_tile<T, 2, FLUX, FAST, VLC>(a, td.bx, td.by);
      }
    } else {
      a.rows_per_block = P.rows_y;
      if (!timing) {
        a.in = P.x.out;
        a.out = const_cast<T*>(P.x.in);
        a.ws_out = &P.ctl->ws_bits[P.ctl->iters & 1];
      }
      sweep_y_tile<T, FLUX, FAST, VLC>(a, td.bx, td.by);
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}

// 1D problems that ran an odd number of steps hold their result in the work
// buffer.  dir 0 copies it to the state buffer; dir 1 copies the state buffer
// back to the work buffer (after the work buffers were used as scratch).
// One workgroup per problem.
template <typename T>
__global__ void __launch_bounds__(256) finish_kernel(const Prob<T>* __restrict__ probs, int nprob,
                                                     int dir) {
  const int p = blockIdx.x;
  if (p >= nprob) return;
  const Prob<T>& P = probs[p];
  if (P.nd != 1 || !(P.ctl->iters & 1)) return;
  const int64_t n = P.x.nx;
  const T* src = dir == 0 ? P.x.out : P.x.in;
  T* dst = dir == 0 ? const_cast<T*>(P.x.in) : P.x.out;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    T u[3];
    load_cell<T, 3>(src, i, P.x.es, P.x.cs, u);
    store_cell<T, 3>(dst, i, P.x.es, P.x.cs, u);
  }
}

// FP64 sums of each conserved component of each problem's state (4 slots per
// problem, the fourth 0 in 1D).  One workgroup per problem; the order of
// additions is fixed, so the sums are reproducible.
template <typename T>
__global__ void __launch_bounds__(256) sums_kernel(const Prob<T>* __restrict__ probs, int nprob,
                                                   double* __restrict__ sums) {
  __shared__ double red[4][256];
  const int p = blockIdx.x;
  if (p >= nprob) return;
  const Prob<T>& P = probs[p];
  const int S = 2 + P.nd;
  const int64_t n = P.x.nx * P.x.ny;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const T* c = P.x.in + i * P.x.es;
    for (int k = 0; k < S; ++k) acc[k] += (double)c[k * P.x.cs];
  }
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    double v = 0.0;
    for (int j = 0; j < (int)blockDim.x; ++j) v += red[threadIdx.x][j];
    sums[p * 4 + threadIdx.x] = v;
  }
}

#define FL_BATCH_PHASES(T, FAST, VLC)                                                    \
  template __global__ void phase_kernel<T, kFluxHllc, 0, FAST, VLC>(                        \
      const Prob<T>*, const Tile*, int, Counters*, int, T);                                  \
  template __global__ void phase_kernel<T, kFluxHllc, 1, FAST, VLC>(                        \
      const Prob<T>*, const Tile*, int, Counters*, int, T);                                  \
  template __global__ void phase_kernel<T, kFluxForce, 0, FAST, VLC>(                       \
      const Prob<T>*, const Tile*, int, Counters*, int, T);                                  \
  template __global__ void phase_kernel<T, kFluxForce, 1, FAST, VLC>(                       \
      const Prob<T>*, const Tile*, int, Counters*, int, T);
#define FL_BATCH_INST(T)                                                                     \
  FL_BATCH_PHASES(T, false, false)                                                           \
  FL_BATCH_PHASES(T, false, true)                                                            \
  template __global__ void finish_kernel<T>(const Prob<T>*, int, int);                        \
  template __global__ void sums_kernel<T>(const Prob<T>*, int, double*);

FL_BATCH_INST(double)
FL_BATCH_INST(float)
// the fast FP64 numerics
FL_BATCH_PHASES(double, true, false)
FL_BATCH_PHASES(double, true, true)

}  // namespace batch
}  // namespace fl
