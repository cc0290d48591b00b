// fl_sweep.hip -- gfx950 kernels of the dimensionally split Euler step on one
// grid: the x and y sweeps (tile bodies in fl_sweep_tiles.hpp), the
// standalone wavespeed max, the device step control and FP64 sums.
#include "fl_sweep_tiles.hpp"

#ifndef FL_XCD_REMAP
#define FL_XCD_REMAP 0
#endif

namespace fl {

// Tile of this workgroup.  With FL_XCD_REMAP, consecutive workgroups of one
// XCD (dispatch is round-robin over the 8 XCDs) take consecutive tiles down a
// column band, so the rows that vertically adjacent tiles both read meet in
// that XCD's L2.
__device__ __forceinline__ void tile_of_block(int64_t& bx, int64_t& by) {
#if FL_XCD_REMAP
  const int64_t gx = gridDim.x, gy = gridDim.y, T = gx * gy;
  const int64_t p = blockIdx.x + gx * (int64_t)blockIdx.y;
  const int64_t xcd = p % 8, k = p / 8;
  const int64_t base = T / 8, extra = T % 8;
  const int64_t tile = xcd * base + (xcd < extra ? xcd : extra) + k;
  bx = tile / gy;
  by = tile % gy;
#else
  bx = blockIdx.x;
  by = blockIdx.y;
#endif
}

// VLC: van Leer in conservative form, fixed at compile time (the launcher
// picks these kernels when the scheme says so).
template <typename T, int ND, int FLUX, bool FAST, bool VLC>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(SweepWavesOf<T, FAST>::value)))
sweep_x_kernel(SweepArgs<T> a) {
  int64_t bx, by;
  tile_of_block(bx, by);
  sweep_x_tile<T, ND, FLUX, FAST, VLC>(a, bx, by);
}

template <typename T, int FLUX, bool FAST, bool VLC>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(SweepWavesOf<T, FAST>::value)))
sweep_y_kernel(SweepArgs<T> a) {
  int64_t bx, by;
  tile_of_block(bx, by);
  sweep_y_tile<T, FLUX, FAST, VLC>(a, bx, by);
}

// The same sweeps over two grids of one shape in one launch (the two
// materials of a two-material step): blockIdx.z selects the grid.
template <typename T, int ND, int FLUX, bool FAST, bool VLC>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(SweepWaves<T>::value)))
sweep_x2_kernel(SweepArgs<T> a0, SweepArgs<T> a1) {
  int64_t bx, by;
  tile_of_block(bx, by);
  sweep_x_tile<T, ND, FLUX, FAST, VLC>(blockIdx.z ? a1 : a0, bx, by);
}

template <typename T, int FLUX, bool FAST, bool VLC>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(SweepWaves<T>::value)))
sweep_y2_kernel(SweepArgs<T> a0, SweepArgs<T> a1) {
  int64_t bx, by;
  tile_of_block(bx, by);
  sweep_y_tile<T, FLUX, FAST, VLC>(blockIdx.z ? a1 : a0, bx, by);
}

// ---------------------------------------------------------------------------
// standalone wavespeed max, step control
// ---------------------------------------------------------------------------

// Grid-stride max over the cells; each thread issues the loads of
// kMaxWsUnroll cells before computing any of them.  The launch keeps the
// grid small (kMaxWsBlocks) because every workgroup ends in one atomicMax on
// the same address, and those serialise in L2.
template <typename T, int ND>
__global__ void __launch_bounds__(kMaxBlock) max_wavespeed_kernel(const T* __restrict__ in,
                                                                   int64_t es, int64_t cs,
                                                                   int64_t ncells,
                                                                   Material<double> m64,
                                                                   unsigned long long* dst) {
  constexpr int U = kMaxWsUnroll;
  unsigned long long v = 0ull;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < ncells; i += U * stride) {
    T u[U][2 + ND];
#pragma unroll
    for (int k = 0; k < U; ++k) load_cell<T, 2 + ND>(in, i + k * stride, es, cs, u[k]);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const unsigned long long b = ws_bits(cell_wavespeed64<T, ND>(u[k], m64));
      v = b > v ? b : v;
    }
  }
  for (; i < ncells; i += stride) {
    T u[2 + ND];
    load_cell<T, 2 + ND>(in, i, es, cs, u);
    const unsigned long long b = ws_bits(cell_wavespeed64<T, ND>(u, m64));
    v = b > v ? b : v;
  }
  block_atomic_max(v, dst);
}

// The max over two grids of one shape (the two materials of a two-material
// step; each grid with its own element and component strides): each thread
// takes its cells of both.
template <typename T, int ND, bool FAST>
__device__ __forceinline__ double ws2_cell(const T* u, const Material<double>& m) {
  if constexpr (FAST && std::is_same<T, double>::value)
    return fast::max_wavespeed<double, ND>(fast::prim_of<double, ND>(u, m), m);
  else
    return cell_wavespeed64<T, ND>(u, m);
}

// FAST (FP64): the fast forms of the wavespeed (fast numerics requested).
template <typename T, int ND, bool FAST>
__global__ void __launch_bounds__(kMaxBlock) max_wavespeed2_kernel(
    const T* __restrict__ in0, const T* __restrict__ in1, int64_t es0, int64_t cs0, int64_t es1,
    int64_t cs1, int64_t ncells, Material<double> m0, Material<double> m1,
    unsigned long long* dst) {
  constexpr int U = kMaxWsUnroll;
  unsigned long long v = 0ull;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U / 2 - 1) * stride < ncells; i += (U / 2) * stride) {
    T u[U][2 + ND];
#pragma unroll
    for (int k = 0; k < U / 2; ++k) {
      load_cell<T, 2 + ND>(in0, i + k * stride, es0, cs0, u[2 * k]);
      load_cell<T, 2 + ND>(in1, i + k * stride, es1, cs1, u[2 * k + 1]);
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const unsigned long long b = ws_bits(ws2_cell<T, ND, FAST>(u[k], (k & 1) ? m1 : m0));
      v = b > v ? b : v;
    }
  }
  for (; i < ncells; i += stride) {
    T u[2][2 + ND];
    load_cell<T, 2 + ND>(in0, i, es0, cs0, u[0]);
    load_cell<T, 2 + ND>(in1, i, es1, cs1, u[1]);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const unsigned long long b = ws_bits(ws2_cell<T, ND, FAST>(u[k], k ? m1 : m0));
      v = b > v ? b : v;
    }
  }
  block_atomic_max(v, dst);
}

// Per-cell FP64 wavespeeds max_d |u_d| + a of a grid into a device array
// (the wavespeed tensor that the reference's set_wavespeeds fills).
template <typename T, int ND>
__global__ void __launch_bounds__(kMaxBlock) wavespeeds_kernel(const T* __restrict__ in,
                                                                int64_t es, int64_t cs,
                                                                int64_t ncells,
                                                                Material<double> m64,
                                                                double* __restrict__ ws) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncells; i += stride) {
    T u[2 + ND];
    load_cell<T, 2 + ND>(in, i, es, cs, u);
    ws[i] = cell_wavespeed64<T, ND>(u, m64);
  }
}

// Maximum of n doubles as an order-preserving 64-bit key (see ordered_key),
// one atomicMax per workgroup into *dst (cleared by the caller to 0, which
// is below every key).
__global__ void __launch_bounds__(kMaxBlock) max_element_kernel(const double* __restrict__ v,
                                                                 int64_t n,
                                                                 unsigned long long* dst) {
  unsigned long long m = 0ull;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const unsigned long long k = ordered_key(v[i]);
    m = k > m ? k : m;
  }
  block_atomic_max(m, dst);
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

// One step of Parameters::update_time_delta / dt_dh / update_simulation_info
// (simulator/parameters.hpp:60-99) and the loop condition of
// SimulatorImpl::simulate (simulator/simulator_impl.hpp:235).
__global__ void step_control_kernel(StepCtl* ctl, int parity, double cfl, double dh,
                                    double sim_time, long long max_iters) {
  const bool active = ctl->t < sim_time && ctl->iters < max_iters;
  if (active) {
    const double maxws = __longlong_as_double((long long)ctl->ws_bits[parity]);
    const double dt = cfl * dh / maxws;
    ctl->dt = dt;
    ctl->dtdh = dt / dh;
    ctl->t += dt;
    ctl->iters += 1;
    ctl->ws_bits[parity ^ 1] = 0ull;  // target of this step's fused CFL epilogue
  }
  ctl->active = active ? 1 : 0;
}

// dt of a two-material step from the maximum wavespeed reduced into
// ws_bits[0] over both grids after the ghost states were written (the loop of
// simulator/multimaterial_simulator.hpp, update_time_delta in
// simulator/parameters.hpp:60-99): dt = cfl dh / max, dt/dh, run time and
// iteration count.  The maximum moves to ws_bits[1] (reported) and ws_bits[0]
// is cleared for the next step.  A NaN, infinite or non-positive maximum
// leaves dt, t and iters unchanged and clears `active`, which the sweeps read.
__global__ void mm_control_kernel(StepCtl* ctl, double cfl, double dh) {
  const unsigned long long bits = ctl->ws_bits[0];
  const double maxws = __longlong_as_double((long long)bits);
  ctl->ws_bits[1] = bits;
  ctl->ws_bits[0] = 0ull;
  const bool ok = maxws > 0.0 && !isnan(maxws) && !isinf(maxws);
  if (ok) {
    const double dt = cfl * dh / maxws;
    ctl->dt = dt;
    ctl->dtdh = dt / dh;
    ctl->t += dt;
    ctl->iters += 1;
  }
  ctl->active = ok ? 1 : 0;
}

// ---------------------------------------------------------------------------
// persistent 1D step loop
// ---------------------------------------------------------------------------

// Grid-wide barrier of a cooperative launch: every workgroup adds one to the
// counter and waits until it reaches `target`.  The wait gives up after
// spin_limit polls, sets *err and returns false (every workgroup then stops:
// the counter cannot reach the target of any later barrier either), so the
// kernel ends even if the barrier never completes.  Release/acquire at agent
// scope make the stores before the barrier visible to every workgroup after it.
__device__ __forceinline__ bool grid_barrier(unsigned int* bar, unsigned int target,
                                             unsigned int spin_limit, int* err) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    // the release fence publishes this workgroup's stores and atomics; the
    // polls are relaxed agent-scope loads (they see the counter at the
    // device's coherence point) and one acquire fence follows the last, so
    // no poll invalidates the caches; the error flag is read every 64 polls
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned int spins = 0;
    int good = 1;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      ++spins;
      if (spins > spin_limit ||
          ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

__device__ __forceinline__ unsigned long long load_u64_agent(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The whole step loop of a 1D grid in one cooperative launch: steps k0 ..
// k0 + nsteps - 1 while run_time < sim_time and iters < max_iters (the
// reference's loop condition, simulator/simulator_impl.hpp:235-252).  Every
// workgroup reads the same control words, so every workgroup computes the
// same dt = cfl dh / max (update_time_delta, simulator/parameters.hpp:60-99)
// and stops at the same step.  Workgroup b sweeps segment b of the line
// (sweep_x_tile, one segment per workgroup), reducing the new wavespeeds into
// the next slot; a grid barrier ends each step.  Step k reads a_ab (state ->
// work) for even k and a_ba for odd k.
template <typename T, int FLUX>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(SweepWaves<T>::value)))
persist1d_kernel(SweepArgs<T> a_ab, SweepArgs<T> a_ba, PersistCtl* pc, double cfl, double dh,
                 double sim_time, long long max_iters, long long k0, long long nsteps,
                 unsigned int spin_limit) {
  const int64_t col = (int64_t)blockIdx.x * (blockDim.x - 2) - 1 + threadIdx.x;
  const int64_t ncell = a_ab.nx;
  for (long long k = k0; k < k0 + nsteps; ++k) {
    // This workgroup's cells of the step's input, read while the control
    // words are on their way (after the barrier they come from device
    // memory): the tile's own loads of them then hit the caches.
    const SweepArgs<T>& ain = (k & 1) ? a_ba : a_ab;
    const int64_t pc_col = col < 0 ? 0 : (col >= ncell ? ncell - 1 : col);
    T warm[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) warm[q] = ain.in[pc_col * ain.es + q * ain.cs];
    const double maxws = __longlong_as_double((long long)load_u64_agent(&pc->ws[k % 3]));
    const double t = __longlong_as_double((long long)load_u64_agent(&pc->t[k & 1]));
    const long long it =
        (long long)load_u64_agent(reinterpret_cast<unsigned long long*>(&pc->iters[k & 1]));
    if (!(t < sim_time && it < max_iters)) return;  // the same for every workgroup
    const double dt = cfl * dh / maxws;
    const double dtdh = dt / dh;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      pc->t[(k + 1) & 1] = (unsigned long long)__double_as_longlong(t + dt);
      pc->iters[(k + 1) & 1] = it + 1;
      pc->dt = dt;
      pc->ws[(k + 2) % 3] = 0ull;
    }
    SweepArgs<T> a = (k & 1) ? a_ba : a_ab;
    a.ctl = nullptr;
    a.dtdh_host = (T)dtdh;
    a.ws_out = &pc->ws[(k + 1) % 3];
    asm volatile("" ::"v"(warm[0]), "v"(warm[1]), "v"(warm[2]));
    sweep_x_tile<T, 1, FLUX>(a, blockIdx.x, 0);
    if (!grid_barrier(&pc->bar, (unsigned int)((k - k0 + 1) * gridDim.x), spin_limit, &pc->err))
      return;
  }
}

// Slot 0 holds the initial maximum (*ws0), run time and count start at 0.
__global__ void persist_init_kernel(PersistCtl* pc, const unsigned long long* ws0) {
  pc->ws[0] = *ws0;
  pc->ws[1] = 0ull;
  pc->ws[2] = 0ull;
  pc->t[0] = 0ull;
  pc->t[1] = 0ull;
  pc->iters[0] = 0;
  pc->iters[1] = 0;
  pc->dt = 0.0;
  pc->bar = 0u;
  pc->err = 0;
}

__global__ void persist_bar_reset_kernel(PersistCtl* pc) { pc->bar = 0u; }

__global__ void init_control_kernel(StepCtl* ctl) {
  ctl->ws_bits[0] = 0ull;
  ctl->ws_bits[1] = 0ull;
  ctl->dt = 0.0;
  ctl->dtdh = 0.0;
  ctl->t = 0.0;
  ctl->iters = 0;
  ctl->active = 1;
  ctl->pad = 0;
}

// ---------------------------------------------------------------------------
// FP64 conservation sums (per component), used beside the FP32 bulk path
// ---------------------------------------------------------------------------

template <typename T, int S>
__global__ void __launch_bounds__(kMaxBlock) component_sums_kernel(const T* __restrict__ in,
                                                                   int64_t es, int64_t cs,
                                                                   int64_t ncells,
                                                                   double* partial) {
  __shared__ double red[S][kMaxBlock / 64];
  double acc[S];
#pragma unroll
  for (int k = 0; k < S; ++k) acc[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncells;
       i += (int64_t)gridDim.x * blockDim.x) {
    T u[S];
    load_cell<T, S>(in, i, es, cs, u);
#pragma unroll
    for (int k = 0; k < S; ++k) acc[k] += (double)u[k];
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[k] = v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < S; ++k) red[k][wave] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      double v = 0.0;
      for (int i = 0; i < nw; ++i) v += red[k][i];
      partial[(int64_t)blockIdx.x * S + k] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// explicit instantiations (launched from fl_abi.cpp)
// ---------------------------------------------------------------------------

#define FL_INST_SWEEPS2_V(T, FLUX, FAST, VLC)                                               \
  template __global__ void sweep_x2_kernel<T, 1, FLUX, FAST, VLC>(SweepArgs<T>, SweepArgs<T>); \
  template __global__ void sweep_x2_kernel<T, 2, FLUX, FAST, VLC>(SweepArgs<T>, SweepArgs<T>); \
  template __global__ void sweep_y2_kernel<T, FLUX, FAST, VLC>(SweepArgs<T>, SweepArgs<T>);
#define FL_INST_SWEEPS2(T, FLUX) FL_INST_SWEEPS2_V(T, FLUX, false, false)

FL_INST_SWEEPS2(double, kFluxHllc)
FL_INST_SWEEPS2(double, kFluxForce)
FL_INST_SWEEPS2(double, kFluxLaxFriedrichs)
FL_INST_SWEEPS2(double, kFluxRichtmyer)
FL_INST_SWEEPS2(float, kFluxHllc)
FL_INST_SWEEPS2(float, kFluxForce)
FL_INST_SWEEPS2(float, kFluxLaxFriedrichs)
FL_INST_SWEEPS2(float, kFluxRichtmyer)
FL_INST_SWEEPS2_V(double, kFluxHllc, false, true)
FL_INST_SWEEPS2_V(double, kFluxForce, false, true)
FL_INST_SWEEPS2_V(float, kFluxHllc, false, true)
FL_INST_SWEEPS2_V(float, kFluxForce, false, true)
// the fast FP64 numerics
FL_INST_SWEEPS2_V(double, kFluxHllc, true, false)
FL_INST_SWEEPS2_V(double, kFluxForce, true, false)
FL_INST_SWEEPS2_V(double, kFluxLaxFriedrichs, true, false)
FL_INST_SWEEPS2_V(double, kFluxRichtmyer, true, false)
FL_INST_SWEEPS2_V(double, kFluxHllc, true, true)
FL_INST_SWEEPS2_V(double, kFluxForce, true, true)

#define FL_INST_MAXWS2_F(T, FAST)                                                          \
  template __global__ void max_wavespeed2_kernel<T, 1, FAST>(                                 \
      const T*, const T*, int64_t, int64_t, int64_t, int64_t, int64_t, Material<double>,      \
      Material<double>, unsigned long long*);                                                \
  template __global__ void max_wavespeed2_kernel<T, 2, FAST>(                                 \
      const T*, const T*, int64_t, int64_t, int64_t, int64_t, int64_t, Material<double>,      \
      Material<double>, unsigned long long*);
#define FL_INST_MAXWS2(T) FL_INST_MAXWS2_F(T, false)

FL_INST_MAXWS2(double)
FL_INST_MAXWS2(float)
FL_INST_MAXWS2_F(double, true)

#define FL_INST_PERSIST(T)                                                                   \
  template __global__ void persist1d_kernel<T, kFluxHllc>(SweepArgs<T>, SweepArgs<T>,        \
                                                          PersistCtl*, double, double, double, \
                                                          long long, long long, long long,     \
                                                          unsigned int);                       \
  template __global__ void persist1d_kernel<T, kFluxForce>(SweepArgs<T>, SweepArgs<T>,       \
                                                           PersistCtl*, double, double, double, \
                                                           long long, long long, long long,     \
                                                           unsigned int);                       \
  template __global__ void persist1d_kernel<T, kFluxLaxFriedrichs>(                          \
      SweepArgs<T>, SweepArgs<T>, PersistCtl*, double, double, double, long long, long long,  \
      long long, unsigned int);                                                                \
  template __global__ void persist1d_kernel<T, kFluxRichtmyer>(                              \
      SweepArgs<T>, SweepArgs<T>, PersistCtl*, double, double, double, long long, long long,  \
      long long, unsigned int);
FL_INST_PERSIST(double)
FL_INST_PERSIST(float)

#define FL_SWEEPS_V(T, FLUX, FAST, VLC)                                          \
  template __global__ void sweep_x_kernel<T, 1, FLUX, FAST, VLC>(SweepArgs<T>);  \
  template __global__ void sweep_x_kernel<T, 2, FLUX, FAST, VLC>(SweepArgs<T>);  \
  template __global__ void sweep_y_kernel<T, FLUX, FAST, VLC>(SweepArgs<T>);
#define FL_SWEEPS(T, FLUX, FAST) FL_SWEEPS_V(T, FLUX, FAST, false)

#define FL_INST_SWEEPS(T)                                                      \
  FL_SWEEPS(T, kFluxHllc, false)                                                          \
  FL_SWEEPS(T, kFluxForce, false)                                                         \
  FL_SWEEPS(T, kFluxLaxFriedrichs, false)                                                 \
  FL_SWEEPS(T, kFluxRichtmyer, false)                                                     \
  template __global__ void max_wavespeed_kernel<T, 1>(const T*, int64_t, int64_t, int64_t, \
                                                      Material<double>, unsigned long long*); \
  template __global__ void max_wavespeed_kernel<T, 2>(const T*, int64_t, int64_t, int64_t, \
                                                      Material<double>, unsigned long long*); \
  template __global__ void wavespeeds_kernel<T, 1>(const T*, int64_t, int64_t, int64_t,        \
                                                   Material<double>, double*);                   \
  template __global__ void wavespeeds_kernel<T, 2>(const T*, int64_t, int64_t, int64_t,        \
                                                   Material<double>, double*);                   \
  template __global__ void component_sums_kernel<T, 3>(const T*, int64_t, int64_t, int64_t, double*); \
  template __global__ void component_sums_kernel<T, 4>(const T*, int64_t, int64_t, int64_t, double*);

FL_INST_SWEEPS(double)
FL_INST_SWEEPS(float)
// the fast FP64 numerics (FP32 sweeps always use them)
FL_SWEEPS(double, kFluxHllc, true)
FL_SWEEPS(double, kFluxForce, true)
FL_SWEEPS(double, kFluxLaxFriedrichs, true)
FL_SWEEPS(double, kFluxRichtmyer, true)
// van Leer, conservative form: HLLC and FORCE
FL_SWEEPS_V(double, kFluxHllc, false, true)
FL_SWEEPS_V(double, kFluxForce, false, true)
FL_SWEEPS_V(double, kFluxHllc, true, true)
FL_SWEEPS_V(double, kFluxForce, true, true)
FL_SWEEPS_V(float, kFluxHllc, false, true)
FL_SWEEPS_V(float, kFluxForce, false, true)

}  // namespace fl
