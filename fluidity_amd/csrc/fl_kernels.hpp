// fl_kernels.hpp -- argument blocks shared by the gfx950 kernels and the
// host-side launch code of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fl_numerics.hpp"

namespace fl {

// Device-resident step control.  One instance drives a whole simulation so
// that the time-step loop never synchronises with the host: the control
// kernel turns the fused maximum wavespeed into dt and dt/dh exactly as
// Parameters::update_time_delta / dt_dh do (simulator/parameters.hpp:60-99)
// and advances run_time / iters (:67-71).
struct StepCtl {
  unsigned long long ws_bits[2];  // max wavespeed as bits, ping-pong by step parity
  double dt;
  double dtdh;
  double t;                       // run time
  long long iters;
  int active;                     // run_time < sim_time && iters < max_iters
  int pad;
};

// Control words of the persistent 1D step loop (persist1d_kernel): the
// fused wavespeed maxima in three slots (step k reads slot k % 3, reduces into
// (k + 1) % 3 and clears (k + 2) % 3), run time and iteration count in two
// (step k reads k % 2, writes (k + 1) % 2), the last dt, the grid-barrier
// counter of the launch and a timeout flag.
struct PersistCtl {
  unsigned long long ws[3];
  unsigned long long t[2];  // run time as bits
  long long iters[2];
  double dt;
  unsigned int bar;
  int err;
};

template <typename T>
struct SweepArgs {
  const T* in;
  T* out;
  int64_t es;          // element stride (in T units): S for AoS, 1 for SoA
  int64_t cs;          // component stride: 1 for AoS, nx*ny for SoA
  int64_t nx, ny;      // interior cells
  Material<T> mat;     // material in the sweep precision
  Material<double> mat64;  // FP64 material for the fused wavespeed
  int limiter, form;
  int bcx, bcy;
  int rows_per_block;  // row chunk marched by one workgroup
  int pad;
  T dtdh_host;         // used when ctl == nullptr
  const StepCtl* ctl;  // nullable: provides dtdh and the active flag
  unsigned long long* ws_out;  // nullable: fused max wavespeed (bit pattern)
};

// Maximum workgroup width of the sweep kernels (LDS arrays are sized for it).
constexpr int kMaxBlock = 256;
constexpr int kHalo = 3;  // extra cells each side staged in LDS for chi

// Dynamic LDS of an x-sweep tile of B lanes over states of S components of
// tsize bytes: S + 3 rows of B + 2 kHalo staged columns, 2 S rows of B.
__host__ __device__ constexpr size_t x_tile_lds_bytes(int B, int S, int tsize) {
  return (size_t)((S + 3) * (B + 2 * kHalo) + 2 * S * B) * (size_t)tsize;
}

// Dynamic LDS of a y-sweep tile of B lanes (2D states, 4 components of tsize
// bytes): 3 rows of B + 2 kHalo columns, a ring of kYRing rows of 4 x B.
constexpr int kYRing = 5;
__host__ __device__ constexpr size_t y_tile_lds_bytes(int B, int tsize) {
  return (size_t)(3 * (B + 2 * kHalo) + kYRing * 4 * B) * (size_t)tsize;
}

// Standalone wavespeed max: at most kMaxWsBlocks workgroups of kMaxBlock
// threads, each thread loading kMaxWsUnroll cells per iteration.
constexpr int kMaxWsBlocks = 512;
constexpr int kMaxWsUnroll = 4;

}  // namespace fl

// Kernels defined in fl_sweep.hip (explicitly instantiated there).
namespace fl {
template <typename T, int ND, int FLUX, bool FAST, bool VLC> __global__ void sweep_x_kernel(SweepArgs<T> a);
template <typename T, int FLUX, bool FAST, bool VLC> __global__ void sweep_y_kernel(SweepArgs<T> a);
template <typename T, int ND, int FLUX, bool FAST, bool VLC>
__global__ void sweep_x2_kernel(SweepArgs<T> a0, SweepArgs<T> a1);
template <typename T, int FLUX, bool FAST, bool VLC>
__global__ void sweep_y2_kernel(SweepArgs<T> a0, SweepArgs<T> a1);
template <typename T, int ND, bool FAST>
__global__ void max_wavespeed2_kernel(const T* in0, const T* in1, int64_t es0, int64_t cs0,
                                      int64_t es1, int64_t cs1, int64_t ncells,
                                      Material<double> m0, Material<double> m1,
                                      unsigned long long* dst);
template <typename T, int ND>
__global__ void max_wavespeed_kernel(const T* in, int64_t es, int64_t cs, int64_t ncells,
                                     Material<double> m64, unsigned long long* dst);
template <typename T, int S>
__global__ void component_sums_kernel(const T* in, int64_t es, int64_t cs, int64_t ncells,
                                      double* partial);
template <typename T, int ND>
__global__ void wavespeeds_kernel(const T* in, int64_t es, int64_t cs, int64_t ncells,
                                  Material<double> m64, double* ws);
__global__ void max_element_kernel(const double* v, int64_t n, unsigned long long* dst);
__global__ void zero_u64_kernel(unsigned long long* p);
__global__ void step_control_kernel(StepCtl* ctl, int parity, double cfl, double dh,
                                    double sim_time, long long max_iters);
__global__ void init_control_kernel(StepCtl* ctl);
__global__ void mm_control_kernel(StepCtl* ctl, double cfl, double dh);
template <typename T, int FLUX>
__global__ void persist1d_kernel(SweepArgs<T> a_ab, SweepArgs<T> a_ba, PersistCtl* pc,
                                 double cfl, double dh, double sim_time, long long max_iters,
                                 long long k0, long long nsteps, unsigned int spin_limit);
__global__ void persist_init_kernel(PersistCtl* pc, const unsigned long long* ws0);
__global__ void persist_bar_reset_kernel(PersistCtl* pc);
}  // namespace fl
