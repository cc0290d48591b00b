// fl_batch.hpp -- device tables and kernels of the batched Euler step: many
// independent grids of uneven size advanced together, each with its own
// material, dt and stop condition.
#pragma once

#include "fl_kernels.hpp"

namespace fl {
namespace batch {

// One problem.  x.in is the problem's state and x.out its work buffer; x.ctl
// and the material fields are the problem's own.  x.rows_per_block is the x
// tile height (2D), rows_y the y tile height.
template <typename T>
struct Prob {
  SweepArgs<T> x;
  StepCtl* ctl;
  int nd;
  int rows_y;
};

// A tile of one sweep phase: problem index and tile coordinates.
struct Tile {
  int prob;
  int bx;
  int by;
  int pad;
};

// Per-problem loop parameters (Parameters of the reference simulator:
// simulator/parameters.hpp).
struct Params {
  double cfl;
  double dh;
  double sim_time;
  long long max_iters;
};

// Work counters of the two sweep phases, cleared by the control kernel at
// the start of every step.
struct Counters {
  int next[2];
  int pad[2];
};

__global__ void init_kernel(StepCtl* ctls, int nprob);
__global__ void control_kernel(StepCtl* ctls, const Params* params, int nprob, Counters* cnt);
__global__ void reset_counters_kernel(Counters* cnt);
template <typename T, int FLUX, int PHASE, bool FAST, bool VLC>
__global__ void phase_kernel(const Prob<T>* probs, const Tile* tiles, int ntiles, Counters* cnt,
                             int timing, T timing_dtdh);
template <typename T>
__global__ void finish_kernel(const Prob<T>* probs, int nprob, int dir);
template <typename T>
__global__ void sums_kernel(const Prob<T>* probs, int nprob, double* sums);

}  // namespace batch
}  // namespace fl
