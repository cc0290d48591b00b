// fl_levelset.hpp -- argument blocks and declarations of the level-set and
// ghost-fluid kernels (fl_levelset.hip), shared with the C ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace fl {
namespace ls {

struct Grid {
  int64_t nx, ny;  // ny = 1 in 1D
  int nd;
  int pad;
};

// A conservative state grid: element stride and component stride in
// elements (AoS: es = S, cs = 1; SoA: es = 1, cs = cells).  f32: the grid
// holds floats (the FP32 bulk path); the level-set kernels compute in FP64
// and convert on load and store.
struct StateView {
  double* data;
  int64_t es, cs;
  int f32;
  int pad;
};

// Control words of the eikonal Jacobi loop: flags[k % 3] records a change in
// sweep k, done ends the loop, iters counts the sweeps run.
struct EikCtl {
  int flags[3];
  int done;
  long long iters;
  int ntiles;  // entries of the job's active-tile list
  int pad;
};

// One eikonal reinitialisation: the level set in, its two Jacobi distance
// buffers, per-cell info bytes, the tile flags, its loop control and the
// result.  Launches carry up to two jobs; blockIdx.z selects one.
struct EikJob {
  const double* phi;
  double* d[2];
  uint8_t* info;
  uint8_t* src;   // per tile: holds a source
  int* list;      // the tiles a sweep visits, in no particular order
  EikCtl* ctl;
  double* out;
  uint8_t* band_mask;  // nullable
};
struct EikJobs {
  EikJob j[2];
};

struct Eos {
  double gamma, p_inf;
  int stiff;  // either stiffened gas (energy and sound speed with p_inf)
  int corr;   // corrected stiffened gas: pressure minus gamma p_inf
};

// Materials of an n-material step (n <= kMaxMat): level sets, states and
// equations of state, passed by value to the kernels that read all of them.
constexpr int kMaxMat = 8;
struct MatSet {
  int nmat;
  int pad;
  const double* phi[kMaxMat];
  StateView u[kMaxMat];
  Eos e[kMaxMat];
};
struct PhiSet {
  int nmat;
  int pad;
  double* phi[kMaxMat];
};

__global__ void masks_kernel(const double* phi, Grid g, double dh, uint8_t* masks);
template <int ND>
__global__ void gfm_star_n_kernel(MatSet ms, int m, Grid g, double dh, double* star,
                                  uint8_t* has_star, int* count_reset);
template <int ND>
__global__ void velocity_n_kernel(MatSet ms, int target, Grid g, double dh, double* V);
__global__ void project_n_kernel(PhiSet ps, int64_t count);
template <int ND>
__global__ void velocity_kernel(const double* phi0, const double* phi1, StateView u0,
                                StateView u1, int nmat, Grid g, double dh, int target, double* V);
template <int ND>
__global__ void velocity2_kernel(const double* phi0, const double* phi1, StateView u0,
                                 StateView u1, Grid g, double dh, double* V0, double* V1);
__global__ void project2_kernel(double* phi0, double* phi1, int64_t count);
// One level set's RK3 stage: phi_n, the stage source, the speed and the output.
struct Rk3Job {
  const double* phi_n;
  const double* src;
  const double* V;
  double* out;
};
template <int ND, bool FAST>
__global__ void rk3_stage2_kernel(int stage, Rk3Job j0, Rk3Job j1, Grid g, const double* dt_dev,
                                  double h1);
template <int ND>
__global__ void rk3_stage_kernel(int stage, const double* phi_n, const double* src,
                                 const double* V, Grid g, double dt, const double* dt_dev,
                                 double h1, double* out);
__global__ void eik_init_kernel(EikJobs jobs, Grid g);
// Eikonal tiles: kEikTileW cells x kEikTileH rows (one workgroup each).
constexpr int kEikTileW = 64;
constexpr int kEikTileH = 4;
__global__ void eik_tile_src_kernel(EikJobs jobs, Grid g, int64_t trows);
__global__ void eik_tile_act_kernel(EikJobs jobs, int64_t trows, int64_t segs, int tr_r,
                                    int segs_r);
__global__ void eik_ctl_init_kernel(EikJobs jobs, int njobs);
// Workgroups per job of an eikonal sweep; they stride over the active tiles.
constexpr int kEikSweepGroups = 512;
__global__ void eik_sweep_kernel(EikJobs jobs, Grid g, int64_t segs, double dh, double limit,
                                 int k);
__global__ void eik_finish_kernel(EikJobs jobs, Grid g, double limit, double clamp);
template <int ND>
__global__ void gfm_star_kernel(const double* phi_m, const double* phi_o, StateView um,
                                StateView uo, Eos em, Eos eo, Grid g, double dh, double* star,
                                uint8_t* has_star, int* count_reset);
// Both materials of a two-material step in one pass each.
template <int ND>
__global__ void gfm_star2_kernel(const double* phi0, const double* phi1, StateView u0,
                                 StateView u1, Eos e0, Eos e1, Grid g, double dh, double* star0,
                                 double* star1, uint8_t* has0, uint8_t* has1, int* count0,
                                 int* count1);
// One extrapolation job: a material's level set and state, its two Jacobi
// buffers (AoS), and its band list with the list's count.
struct ExtrapJob {
  const double* phi;
  StateView um;
  double* A;
  double* B;
  int* list;  // written by the band seed, read by the sweeps
  int* count;
};
constexpr int kExtrapBlocks = 128;  // workgroups per job of one sweep launch

template <int ND>
__global__ void gfm_band_seed_kernel(const double* phi_m, double upper, StateView um,
                                     const double* star, const uint8_t* has_star, Grid g,
                                     double* A, double* B, int* list, int* count);
template <int ND>
__global__ void gfm_band_seed2_kernel(ExtrapJob j0, ExtrapJob j1, double upper,
                                      const double* star0, const double* star1,
                                      const uint8_t* has0, const uint8_t* has1, Grid g);
__global__ void count_clear_kernel(int* count);
template <int ND>
__global__ void gfm_extrap_kernel(ExtrapJob j0, ExtrapJob j1, Grid g, double dh, int sweep,
                                  int last);
template <int ND>
__global__ void combine_kernel(const double* phi0, StateView u0, StateView u1, Grid g,
                               StateView out);

}  // namespace ls
}  // namespace fl
