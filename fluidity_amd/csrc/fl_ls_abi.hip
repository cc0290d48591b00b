// fl_ls_abi.hip -- host side of the level-set and ghost-fluid entry points of
// include/fluidity_amd.h: argument checks and launches of the kernels in
// fl_levelset.hip, and the two-material step loop.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include "fl_abi_internal.hpp"

using namespace fl::abi;
namespace ls = fl::ls;

namespace {

constexpr int kBlock = 256;
constexpr int kEikBatch = 8;  // eikonal sweeps launched per convergence test

dim3 grid_for(int64_t cells) { return dim3((unsigned)((cells + kBlock - 1) / kBlock)); }

int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

int check_shape(int ndim, const int64_t* n, ls::Grid* g) {
  if (ndim != 1 && ndim != 2) return fail(FL_ERR_ARGUMENT, "ndim %d must be 1 or 2", ndim);
  if (!n) return fail(FL_ERR_ARGUMENT, "n is NULL");
  for (int d = 0; d < ndim; ++d)
    if (n[d] < 4 || n[d] > ((int64_t)1 << 30))
      return fail(FL_ERR_ARGUMENT, "n[%d] = %lld out of range [4, 2^30]", d, (long long)n[d]);
  g->nx = n[0];
  g->ny = ndim > 1 ? n[1] : 1;
  g->nd = ndim;
  g->pad = 0;
  // band lists and active-tile lists hold cell and tile indices as int
  if (g->nx * g->ny >= ((int64_t)1 << 31))
    return fail(FL_ERR_ARGUMENT, "grid of %lld cells too large (at most 2^31 - 1)",
                (long long)(g->nx * g->ny));
  return FL_OK;
}

int check_dh(double dh) {
  if (!(dh > 0.0) || !isfinite(dh)) return fail(FL_ERR_ARGUMENT, "dh %g must be > 0", dh);
  return FL_OK;
}

// A state field (FP64, or FP32 bulk storage) over grid g.
int check_state(const fl_field* f, const ls::Grid& g, const char* name) {
  int rc = check_field(f, name);
  if (rc) return rc;
  if (f->ndim != g.nd || f->n[0] != g.nx || (g.nd > 1 && f->n[1] != g.ny))
    return fail(FL_ERR_ARGUMENT, "%s does not match the level-set grid", name);
  return FL_OK;
}

ls::StateView view_of(const fl_field* f) {
  ls::StateView v;
  v.data = (double*)f->data;
  strides(f, v.es, v.cs);
  v.f32 = f->dtype == FL_F32 ? 1 : 0;
  v.pad = 0;
  return v;
}

ls::Eos eos_of(const fl_scheme* s) {
  ls::Eos e;
  e.gamma = s->gamma;
  e.p_inf = s->p_inf;
  e.stiff = (s->eos == FL_EOS_STIFF || s->eos == FL_EOS_STIFF_CORRECTED) ? 1 : 0;
  e.corr = s->eos == FL_EOS_STIFF_CORRECTED ? 1 : 0;
  return e;
}

int launch_masks(const double* phi, const ls::Grid& g, double dh, uint8_t* masks,
                 hipStream_t st) {
  hipLaunchKernelGGL(ls::masks_kernel, grid_for(g.nx * g.ny), dim3(kBlock), 0, st, phi, g, dh,
                     masks);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int launch_velocity(int nmat, const double* const* phis, const fl_field* states,
                    const ls::Grid& g, double dh, int target, double* V, hipStream_t st) {
  const ls::StateView u0 = view_of(&states[0]);
  const ls::StateView u1 = nmat > 1 ? view_of(&states[1]) : u0;
  const double* p1 = nmat > 1 ? phis[1] : phis[0];
  if (g.nd == 1)
    hipLaunchKernelGGL(ls::velocity_kernel<1>, grid_for(g.nx), dim3(kBlock), 0, st, phis[0], p1,
                       u0, u1, nmat, g, dh, target, V);
  else
    hipLaunchKernelGGL(ls::velocity_kernel<2>, grid_for(g.nx * g.ny), dim3(kBlock), 0, st,
                       phis[0], p1, u0, u1, nmat, g, dh, target, V);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Both level sets' speeds from the level sets phis and the two states.
int launch_velocity2(double* const* phis, const fl_field& s0, const fl_field& s1,
                     const ls::Grid& g, double dh, double* const* V, hipStream_t st) {
  const dim3 grid = grid_for(g.nx * g.ny);
  if (g.nd == 1)
    hipLaunchKernelGGL(ls::velocity2_kernel<1>, grid, dim3(kBlock), 0, st,
                       (const double*)phis[0], (const double*)phis[1], view_of(&s0),
                       view_of(&s1), g, dh, V[0], V[1]);
  else
    hipLaunchKernelGGL(ls::velocity2_kernel<2>, grid, dim3(kBlock), 0, st,
                       (const double*)phis[0], (const double*)phis[1], view_of(&s0),
                       view_of(&s1), g, dh, V[0], V[1]);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// RK3: phi -> s1 -> s2 -> out, s1/s2 in work.  dt_dev (nullable): dt in
// device memory, used instead of dt.
int launch_advect(const double* phi, const double* V, const ls::Grid& g, double dt, double dh,
                  double* out, double* work, hipStream_t st, const double* dt_dev = nullptr) {
  const int64_t cells = g.nx * g.ny;
  double* s1 = work;
  double* s2 = work + cells;
  const dim3 grid((unsigned)((g.nx + kBlock - 1) / kBlock),
                  (unsigned)(g.ny < 65535 ? g.ny : 65535));
  const double h1 = 1.0 / dh;
  auto k = g.nd == 1 ? ls::rk3_stage_kernel<1> : ls::rk3_stage_kernel<2>;
  hipLaunchKernelGGL(k, grid, dim3(kBlock), 0, st, 0, phi, phi, V, g, dt, dt_dev, h1, s1);
  hipLaunchKernelGGL(k, grid, dim3(kBlock), 0, st, 1, phi, (const double*)s1, V, g, dt, dt_dev,
                     h1, s2);
  hipLaunchKernelGGL(k, grid, dim3(kBlock), 0, st, 2, phi, (const double*)s2, V, g, dt, dt_dev,
                     h1, out);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// RK3 of both level sets, each stage of both in one launch: cur[m] -> out[m]
// with speeds V[m], dt in device memory; work holds 4 cells of scratch.
// fast: the fast FP64 numerics in the WENO derivatives.
int launch_advect2(double* const cur[2], double* const V[2], const ls::Grid& g, double dh,
                   double* const out[2], double* work, const double* dt_dev, bool fast,
                   hipStream_t st) {
  const int64_t cells = g.nx * g.ny;
  double* s1[2] = {work, work + cells};
  double* s2[2] = {work + 2 * cells, work + 3 * cells};
  const dim3 grid((unsigned)((g.nx + kBlock - 1) / kBlock),
                  (unsigned)(g.ny < 65535 ? g.ny : 65535), 2);
  const double h1 = 1.0 / dh;
  auto k = g.nd == 1 ? (fast ? ls::rk3_stage2_kernel<1, true> : ls::rk3_stage2_kernel<1, false>)
                     : (fast ? ls::rk3_stage2_kernel<2, true> : ls::rk3_stage2_kernel<2, false>);
  ls::Rk3Job j[3][2];
  for (int m = 0; m < 2; ++m) {
    j[0][m] = {cur[m], cur[m], V[m], s1[m]};
    j[1][m] = {cur[m], s1[m], V[m], s2[m]};
    j[2][m] = {cur[m], s2[m], V[m], out[m]};
  }
  for (int stage = 0; stage < 3; ++stage)
    hipLaunchKernelGGL(k, grid, dim3(kBlock), 0, st, stage, j[stage][0], j[stage][1], g, dt_dev,
                       h1);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int launch_project(double* phi0, double* phi1, int64_t cells, hipStream_t st) {
  hipLaunchKernelGGL(ls::project2_kernel, grid_for(cells), dim3(kBlock), 0, st, phi0, phi1,
                     cells);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int64_t eik_segs(const ls::Grid& g) { return (g.nx + ls::kEikTileW - 1) / ls::kEikTileW; }
int64_t eik_trows(const ls::Grid& g) { return (g.ny + ls::kEikTileH - 1) / ls::kEikTileH; }

// Two distance buffers, the per-cell info bytes, the per-tile source flags
// and the active-tile list.
int64_t reinit_work_bytes(const ls::Grid& g) {
  const int64_t cells = g.nx * g.ny;
  const int64_t tiles = eik_segs(g) * eik_trows(g);
  return 2 * align256(cells * 8) + align256(cells) + align256(tiles) + align256(tiles * 4);
}

// One reinitialisation job: phi in, out (and the nullable band mask) out,
// `work` of reinit_work_bytes, control words ctl.
ls::EikJob eik_job(const double* phi, const ls::Grid& g, double* out, uint8_t* band_mask,
                   void* work, ls::EikCtl* ctl) {
  const int64_t cells = g.nx * g.ny;
  const int64_t tiles = eik_segs(g) * eik_trows(g);
  char* w = (char*)work;
  ls::EikJob j;
  j.phi = phi;
  j.d[0] = (double*)w;
  j.d[1] = (double*)(w + align256(cells * 8));
  j.info = (uint8_t*)(w + 2 * align256(cells * 8));
  j.src = j.info + align256(cells);
  j.list = (int*)(j.src + align256(tiles));
  j.ctl = ctl;
  j.out = out;
  j.band_mask = band_mask;
  return j;
}

// Eikonal reinitialisation of njobs (1 or 2) level sets on one grid, every
// launch covering all jobs; iterations (nullable) receives each job's sweep
// count.  The host tests convergence after every kEikBatch sweeps.
int run_reinit(fl_context* c, const ls::EikJobs& jobs, int njobs, const ls::Grid& g, double dh,
               int band, int64_t* iterations, hipStream_t st) {
  const int64_t cells = g.nx * g.ny;
  const int64_t segs = eik_segs(g), trows = eik_trows(g), tiles = segs * trows;
  const double limit = band * dh;
  const double clamp = (band + 1) * dh;
  dim3 grid = grid_for(cells);
  grid.z = njobs;
  const dim3 tgrid((unsigned)segs, (unsigned)(trows < 65535 ? trows : 65535), njobs);
  const dim3 tblock(ls::kEikTileW * ls::kEikTileH);
  const dim3 sgrid((unsigned)(tiles < ls::kEikSweepGroups ? tiles : ls::kEikSweepGroups), 1,
                   njobs);
  dim3 agrid = grid_for(tiles);
  agrid.z = njobs;
  // cells within which a source can pull a cell into the band, in tiles
  const int reach = (int)ceil(band * sqrt(2.0)) + 2;
  const int segs_r = (reach + ls::kEikTileW - 1) / ls::kEikTileW;
  const int tr_r = g.nd > 1 ? (reach + ls::kEikTileH - 1) / ls::kEikTileH : 0;
  hipLaunchKernelGGL(ls::eik_init_kernel, grid, dim3(kBlock), 0, st, jobs, g);
  hipLaunchKernelGGL(ls::eik_tile_src_kernel, tgrid, tblock, 0, st, jobs, g, trows);
  hipLaunchKernelGGL(ls::eik_ctl_init_kernel, dim3(1), dim3(64), 0, st, jobs, njobs);
  hipLaunchKernelGGL(ls::eik_tile_act_kernel, agrid, dim3(kBlock), 0, st, jobs, trows, segs,
                     tr_r, segs_r);
  FL_HIP(hipGetLastError());
  // Every sweep strictly lowers some cell of an unconverged job, so the loop
  // ends; the bound only guards the host against a device that never
  // reports convergence.
  // The first batch is as long as the context's previous solve took (a
  // level set reinitialised every few steps needs about the same number of
  // sweeps each time), so that one test usually suffices; sweeps past
  // convergence return at once.
  const int64_t max_sweeps = cells + 2;
  int64_t launched = 0;
  int64_t batch = c->eik_sweeps > kEikBatch ? c->eik_sweeps + 1 : kEikBatch;
  if (batch > max_sweeps) batch = max_sweeps;
  for (;;) {
    for (int64_t b = 0; b < batch; ++b, ++launched)
      hipLaunchKernelGGL(ls::eik_sweep_kernel, sgrid, tblock, 0, st, jobs, g, segs, dh, limit,
                         (int)launched);
    batch = kEikBatch;
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(c->host_eik, c->eik, njobs * sizeof(ls::EikCtl),
                          hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    bool converged = true;
    for (int j = 0; j < njobs; ++j) {
      const ls::EikCtl& h = c->host_eik[j];
      converged = converged && (h.done || h.flags[(h.iters - 1) % 3] == 0);
    }
    if (converged) break;
    if (launched >= max_sweeps)
      return fail(FL_ERR_NUMERIC, "eikonal solve did not converge in %lld sweeps",
                  (long long)launched);
  }
  int64_t took = 0;
  for (int j = 0; j < njobs; ++j)
    took = c->host_eik[j].iters > took ? c->host_eik[j].iters : took;
  c->eik_sweeps = took;
  hipLaunchKernelGGL(ls::eik_finish_kernel, grid, dim3(kBlock), 0, st, jobs, g, limit, clamp);
  FL_HIP(hipGetLastError());
  if (iterations)
    for (int j = 0; j < njobs; ++j) iterations[j] = c->host_eik[j].iters;
  return FL_OK;
}

// count_reset (nullable): a band-list count to zero for the extrapolation
// that follows.
int launch_star(const double* phi_m, const double* phi_o, const fl_field* sm,
                const fl_field* so, const fl_scheme* em, const fl_scheme* eo, const ls::Grid& g,
                double dh, double* star, uint8_t* has, hipStream_t st,
                int* count_reset = nullptr) {
  const dim3 grid = grid_for(g.nx * g.ny);
  if (g.nd == 1)
    hipLaunchKernelGGL(ls::gfm_star_kernel<1>, grid, dim3(kBlock), 0, st, phi_m, phi_o,
                       view_of(sm), view_of(so), eos_of(em), eos_of(eo), g, dh, star, has,
                       count_reset);
  else
    hipLaunchKernelGGL(ls::gfm_star_kernel<2>, grid, dim3(kBlock), 0, st, phi_m, phi_o,
                       view_of(sm), view_of(so), eos_of(em), eos_of(eo), g, dh, star, has,
                       count_reset);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Extrapolation scratch: two AoS Jacobi buffers, the band list and its count.
int64_t extrap_work_bytes(const ls::Grid& g) {
  const int64_t cells = g.nx * g.ny;
  return 2 * align256(cells * (2 + g.nd) * 8) + align256(cells * 4) + 256;
}

// The pieces of one material's extrapolation scratch (extrap_work_bytes).
ls::ExtrapJob extrap_job(const double* phi_m, ls::StateView um, const ls::Grid& g, void* work) {
  const int64_t cells = g.nx * g.ny;
  char* w = (char*)work;
  const int64_t bb = align256(cells * (2 + g.nd) * 8);
  ls::ExtrapJob j;
  j.phi = phi_m;
  j.um = um;
  j.A = (double*)w;
  j.B = (double*)(w + bb);
  j.list = (int*)(w + 2 * bb);
  j.count = (int*)(w + 2 * bb + align256(cells * 4));
  return j;
}

double extrap_upper(double dh) { return 6.0 * dh * (1.0 / sqrt(2.0)); }

// The two materials' star-state planes and star flags of a step.
struct MmGhost {
  double* star[2];
  uint8_t* has[2];
};

// Seeds the Jacobi buffers and lists the band of one material (the job's
// count must be zero).
void band_seed(const ls::ExtrapJob& j, const double* star, const uint8_t* has,
               const ls::Grid& g, double dh, hipStream_t st) {
  const dim3 grid = grid_for(g.nx * g.ny);
  if (g.nd == 1)
    hipLaunchKernelGGL(ls::gfm_band_seed_kernel<1>, grid, dim3(kBlock), 0, st, j.phi,
                       extrap_upper(dh), j.um, star, has, g, j.A, j.B, j.list, j.count);
  else
    hipLaunchKernelGGL(ls::gfm_band_seed_kernel<2>, grid, dim3(kBlock), 0, st, j.phi,
                       extrap_upper(dh), j.um, star, has, g, j.A, j.B, j.list, j.count);
}

// Star states of both materials (each from the same level sets and states)
// and zeroed band-list counts of both jobs.
void launch_star2(double* const cur[2], const fl_material* mats, const ls::Grid& g, double dh,
                  const MmGhost& w, const ls::ExtrapJob jobs[2], hipStream_t st) {
  const dim3 grid = grid_for(g.nx * g.ny);
  const ls::StateView u0 = view_of(&mats[0].state), u1 = view_of(&mats[1].state);
  const ls::Eos e0 = eos_of(&mats[0].scheme), e1 = eos_of(&mats[1].scheme);
  if (g.nd == 1)
    hipLaunchKernelGGL(ls::gfm_star2_kernel<1>, grid, dim3(kBlock), 0, st,
                       (const double*)cur[0], (const double*)cur[1], u0, u1, e0, e1, g, dh,
                       w.star[0], w.star[1], w.has[0], w.has[1], jobs[0].count, jobs[1].count);
  else
    hipLaunchKernelGGL(ls::gfm_star2_kernel<2>, grid, dim3(kBlock), 0, st,
                       (const double*)cur[0], (const double*)cur[1], u0, u1, e0, e1, g, dh,
                       w.star[0], w.star[1], w.has[0], w.has[1], jobs[0].count, jobs[1].count);
}

// Band seeds of both materials (the jobs' counts must be zero).
void band_seed2(const ls::ExtrapJob jobs[2], const MmGhost& w, const ls::Grid& g, double dh,
                hipStream_t st) {
  const dim3 grid = grid_for(g.nx * g.ny);
  if (g.nd == 1)
    hipLaunchKernelGGL(ls::gfm_band_seed2_kernel<1>, grid, dim3(kBlock), 0, st, jobs[0],
                       jobs[1], extrap_upper(dh), (const double*)w.star[0],
                       (const double*)w.star[1], (const uint8_t*)w.has[0],
                       (const uint8_t*)w.has[1], g);
  else
    hipLaunchKernelGGL(ls::gfm_band_seed2_kernel<2>, grid, dim3(kBlock), 0, st, jobs[0],
                       jobs[1], extrap_upper(dh), (const double*)w.star[0],
                       (const double*)w.star[1], (const uint8_t*)w.has[0],
                       (const uint8_t*)w.has[1], g);
}

// The sweeps and write-back of `njobs` (1 or 2) materials: one launch per
// sweep covers every job; the last sweep writes the band into the states.
void extrap_sweeps(const ls::ExtrapJob& j0, const ls::ExtrapJob& j1, int njobs,
                   const ls::Grid& g, double dh, int sweeps, hipStream_t st) {
  const dim3 grid(ls::kExtrapBlocks * njobs);
  const int launches = sweeps > 0 ? sweeps : 1;
  for (int k = 0; k < launches; ++k) {
    const int sweep = sweeps > 0 ? k : -1;
    const int last = k == launches - 1;
    if (g.nd == 1)
      hipLaunchKernelGGL(ls::gfm_extrap_kernel<1>, grid, dim3(kBlock), 0, st, j0, j1, g, dh,
                         sweep, last);
    else
      hipLaunchKernelGGL(ls::gfm_extrap_kernel<2>, grid, dim3(kBlock), 0, st, j0, j1, g, dh,
                         sweep, last);
  }
}

int launch_extrapolate(const double* phi_m, const fl_field* sm, const double* star,
                       const uint8_t* has, const ls::Grid& g, double dh, int sweeps,
                       void* work, hipStream_t st) {
  const ls::ExtrapJob j = extrap_job(phi_m, view_of(sm), g, work);
  hipLaunchKernelGGL(ls::count_clear_kernel, dim3(1), dim3(1), 0, st, j.count);
  band_seed(j, star, has, g, dh, st);
  extrap_sweeps(j, j, 1, g, dh, sweeps, st);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int launch_combine(const double* phi0, const fl_field* s0, const fl_field* s1,
                   const fl_field* out, const ls::Grid& g, hipStream_t st) {
  const dim3 grid = grid_for(g.nx * g.ny);
  if (g.nd == 1)
    hipLaunchKernelGGL(ls::combine_kernel<1>, grid, dim3(kBlock), 0, st, phi0, view_of(s0),
                       view_of(s1), g, view_of(out));
  else
    hipLaunchKernelGGL(ls::combine_kernel<2>, grid, dim3(kBlock), 0, st, phi0, view_of(s0),
                       view_of(s1), g, view_of(out));
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Scratch layout of fl_mm_simulate for nmat materials: per material a star
// plane, extrapolation scratch, speed, advection buffer, reinitialisation
// scratch and star flags, and the RK3 stages of two level sets.
struct MmWork {
  double* star[ls::kMaxMat];
  void* extrap[ls::kMaxMat];  // extrap_work_bytes each
  double* V[ls::kMaxMat];
  double* adv;                // 4 * cells (RK3 stages of two level sets)
  double* tphi[ls::kMaxMat];
  void* reinit[ls::kMaxMat];  // reinit_work_bytes each
  uint8_t* has[ls::kMaxMat];
  int64_t bytes;
};

MmWork mm_layout(void* base, const ls::Grid& g, int nmat) {
  const int64_t cells = g.nx * g.ny;
  const int S = 2 + g.nd;
  char* p = (char*)base;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* r = p ? p + off : nullptr;
    off += align256(bytes);
    return r;
  };
  MmWork w;
  memset(&w, 0, sizeof(w));
  for (int m = 0; m < nmat; ++m) w.star[m] = (double*)take(cells * S * 8);
  for (int m = 0; m < nmat; ++m) w.extrap[m] = take(extrap_work_bytes(g));
  for (int m = 0; m < nmat; ++m) w.V[m] = (double*)take(cells * 8);
  w.adv = (double*)take(4 * cells * 8);
  for (int m = 0; m < nmat; ++m) w.tphi[m] = (double*)take(cells * 8);
  for (int m = 0; m < nmat; ++m) w.reinit[m] = take(reinit_work_bytes(g));
  for (int m = 0; m < nmat; ++m) w.has[m] = (uint8_t*)take(cells);
  w.bytes = off;
  return w;
}

bool aligned(const void* p) { return ((uintptr_t)p & 255) == 0; }

// Materials' level sets (cur) and states as a MatSet for the n-material kernels.
ls::MatSet mat_set(double* const* cur, const fl_material* mats, int nmat) {
  ls::MatSet ms;
  memset(&ms, 0, sizeof(ms));
  ms.nmat = nmat;
  for (int m = 0; m < nmat; ++m) {
    ms.phi[m] = cur[m];
    ms.u[m] = view_of(&mats[m].state);
    ms.e[m] = eos_of(&mats[m].scheme);
  }
  return ms;
}

// fl_mm_simulate's loop for nmat > 2 materials (the two-material loop's
// order, decisions L7-L9): star states of every material from the same
// data, each material's extrapolation (two per launch), dt on the device
// from the maximum over every grid, the split solves (two grids per launch
// when their flux methods match), every level set's speed, advection (two
// per launch), the n-level-set projection and reinitialisation (two per
// solve).
int mm_simulate_n(fl_context* c, fl_material* mats, int nmat, const fl_sim_params* p,
                  const ls::Grid& g, int band, int sweeps, int every, const MmWork& w,
                  fl_sim_result* result, hipStream_t st) {
  const int64_t cells = g.nx * g.ny;
  const int S = 2 + g.nd;
  const int64_t esize = mats[0].state.dtype == FL_F64 ? 8 : 4;
  const double dh = p->dh;
  const bool fast = mats[0].scheme.numerics == FL_NUMERICS_FAST;  // one for every material
  double t = 0.0, dt = 0.0, maxws = 0.0;
  int64_t it = 0;
  int rc;
  double* cur[ls::kMaxMat];
  double* nxt[ls::kMaxMat];
  for (int m = 0; m < nmat; ++m) {
    cur[m] = mats[m].phi;
    nxt[m] = w.tphi[m];
  }
  auto settle = [&]() -> int {
    for (int m = 0; m < nmat; ++m)
      if (cur[m] != mats[m].phi)
        FL_HIP(hipMemcpyAsync(mats[m].phi, cur[m], (size_t)(cells * 8),
                              hipMemcpyDeviceToDevice, st));
    return FL_OK;
  };
  const dim3 grid = grid_for(cells);
  fl::StepCtl* ctl = c->ctl;
  hipLaunchKernelGGL(fl::init_control_kernel, dim3(1), dim3(1), 0, st, ctl);
  FL_HIP(hipGetLastError());
  while (t < p->sim_time && it < p->max_iters) {
    ls::ExtrapJob jobs[ls::kMaxMat];
    for (int m = 0; m < nmat; ++m)
      jobs[m] = extrap_job(cur[m], view_of(&mats[m].state), g, w.extrap[m]);
    const ls::MatSet ms = mat_set(cur, mats, nmat);
    for (int m = 0; m < nmat; ++m) {
      if (g.nd == 1)
        hipLaunchKernelGGL(ls::gfm_star_n_kernel<1>, grid, dim3(kBlock), 0, st, ms, m, g, dh,
                           w.star[m], w.has[m], jobs[m].count);
      else
        hipLaunchKernelGGL(ls::gfm_star_n_kernel<2>, grid, dim3(kBlock), 0, st, ms, m, g, dh,
                           w.star[m], w.has[m], jobs[m].count);
    }
    for (int m = 0; m < nmat; m += 2) {
      if (m + 1 < nmat) {
        const MmGhost gw = {{w.star[m], w.star[m + 1]}, {w.has[m], w.has[m + 1]}};
        band_seed2(jobs + m, gw, g, dh, st);
        extrap_sweeps(jobs[m], jobs[m + 1], 2, g, dh, sweeps, st);
      } else {
        band_seed(jobs[m], w.star[m], w.has[m], g, dh, st);
        extrap_sweeps(jobs[m], jobs[m], 1, g, dh, sweeps, st);
      }
    }
    FL_HIP(hipGetLastError());
    for (int m = 0; m < nmat; ++m)
      if ((rc = launch_max_ws(&mats[m].state, &mats[m].scheme, &ctl->ws_bits[0], st))) return rc;
    hipLaunchKernelGGL(fl::mm_control_kernel, dim3(1), dim3(1), 0, st, ctl, p->cfl, dh);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(c->host_ctl, ctl, sizeof(fl::StepCtl), hipMemcpyDeviceToHost, st));
    FL_HIP(hipEventRecord(c->step_ev, st));
    for (int m = 0; m < nmat; m += 2) {
      if (m + 1 < nmat) {
        const fl_field* const u[2] = {&mats[m].state, &mats[m + 1].state};
        const fl_field* const v[2] = {&mats[m].work, &mats[m + 1].work};
        const fl_scheme* const sc[2] = {&mats[m].scheme, &mats[m + 1].scheme};
        if ((rc = launch_sweep2(c, u, v, 0, sc, 0.0, st, ctl))) return rc;
        if (g.nd > 1 && (rc = launch_sweep2(c, v, u, 1, sc, 0.0, st, ctl))) return rc;
      } else {
        if ((rc = launch_sweep(c, &mats[m].state, &mats[m].work, 0, &mats[m].scheme, 0.0, ctl,
                               nullptr, st)))
          return rc;
        if (g.nd > 1 && (rc = launch_sweep(c, &mats[m].work, &mats[m].state, 1, &mats[m].scheme,
                                           0.0, ctl, nullptr, st)))
          return rc;
      }
    }
    FL_HIP(hipEventSynchronize(c->step_ev));
    const fl::StepCtl hc = *c->host_ctl;
    maxws = bits_to_double(hc.ws_bits[1]);
    if (!hc.active) {
      if (result) {
        result->run_time = t;
        result->iters = it;
        result->last_dt = dt;
        result->max_wavespeed = maxws;
      }
      if ((rc = settle())) return rc;
      FL_HIP(hipStreamSynchronize(st));
      return fail(FL_ERR_NUMERIC, "max wavespeed %g at step %lld (NaN or vacuum)", maxws,
                  (long long)it);
    }
    if (g.nd == 1)
      for (int m = 0; m < nmat; ++m)
        FL_HIP(hipMemcpyAsync(mats[m].state.data, mats[m].work.data, (size_t)(cells * S * esize),
                              hipMemcpyDeviceToDevice, st));
    const ls::MatSet ms2 = mat_set(cur, mats, nmat);
    for (int m = 0; m < nmat; ++m) {
      if (g.nd == 1)
        hipLaunchKernelGGL(ls::velocity_n_kernel<1>, grid, dim3(kBlock), 0, st, ms2, m, g, dh,
                           w.V[m]);
      else
        hipLaunchKernelGGL(ls::velocity_n_kernel<2>, grid, dim3(kBlock), 0, st, ms2, m, g, dh,
                           w.V[m]);
    }
    for (int m = 0; m < nmat; m += 2) {
      if (m + 1 < nmat) {
        double* const c2[2] = {cur[m], cur[m + 1]};
        double* const v2[2] = {w.V[m], w.V[m + 1]};
        double* const o2[2] = {nxt[m], nxt[m + 1]};
        if ((rc = launch_advect2(c2, v2, g, dh, o2, w.adv, &ctl->dt, fast, st))) return rc;
      } else if ((rc = launch_advect(cur[m], w.V[m], g, 0.0, dh, nxt[m], w.adv, st, &ctl->dt))) {
        return rc;
      }
    }
    {
      ls::PhiSet ps;
      memset(&ps, 0, sizeof(ps));
      ps.nmat = nmat;
      for (int m = 0; m < nmat; ++m) ps.phi[m] = nxt[m];
      hipLaunchKernelGGL(ls::project_n_kernel, grid, dim3(kBlock), 0, st, ps, cells);
      FL_HIP(hipGetLastError());
    }
    const bool reinit = (it % every) == every - 1;
    if (reinit) {  // two level sets per solve, written back into cur[m]
      for (int m = 0; m < nmat; m += 2) {
        ls::EikJobs ej;
        const int nj = m + 1 < nmat ? 2 : 1;
        ej.j[0] = eik_job(nxt[m], g, cur[m], nullptr, w.reinit[m], c->eik);
        ej.j[1] = nj == 2 ? eik_job(nxt[m + 1], g, cur[m + 1], nullptr, w.reinit[m + 1], c->eik + 1)
                          : ej.j[0];
        if ((rc = run_reinit(c, ej, nj, g, dh, band, nullptr, st))) return rc;
      }
    } else {
      for (int m = 0; m < nmat; ++m) {
        double* tmp = cur[m];
        cur[m] = nxt[m];
        nxt[m] = tmp;
      }
    }
    dt = hc.dt;
    t = hc.t;
    it = hc.iters;
  }
  if ((rc = settle())) return rc;
  FL_HIP(hipStreamSynchronize(st));
  if (result) {
    result->run_time = t;
    result->iters = it;
    result->last_dt = dt;
    result->max_wavespeed = maxws;
  }
  return FL_OK;
}

}  // namespace

extern "C" {

int fl_ls_masks(const double* phi, int ndim, const int64_t* n, double dh, uint8_t* masks,
                void* stream) {
  ls::Grid g;
  int rc = check_shape(ndim, n, &g);
  if (rc) return rc;
  if (!phi || !masks) return fail(FL_ERR_ARGUMENT, "phi or masks is NULL");
  if ((rc = check_dh(dh))) return rc;
  return launch_masks(phi, g, dh, masks, (hipStream_t)stream);
}

int fl_ls_velocity(int nmat, const double* const* phis, const fl_field* states, double dh,
                   int target, double* V, void* stream) {
  if (nmat < 1 || nmat > 2) return fail(FL_ERR_UNSUPPORTED, "nmat %d: 1 or 2 materials", nmat);
  if (!phis || !states || !V) return fail(FL_ERR_ARGUMENT, "phis, states or V is NULL");
  if (target >= nmat) return fail(FL_ERR_ARGUMENT, "target %d >= nmat %d", target, nmat);
  ls::Grid g;
  int rc = check_shape(states[0].ndim, states[0].n, &g);
  if (rc) return rc;
  if ((rc = check_dh(dh))) return rc;
  for (int m = 0; m < nmat; ++m) {
    if (!phis[m]) return fail(FL_ERR_ARGUMENT, "phis[%d] is NULL", m);
    if ((rc = check_state(&states[m], g, "states[m]"))) return rc;
  }
  return launch_velocity(nmat, phis, states, g, dh, target, V, (hipStream_t)stream);
}

int fl_ls_advect(const double* phi, const double* V, int ndim, const int64_t* n, double dt,
                 double dh, double* out, double* work, void* stream) {
  ls::Grid g;
  int rc = check_shape(ndim, n, &g);
  if (rc) return rc;
  if (!phi || !V || !out || !work) return fail(FL_ERR_ARGUMENT, "NULL pointer argument");
  if ((rc = check_dh(dh))) return rc;
  if (!(dt >= 0.0) || !isfinite(dt)) return fail(FL_ERR_ARGUMENT, "dt %g must be >= 0", dt);
  if (out == phi || out == work) return fail(FL_ERR_ARGUMENT, "out aliases phi or work");
  return launch_advect(phi, V, g, dt, dh, out, work, (hipStream_t)stream);
}

int fl_ls_project(double* phi0, double* phi1, int64_t cells, void* stream) {
  if (!phi0 || !phi1) return fail(FL_ERR_ARGUMENT, "phi0 or phi1 is NULL");
  if (phi0 == phi1) return fail(FL_ERR_ARGUMENT, "phi0 and phi1 alias");
  if (cells < 0) return fail(FL_ERR_ARGUMENT, "cells < 0");
  if (cells == 0) return FL_OK;
  return launch_project(phi0, phi1, cells, (hipStream_t)stream);
}

int64_t fl_ls_reinit_work_bytes(int ndim, const int64_t* n) {
  ls::Grid g;
  if (check_shape(ndim, n, &g)) return -1;
  return reinit_work_bytes(g);
}

int fl_ls_reinit(fl_context* c, const double* phi, int ndim, const int64_t* n, double dh,
                 int band, double* out, uint8_t* band_mask, void* work, int64_t* iterations,
                 void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  ls::Grid g;
  int rc = check_shape(ndim, n, &g);
  if (rc) return rc;
  if (!phi || !out || !work) return fail(FL_ERR_ARGUMENT, "NULL pointer argument");
  if ((rc = check_dh(dh))) return rc;
  if (band < 1 || band > (1 << 20)) return fail(FL_ERR_ARGUMENT, "band %d out of range", band);
  if (out == phi) return fail(FL_ERR_ARGUMENT, "out aliases phi");
  if (!aligned(work)) return fail(FL_ERR_ARGUMENT, "work must be 256-byte aligned");
  DeviceGuard dg(c->device);
  ls::EikJobs jobs;
  jobs.j[0] = jobs.j[1] = eik_job(phi, g, out, band_mask, work, c->eik);
  return run_reinit(c, jobs, 1, g, dh, band, iterations, (hipStream_t)stream);
}

int fl_gfm_star(const double* phi_m, const double* phi_o, const fl_field* state_m,
                const fl_field* state_o, const fl_scheme* eos_m, const fl_scheme* eos_o,
                double dh, double* star, uint8_t* has_star, void* stream) {
  if (!phi_m || !phi_o || !star || !has_star || !state_m)
    return fail(FL_ERR_ARGUMENT, "NULL pointer argument");
  ls::Grid g;
  int rc = check_shape(state_m->ndim, state_m->n, &g);
  if (rc) return rc;
  if ((rc = check_state(state_m, g, "state_m"))) return rc;
  if ((rc = check_state(state_o, g, "state_o"))) return rc;
  if ((rc = check_scheme(eos_m, g.nd))) return rc;
  if ((rc = check_scheme(eos_o, g.nd))) return rc;
  if ((rc = check_dh(dh))) return rc;
  return launch_star(phi_m, phi_o, state_m, state_o, eos_m, eos_o, g, dh, star, has_star,
                     (hipStream_t)stream);
}

int64_t fl_gfm_extrapolate_work_bytes(int ndim, const int64_t* n) {
  ls::Grid g;
  if (check_shape(ndim, n, &g)) return -1;
  return extrap_work_bytes(g);
}

int fl_gfm_extrapolate(const double* phi_m, fl_field* state_m, const double* star,
                       const uint8_t* has_star, double dh, int sweeps, void* work,
                       void* stream) {
  if (!phi_m || !state_m || !star || !has_star || !work)
    return fail(FL_ERR_ARGUMENT, "NULL pointer argument");
  ls::Grid g;
  int rc = check_shape(state_m->ndim, state_m->n, &g);
  if (rc) return rc;
  if ((rc = check_state(state_m, g, "state_m"))) return rc;
  if ((rc = check_dh(dh))) return rc;
  if (sweeps < 0 || sweeps > 4096) return fail(FL_ERR_ARGUMENT, "sweeps %d out of range", sweeps);
  if ((uintptr_t)work & 255) return fail(FL_ERR_ARGUMENT, "work must be 256-byte aligned");
  return launch_extrapolate(phi_m, state_m, star, has_star, g, dh, sweeps, work,
                            (hipStream_t)stream);
}

int fl_mm_combine(const double* phi0, const fl_field* state0, const fl_field* state1,
                  fl_field* out, void* stream) {
  if (!phi0 || !state0) return fail(FL_ERR_ARGUMENT, "NULL pointer argument");
  ls::Grid g;
  int rc = check_shape(state0->ndim, state0->n, &g);
  if (rc) return rc;
  if ((rc = check_state(state0, g, "state0"))) return rc;
  if ((rc = check_state(state1, g, "state1"))) return rc;
  if ((rc = check_state(out, g, "out"))) return rc;
  return launch_combine(phi0, state0, state1, out, g, (hipStream_t)stream);
}

int64_t fl_mm_work_bytes(int ndim, const int64_t* n) {
  ls::Grid g;
  if (check_shape(ndim, n, &g)) return -1;
  return mm_layout(nullptr, g, 2).bytes;
}

int64_t fl_mm_work_bytes_n(int nmat, int ndim, const int64_t* n) {
  ls::Grid g;
  if (nmat < 2 || nmat > ls::kMaxMat || check_shape(ndim, n, &g)) return -1;
  return mm_layout(nullptr, g, nmat).bytes;
}

int fl_mm_simulate(fl_context* c, fl_material* mats, int nmat, const fl_sim_params* p,
                   const fl_mm_params* mp, void* work, int64_t work_bytes,
                   fl_sim_result* result, void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  if (!mats || !p || !work) return fail(FL_ERR_ARGUMENT, "NULL pointer argument");
  if (nmat < 2 || nmat > ls::kMaxMat)
    return fail(FL_ERR_UNSUPPORTED, "nmat %d: 2 to %d materials", nmat, ls::kMaxMat);
  ls::Grid g;
  int rc = check_shape(mats[0].state.ndim, mats[0].state.n, &g);
  if (rc) return rc;
  for (int m = 0; m < nmat; ++m) {
    if ((rc = check_state(&mats[m].state, g, "mats[m].state"))) return rc;
    if ((rc = check_state(&mats[m].work, g, "mats[m].work"))) return rc;
    if (mats[m].state.layout != mats[m].work.layout || mats[m].state.data == mats[m].work.data)
      return fail(FL_ERR_ARGUMENT, "mats[%d]: work must be a distinct buffer of the same layout", m);
    if (mats[m].state.dtype != mats[0].state.dtype || mats[m].work.dtype != mats[0].state.dtype)
      return fail(FL_ERR_ARGUMENT, "mats[%d]: every state and work grid needs one dtype", m);
    if (!mats[m].phi) return fail(FL_ERR_ARGUMENT, "mats[%d].phi is NULL", m);
    if ((rc = check_scheme(&mats[m].scheme, g.nd))) return rc;
    if (mats[m].scheme.numerics != mats[0].scheme.numerics)
      return fail(FL_ERR_ARGUMENT, "mats[%d]: every material needs the same numerics", m);
  }
  if (!(p->cfl > 0.0) || !isfinite(p->cfl)) return fail(FL_ERR_ARGUMENT, "cfl must be > 0");
  if ((rc = check_dh(p->dh))) return rc;
  if (p->max_iters < 0) return fail(FL_ERR_ARGUMENT, "max_iters must be >= 0");
  if (isnan(p->sim_time)) return fail(FL_ERR_ARGUMENT, "sim_time is NaN");
  const int band = mp && mp->band ? mp->band : 6;
  const int sweeps = mp && mp->extrap_sweeps ? mp->extrap_sweeps : 8;
  const int every = mp && mp->reinit_every ? mp->reinit_every : 1;
  if (band < 1 || band > (1 << 20) || sweeps < 0 || sweeps > 4096 || every < 1)
    return fail(FL_ERR_ARGUMENT, "band %d, extrap_sweeps %d or reinit_every %d out of range",
                band, sweeps, every);
  if (!aligned(work)) return fail(FL_ERR_ARGUMENT, "work must be 256-byte aligned");
  const MmWork w = mm_layout(work, g, nmat);
  if (work_bytes < w.bytes)
    return fail(FL_ERR_ARGUMENT, "work of %lld bytes, need %lld", (long long)work_bytes,
                (long long)w.bytes);
  DeviceGuard dg(c->device);
  hipStream_t st = (hipStream_t)stream;
  if (nmat > 2) return mm_simulate_n(c, mats, nmat, p, g, band, sweeps, every, w, result, st);
  const int64_t cells = g.nx * g.ny;
  const int S = 2 + g.nd;
  const int64_t esize = mats[0].state.dtype == FL_F64 ? 8 : 4;
  const double dh = p->dh;
  const bool fast = mats[0].scheme.numerics == FL_NUMERICS_FAST;  // one for every material
  double t = 0.0, dt = 0.0, maxws = 0.0;
  int64_t it = 0;
  // Each level set lives in the caller's phi or in the advection buffer tphi:
  // cur[m] holds it, nxt[m] receives the next advection.  A step without
  // reinitialisation swaps the two; one with it writes back into cur[m].
  double* cur[2] = {mats[0].phi, mats[1].phi};
  double* nxt[2] = {w.tphi[0], w.tphi[1]};
  auto settle = [&]() -> int {  // the level sets back in the caller's buffers
    for (int m = 0; m < 2; ++m)
      if (cur[m] != mats[m].phi)
        FL_HIP(hipMemcpyAsync(mats[m].phi, cur[m], (size_t)(cells * 8),
                              hipMemcpyDeviceToDevice, st));
    return FL_OK;
  };
  // dt lives on the device: the control kernel turns the maximum wavespeed
  // into dt, dt/dh, run time and iteration count in the context's control
  // block, the sweeps read dt/dh and the level-set advection dt from it.  The
  // host reads a copy of the block once per step, after the step's sweeps are
  // queued (so the device keeps working while it waits), to test the loop
  // condition and to stop on a NaN or vacuum maximum, as the reference loop
  // does, before the level sets move.
  fl::StepCtl* ctl = c->ctl;
  hipLaunchKernelGGL(fl::init_control_kernel, dim3(1), dim3(1), 0, st, ctl);
  FL_HIP(hipGetLastError());
  while (t < p->sim_time && it < p->max_iters) {
    // ghost cells: star states of both materials from the same data, then
    // extrapolation into each material's ghost band
    ls::ExtrapJob jobs[2];
    for (int m = 0; m < 2; ++m)
      jobs[m] = extrap_job(cur[m], view_of(&mats[m].state), g, w.extrap[m]);
    const MmGhost gw = {{w.star[0], w.star[1]}, {w.has[0], w.has[1]}};
    launch_star2(cur, mats, g, dh, gw, jobs, st);
    band_seed2(jobs, gw, g, dh, st);
    extrap_sweeps(jobs[0], jobs[1], 2, g, dh, sweeps, st);
    FL_HIP(hipGetLastError());
    // dt from the maximum wavespeed over both grids (ws_bits[0] is clear)
    if ((rc = launch_max_ws2(&mats[0].state, &mats[1].state, &mats[0].scheme, &mats[1].scheme,
                             &ctl->ws_bits[0], st)))
      return rc;
    hipLaunchKernelGGL(fl::mm_control_kernel, dim3(1), dim3(1), 0, st, ctl, p->cfl, dh);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(c->host_ctl, ctl, sizeof(fl::StepCtl), hipMemcpyDeviceToHost, st));
    FL_HIP(hipEventRecord(c->step_ev, st));
    // both materials' split solves, each sweep of both grids in one launch;
    // they read dt/dh from the control block and do nothing when it is not
    // active
    {
      const fl_field* const u[2] = {&mats[0].state, &mats[1].state};
      const fl_field* const v[2] = {&mats[0].work, &mats[1].work};
      const fl_scheme* const sc[2] = {&mats[0].scheme, &mats[1].scheme};
      if ((rc = launch_sweep2(c, u, v, 0, sc, 0.0, st, ctl))) return rc;
      if (g.nd > 1 && (rc = launch_sweep2(c, v, u, 1, sc, 0.0, st, ctl))) return rc;
    }
    FL_HIP(hipEventSynchronize(c->step_ev));
    const fl::StepCtl hc = *c->host_ctl;
    maxws = bits_to_double(hc.ws_bits[1]);
    if (!hc.active) {
      if (result) {
        result->run_time = t;
        result->iters = it;
        result->last_dt = dt;
        result->max_wavespeed = maxws;
      }
      if ((rc = settle())) return rc;
      FL_HIP(hipStreamSynchronize(st));
      return fail(FL_ERR_NUMERIC, "max wavespeed %g at step %lld (NaN or vacuum)", maxws,
                  (long long)it);
    }
    if (g.nd == 1)
      for (int m = 0; m < 2; ++m)
        FL_HIP(hipMemcpyAsync(mats[m].state.data, mats[m].work.data, (size_t)(cells * S * esize),
                              hipMemcpyDeviceToDevice, st));
    // level sets: speeds from the old level sets and new states (both in one
    // pass), advection with the device dt, projection, reinitialisation every
    // `every` steps
    if ((rc = launch_velocity2(cur, mats[0].state, mats[1].state, g, dh, w.V, st))) return rc;
    if ((rc = launch_advect2(cur, w.V, g, dh, nxt, w.adv, &ctl->dt, fast, st))) return rc;
    if ((rc = launch_project(nxt[0], nxt[1], cells, st))) return rc;
    const bool reinit = (it % every) == every - 1;
    if (reinit) {  // both level sets in one solve, written back into cur[m]
      ls::EikJobs ej;
      for (int m = 0; m < 2; ++m)
        ej.j[m] = eik_job(nxt[m], g, cur[m], nullptr, w.reinit[m], c->eik + m);
      if ((rc = run_reinit(c, ej, 2, g, dh, band, nullptr, st))) return rc;
    } else {
      for (int m = 0; m < 2; ++m) {
        double* tmp = cur[m];
        cur[m] = nxt[m];
        nxt[m] = tmp;
      }
    }
    dt = hc.dt;
    t = hc.t;
    it = hc.iters;
  }
  if ((rc = settle())) return rc;
  FL_HIP(hipStreamSynchronize(st));
  if (result) {
    result->run_time = t;
    result->iters = it;
    result->last_dt = dt;
    result->max_wavespeed = maxws;
  }
  return FL_OK;
}

}  // extern "C"
