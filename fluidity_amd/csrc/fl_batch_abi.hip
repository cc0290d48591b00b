// fl_batch_abi.hip -- host side of the batched entry points of
// include/fluidity_amd.h: problem tables, tile lists, the launch-shape
// autotuner and the hipGraph of one step.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "fl_abi_internal.hpp"
#include "fl_batch.hpp"

using namespace fl::abi;
namespace bt = fl::batch;

struct fl_batch {
  int device;
  int nprob;
  int dtype;
  int flux;
  std::vector<fl_batch_problem> problems;
  void* probs;               // bt::Prob<T>[nprob] (device)
  fl::StepCtl* ctls;         // [nprob] (device)
  fl::StepCtl* host_ctls;    // pinned mirror
  bt::Params* params;        // [nprob] (device)
  bt::Counters* cnt;         // device
  bt::Tile* tiles[2];        // device tile lists of the two phases
  int64_t cap[2];            // their capacity in tiles
  int ntiles[2];
  double* sums;              // [nprob * 4] (device)
  double* host_sums;         // pinned mirror
  fl_tuning tuning;
  int grid[2];               // persistent workgroups per phase
  hipStream_t cap_stream;    // stream used to capture the step graph
  hipGraph_t graph;
  hipGraphExec_t exec;
  bool graph_ok;
  bool vlc;                  // every problem: van Leer, conservative form
  bool fast;                 // FP64 batch in the fast numerics
};

namespace {

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Segments of B-2 cells a 1D tile marches (loads of the next segments overlap
// the current one).
constexpr int kSegments1D = 8;

bool valid_width(int b) { return b >= 64 && b <= fl::kMaxBlock && (b % 64) == 0; }

// Height of the x or y tiles of a 2D problem with ny rows: the tuning's base
// height, doubled (up to 8x) while the problem still gives at least
// kMinRowTiles tiles per column band.  Tall problems pay the warm-up rows of
// a tile (6 in x, 8 in y) less often; short ones keep short tiles, which end
// the phase evenly.
#ifndef FL_BATCH_MIN_ROW_TILES
#define FL_BATCH_MIN_ROW_TILES 8
#endif
constexpr int kMinRowTiles = FL_BATCH_MIN_ROW_TILES;
int64_t tile_rows(int64_t ny, int base) {
  int64_t r = base;
  for (int k = 0; k < 3 && ny >= 2 * r * kMinRowTiles; ++k) r *= 2;
  return r;
}

// Tile lists for a tuning: phase 0 holds the x tiles of the 2D problems
// (largest problems first) and then the 1D problems (longest first); phase 1
// the y tiles of the 2D problems.
void build_tiles(const fl_batch* b, const fl_tuning& t, std::vector<bt::Tile>& t0,
                 std::vector<bt::Tile>& t1) {
  std::vector<int> order(b->nprob);
  for (int p = 0; p < b->nprob; ++p) order[p] = p;
  auto cells = [&](int p) {
    const fl_field& f = b->problems[p].state;
    return f.n[0] * (f.ndim > 1 ? f.n[1] : 1);
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const int da = b->problems[a].state.ndim, dc = b->problems[c].state.ndim;
    if (da != dc) return da > dc;
    return cells(a) > cells(c);
  });
  t0.clear();
  t1.clear();
  for (int p : order) {
    const fl_field& f = b->problems[p].state;
    const int64_t nx = f.n[0];
    if (f.ndim == 1) {
      for (int64_t bx = 0; bx < ceil_div(ceil_div(nx, t.x_block - 2), kSegments1D); ++bx)
        t0.push_back({p, (int)bx, 0, 0});
    } else {
      const int64_t ny = f.n[1];
      const int64_t gx = ceil_div(nx, t.x_block - 2), gy = ceil_div(ny, tile_rows(ny, t.x_rows));
      for (int64_t by = 0; by < gy; ++by)
        for (int64_t bx = 0; bx < gx; ++bx) t0.push_back({p, (int)bx, (int)by, 0});
      const int64_t hx = ceil_div(nx, t.y_block), hy = ceil_div(ny, tile_rows(ny, t.y_rows));
      for (int64_t by = 0; by < hy; ++by)
        for (int64_t bx = 0; bx < hx; ++bx) t1.push_back({p, (int)bx, (int)by, 0});
    }
  }
}

template <typename T>
std::vector<bt::Prob<T>> make_probs(const fl_batch* b, const fl_tuning& t) {
  std::vector<bt::Prob<T>> v(b->nprob);
  for (int p = 0; p < b->nprob; ++p) {
    const fl_batch_problem& q = b->problems[p];
    bt::Prob<T>& P = v[p];
    memset(&P, 0, sizeof(P));
    fl::SweepArgs<T>& a = P.x;
    a.in = (const T*)q.state.data;
    a.out = (T*)q.work.data;
    strides(&q.state, a.es, a.cs);
    a.nx = q.state.n[0];
    a.ny = q.state.ndim > 1 ? q.state.n[1] : 1;
    a.mat = make_material<T>(&q.scheme);
    a.mat64 = make_material<double>(&q.scheme);
    a.limiter = q.scheme.limiter;
    a.form = q.scheme.form;
    a.bcx = q.scheme.bc[0];
    a.bcy = q.state.ndim > 1 ? q.scheme.bc[1] : 0;
    a.rows_per_block = q.state.ndim > 1 ? (int)tile_rows(a.ny, t.x_rows) : kSegments1D;
    a.ctl = b->ctls + p;
    a.ws_out = nullptr;
    P.ctl = b->ctls + p;
    P.nd = q.state.ndim;
    P.rows_y = (int)tile_rows(q.state.ndim > 1 ? q.state.n[1] : 1, t.y_rows);
  }
  return v;
}

template <typename T, int FLUX, bool FAST, bool VLC>
const void* phase_fn(int phase) {
  return phase ? reinterpret_cast<const void*>(&bt::phase_kernel<T, FLUX, 1, FAST, VLC>)
               : reinterpret_cast<const void*>(&bt::phase_kernel<T, FLUX, 0, FAST, VLC>);
}

template <typename T, bool FAST, bool VLC>
const void* phase_fn_f(int flux, int phase) {
  return flux == FL_FLUX_HLLC ? phase_fn<T, fl::kFluxHllc, FAST, VLC>(phase)
                              : phase_fn<T, fl::kFluxForce, FAST, VLC>(phase);
}

const void* phase_kernel_ptr(const fl_batch* b, int phase) {
  if (b->dtype == FL_F64) {
    if (b->fast)
      return b->vlc ? phase_fn_f<double, true, true>(b->flux, phase)
                    : phase_fn_f<double, true, false>(b->flux, phase);
    return b->vlc ? phase_fn_f<double, false, true>(b->flux, phase)
                  : phase_fn_f<double, false, false>(b->flux, phase);
  }
  return b->vlc ? phase_fn_f<float, false, true>(b->flux, phase)
                : phase_fn_f<float, false, false>(b->flux, phase);
}

// Uploads the problem table and tile lists of tuning t; sizes the persistent
// grids.  Synchronous (the host vectors are staging).
int apply_tuning(fl_batch* b, const fl_tuning& t) {
  std::vector<bt::Tile> t0, t1;
  build_tiles(b, t, t0, t1);
  if ((int64_t)t0.size() > b->cap[0] || (int64_t)t1.size() > b->cap[1])
    return fail(FL_ERR_ARGUMENT, "tile lists exceed their capacity");
  if (t0.size() > (size_t)0x7fffffff || t1.size() > (size_t)0x7fffffff)
    return fail(FL_ERR_UNSUPPORTED, "too many tiles");
  if (b->dtype == FL_F64) {
    auto v = make_probs<double>(b, t);
    FL_HIP(hipMemcpy(b->probs, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice));
  } else {
    auto v = make_probs<float>(b, t);
    FL_HIP(hipMemcpy(b->probs, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice));
  }
  if (!t0.empty())
    FL_HIP(hipMemcpy(b->tiles[0], t0.data(), sizeof(bt::Tile) * t0.size(), hipMemcpyHostToDevice));
  if (!t1.empty())
    FL_HIP(hipMemcpy(b->tiles[1], t1.data(), sizeof(bt::Tile) * t1.size(), hipMemcpyHostToDevice));
  b->ntiles[0] = (int)t0.size();
  b->ntiles[1] = (int)t1.size();
  hipDeviceProp_t prop;
  FL_HIP(hipGetDeviceProperties(&prop, b->device));
  for (int ph = 0; ph < 2; ++ph) {
    const int B = ph ? t.y_block : t.x_block;
    int per_cu = 0;
    FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, phase_kernel_ptr(b, ph), B,
        ph == 0 ? fl::x_tile_lds_bytes(B, 4, b->dtype == FL_F64 ? 8 : 4)
                : fl::y_tile_lds_bytes(B, b->dtype == FL_F64 ? 8 : 4)));
    if (per_cu < 1) per_cu = 1;
    b->grid[ph] = per_cu * prop.multiProcessorCount;
  }
  b->tuning = t;
  b->graph_ok = false;
  return FL_OK;
}

template <typename T, int FLUX, bool FAST, bool VLC>
void launch_phases_v(const fl_batch* b, hipStream_t st, int timing, double timing_dtdh,
                     int only_phase) {
  const bt::Prob<T>* probs = (const bt::Prob<T>*)b->probs;
  if (only_phase != 1 && b->ntiles[0] > 0)
    hipLaunchKernelGGL((bt::phase_kernel<T, FLUX, 0, FAST, VLC>), dim3(b->grid[0]),
                       dim3(b->tuning.x_block), fl::x_tile_lds_bytes(b->tuning.x_block, 4, sizeof(T)),
                       st, probs, (const bt::Tile*)b->tiles[0],
                       b->ntiles[0], b->cnt, timing, (T)timing_dtdh);
  if (only_phase != 0 && b->ntiles[1] > 0)
    hipLaunchKernelGGL((bt::phase_kernel<T, FLUX, 1, FAST, VLC>), dim3(b->grid[1]),
                       dim3(b->tuning.y_block), fl::y_tile_lds_bytes(b->tuning.y_block, sizeof(T)),
                       st, probs, (const bt::Tile*)b->tiles[1],
                       b->ntiles[1], b->cnt, timing, (T)timing_dtdh);
}

template <typename T, int FLUX>
void launch_phases_t(const fl_batch* b, hipStream_t st, int timing, double timing_dtdh,
                     int only_phase) {
  if constexpr (std::is_same<T, double>::value) {
    if (b->fast) {
      if (b->vlc)
        launch_phases_v<T, FLUX, true, true>(b, st, timing, timing_dtdh, only_phase);
      else
        launch_phases_v<T, FLUX, true, false>(b, st, timing, timing_dtdh, only_phase);
      return;
    }
  }
  if (b->vlc)
    launch_phases_v<T, FLUX, false, true>(b, st, timing, timing_dtdh, only_phase);
  else
    launch_phases_v<T, FLUX, false, false>(b, st, timing, timing_dtdh, only_phase);
}

// only_phase: -1 both phases, else just that one.
void launch_phases(const fl_batch* b, hipStream_t st, int timing, double timing_dtdh,
                   int only_phase) {
  if (b->dtype == FL_F64) {
    if (b->flux == FL_FLUX_HLLC)
      launch_phases_t<double, fl::kFluxHllc>(b, st, timing, timing_dtdh, only_phase);
    else
      launch_phases_t<double, fl::kFluxForce>(b, st, timing, timing_dtdh, only_phase);
  } else {
    if (b->flux == FL_FLUX_HLLC)
      launch_phases_t<float, fl::kFluxHllc>(b, st, timing, timing_dtdh, only_phase);
    else
      launch_phases_t<float, fl::kFluxForce>(b, st, timing, timing_dtdh, only_phase);
  }
}

int launch_step(const fl_batch* b, hipStream_t st) {
  const int np = b->nprob;
  hipLaunchKernelGGL(bt::control_kernel, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, st,
                     b->ctls, (const bt::Params*)b->params, np, b->cnt);
  launch_phases(b, st, 0, 0.0, -1);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

void launch_finish(const fl_batch* b, hipStream_t st, int dir) {
  if (b->dtype == FL_F64)
    hipLaunchKernelGGL(bt::finish_kernel<double>, dim3(b->nprob), dim3(256), 0, st,
                       (const bt::Prob<double>*)b->probs, b->nprob, dir);
  else
    hipLaunchKernelGGL(bt::finish_kernel<float>, dim3(b->nprob), dim3(256), 0, st,
                       (const bt::Prob<float>*)b->probs, b->nprob, dir);
}

int reset(fl_batch* b, hipStream_t st) {
  hipLaunchKernelGGL(bt::init_kernel, dim3((unsigned)ceil_div(b->nprob, 256)), dim3(256), 0, st,
                     b->ctls, b->nprob);
  FL_HIP(hipGetLastError());
  for (int p = 0; p < b->nprob; ++p) {
    int rc = launch_max_ws(&b->problems[p].state, &b->problems[p].scheme,
                           &b->ctls[p].ws_bits[0], st);
    if (rc) return rc;
  }
  return FL_OK;
}

void release(fl_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->probs) (void)hipFree(b->probs);
  if (b->ctls) (void)hipFree(b->ctls);
  if (b->host_ctls) (void)hipHostFree(b->host_ctls);
  if (b->params) (void)hipFree(b->params);
  if (b->cnt) (void)hipFree(b->cnt);
  if (b->tiles[0]) (void)hipFree(b->tiles[0]);
  if (b->tiles[1]) (void)hipFree(b->tiles[1]);
  if (b->sums) (void)hipFree(b->sums);
  if (b->host_sums) (void)hipHostFree(b->host_sums);
}

}  // namespace

extern "C" {

int fl_batch_create(fl_context* c, const fl_batch_problem* problems, int nprob, fl_batch** out,
                    void* stream) {
  if (!c || !problems || !out) return fail(FL_ERR_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (nprob < 1 || nprob > (1 << 20)) return fail(FL_ERR_ARGUMENT, "nprob %d out of range", nprob);
  const int dtype = problems[0].state.dtype;
  const int flux = problems[0].scheme.flux;
  if (flux != FL_FLUX_HLLC && flux != FL_FLUX_FORCE)
    return fail(FL_ERR_UNSUPPORTED, "batches run HLLC or FORCE (flux %d)", flux);
  for (int p = 0; p < nprob; ++p) {
    const fl_batch_problem& q = problems[p];
    int rc = check_field(&q.state, "problem state");
    if (!rc) rc = check_field(&q.work, "problem work");
    if (!rc) rc = check_scheme(&q.scheme, q.state.ndim);
    if (rc) return fail(rc, "problem %d: %s", p, fl_last_error());
    if (q.state.dtype == FL_F64 && q.scheme.numerics != problems[0].scheme.numerics)
      return fail(FL_ERR_UNSUPPORTED, "problem %d: one FP64 numerics per batch", p);
    if (q.work.dtype != q.state.dtype || q.work.layout != q.state.layout ||
        q.work.ndim != q.state.ndim || q.work.n[0] != q.state.n[0] ||
        (q.state.ndim > 1 && q.work.n[1] != q.state.n[1]) || q.work.data == q.state.data)
      return fail(FL_ERR_ARGUMENT, "problem %d: work must be a distinct buffer like state", p);
    if (q.state.dtype != dtype) return fail(FL_ERR_ARGUMENT, "problem %d: mixed dtypes", p);
    if (q.scheme.flux != flux)
      return fail(FL_ERR_UNSUPPORTED, "problem %d: one flux per batch", p);
    const fl_sim_params& sp = q.params;
    if (!(sp.cfl > 0.0) || !isfinite(sp.cfl) || !(sp.dh > 0.0) || !isfinite(sp.dh) ||
        sp.max_iters < 0 || isnan(sp.sim_time))
      return fail(FL_ERR_ARGUMENT, "problem %d: bad cfl/dh/sim_time/max_iters", p);
    const int64_t ny = q.state.ndim > 1 ? q.state.n[1] : 1;
    if (q.state.n[0] > 0x3fffffff || ny > 0x3fffffff)
      return fail(FL_ERR_UNSUPPORTED, "problem %d too large for the tile index", p);
  }
  DeviceGuard g(c->device);
  fl_batch* b = new fl_batch();
  b->device = c->device;
  b->nprob = nprob;
  b->dtype = dtype;
  b->flux = flux;
  b->fast = dtype == FL_F64 && problems[0].scheme.numerics == FL_NUMERICS_FAST;
  b->vlc = FL_VLC_KERNELS != 0;
  for (int p = 0; p < nprob; ++p)
    b->vlc = b->vlc && problems[p].scheme.limiter == FL_LIMITER_VAN_LEER &&
             problems[p].scheme.form == FL_FORM_CONSERVATIVE;
  b->problems.assign(problems, problems + nprob);
  b->probs = nullptr;
  b->ctls = b->host_ctls = nullptr;
  b->params = nullptr;
  b->cnt = nullptr;
  b->tiles[0] = b->tiles[1] = nullptr;
  b->sums = b->host_sums = nullptr;
  b->cap_stream = nullptr;
  b->graph = nullptr;
  b->exec = nullptr;
  b->graph_ok = false;
  // capacity: the finest tiling the tuning limits allow (64-lane tiles, 1 row)
  int64_t c0 = 0, c1 = 0;
  for (int p = 0; p < nprob; ++p) {
    const fl_field& f = problems[p].state;
    const int64_t ny = f.ndim > 1 ? f.n[1] : 1;
    c0 += ceil_div(f.n[0], 62) * ny;
    if (f.ndim > 1) c1 += ceil_div(f.n[0], 64) * ny;
  }
  b->cap[0] = c0;
  b->cap[1] = c1;
  const size_t prob_bytes =
      dtype == FL_F64 ? sizeof(bt::Prob<double>) : sizeof(bt::Prob<float>);
  hipError_t e = hipMalloc(&b->probs, prob_bytes * nprob);
  if (e == hipSuccess) e = hipMalloc((void**)&b->ctls, sizeof(fl::StepCtl) * nprob);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->host_ctls, sizeof(fl::StepCtl) * nprob, 0);
  if (e == hipSuccess) e = hipMalloc((void**)&b->params, sizeof(bt::Params) * nprob);
  if (e == hipSuccess) e = hipMalloc((void**)&b->cnt, sizeof(bt::Counters));
  if (e == hipSuccess) e = hipMalloc((void**)&b->tiles[0], sizeof(bt::Tile) * (c0 ? c0 : 1));
  if (e == hipSuccess) e = hipMalloc((void**)&b->tiles[1], sizeof(bt::Tile) * (c1 ? c1 : 1));
  if (e == hipSuccess) e = hipMalloc((void**)&b->sums, sizeof(double) * 4 * nprob);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->host_sums, sizeof(double) * 4 * nprob, 0);
  if (e == hipSuccess) e = hipStreamCreate(&b->cap_stream);
  if (e != hipSuccess) {
    release(b);
    delete b;
    return fail(FL_ERR_HIP, "batch allocation failed: %s", hipGetErrorString(e));
  }
  std::vector<bt::Params> hp(nprob);
  for (int p = 0; p < nprob; ++p) {
    hp[p].cfl = problems[p].params.cfl;
    hp[p].dh = problems[p].params.dh;
    hp[p].sim_time = problems[p].params.sim_time;
    hp[p].max_iters = problems[p].params.max_iters;
  }
  fl_tuning t;
  t.x_block = 256;
  t.x_rows = 32;
  t.y_block = 256;
  t.y_rows = 64;
  int rc = FL_OK;
  e = hipMemcpy(b->params, hp.data(), sizeof(bt::Params) * nprob, hipMemcpyHostToDevice);
  if (e != hipSuccess) rc = fail(FL_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(e));
  if (!rc) rc = apply_tuning(b, t);
  if (!rc) rc = reset(b, (hipStream_t)stream);
  if (rc) {
    release(b);
    delete b;
    return rc;
  }
  *out = b;
  return FL_OK;
}

int fl_batch_destroy(fl_batch* b) {
  if (!b) return FL_OK;
  DeviceGuard g(b->device);
  release(b);
  delete b;
  return FL_OK;
}

int fl_batch_set_tuning(fl_batch* b, const fl_tuning* t) {
  if (!b || !t) return fail(FL_ERR_ARGUMENT, "NULL batch or tuning");
  fl_tuning n = b->tuning;
  if (t->x_block) n.x_block = t->x_block;
  if (t->x_rows) n.x_rows = t->x_rows;
  if (t->y_block) n.y_block = t->y_block;
  if (t->y_rows) n.y_rows = t->y_rows;
  if (!valid_width(n.x_block) || !valid_width(n.y_block))
    return fail(FL_ERR_ARGUMENT, "block widths must be multiples of 64 in [64, %d]", fl::kMaxBlock);
  if (n.x_rows < 1 || n.x_rows > 4096 || n.y_rows < 1 || n.y_rows > 4096)
    return fail(FL_ERR_ARGUMENT, "rows per tile must be in [1, 4096]");
  DeviceGuard g(b->device);
  FL_HIP(hipDeviceSynchronize());  // tables in use by queued steps
  return apply_tuning(b, n);
}

int fl_batch_get_tuning(const fl_batch* b, fl_tuning* t) {
  if (!b || !t) return fail(FL_ERR_ARGUMENT, "NULL batch or tuning");
  *t = b->tuning;
  return FL_OK;
}

int fl_batch_reset(fl_batch* b, void* stream) {
  if (!b) return fail(FL_ERR_ARGUMENT, "batch is NULL");
  DeviceGuard g(b->device);
  return reset(b, (hipStream_t)stream);
}

int fl_batch_autotune(fl_batch* b, int reps, fl_tuning* chosen, void* stream) {
  if (!b) return fail(FL_ERR_ARGUMENT, "batch is NULL");
  if (reps < 1 || reps > 100) return fail(FL_ERR_ARGUMENT, "reps %d out of [1, 100]", reps);
  DeviceGuard g(b->device);
  hipStream_t st = (hipStream_t)stream;
  // The timing launches use the work buffers as output: move 1D results that
  // live there into the state buffers first and back afterwards.
  launch_finish(b, st, 0);
  FL_HIP(hipStreamSynchronize(st));  // the tables are rewritten below
  static const int widths[] = {64, 128, 256};
  static const int xrows[] = {4, 8, 16, 32, 64};
  static const int yrows[] = {16, 32, 64, 128};
  hipEvent_t e0, e1;
  FL_HIP(hipEventCreate(&e0));
  FL_HIP(hipEventCreate(&e1));
  const double dtdh = 1e-3;
  fl_tuning best = b->tuning;
  int rc = FL_OK;
  for (int ph = 0; ph < 2 && !rc; ++ph) {
    float best_ms = 1e30f;
    for (int w : widths) {
      const int* rows = ph ? yrows : xrows;
      const int nr = ph ? 4 : 5;
      for (int k = 0; k < nr && !rc; ++k) {
        fl_tuning t = best;
        if (ph == 0) {
          t.x_block = w;
          t.x_rows = rows[k];
        } else {
          t.y_block = w;
          t.y_rows = rows[k];
        }
        if ((rc = apply_tuning(b, t))) break;
        if (b->ntiles[ph] == 0) break;
        float total = 0.0f;
        for (int r = 0; r <= reps && !rc; ++r) {  // r = 0 warms up
          hipLaunchKernelGGL(bt::reset_counters_kernel, dim3(1), dim3(1), 0, st, b->cnt);
          if (hipEventRecord(e0, st) != hipSuccess) rc = fail(FL_ERR_HIP, "hipEventRecord");
          launch_phases(b, st, 1, dtdh, ph);
          if (hipEventRecord(e1, st) != hipSuccess) rc = fail(FL_ERR_HIP, "hipEventRecord");
          if (hipEventSynchronize(e1) != hipSuccess) rc = fail(FL_ERR_HIP, "hipEventSynchronize");
          float ms = 0.0f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (r > 0) total += ms;
        }
        if (!rc && total < best_ms) {
          best_ms = total;
          best = t;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) rc = apply_tuning(b, best);
  if (!rc) launch_finish(b, st, 1);
  if (!rc && chosen) *chosen = best;
  return rc;
}

int fl_batch_step(fl_batch* b, int64_t steps, int use_graph, void* stream) {
  if (!b) return fail(FL_ERR_ARGUMENT, "batch is NULL");
  if (steps < 0) return fail(FL_ERR_ARGUMENT, "steps < 0");
  DeviceGuard g(b->device);
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) {
    for (int64_t k = 0; k < steps; ++k) {
      int rc = launch_step(b, st);
      if (rc) return rc;
    }
    return FL_OK;
  }
  if (!b->graph_ok) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    FL_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(b, b->cap_stream);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(FL_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    b->graph = gr;
    FL_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_ok = true;
  }
  for (int64_t k = 0; k < steps; ++k) FL_HIP(hipGraphLaunch(b->exec, st));
  return FL_OK;
}

int fl_batch_finish(fl_batch* b, fl_sim_result* results, void* stream) {
  if (!b) return fail(FL_ERR_ARGUMENT, "batch is NULL");
  DeviceGuard g(b->device);
  hipStream_t st = (hipStream_t)stream;
  launch_finish(b, st, 0);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(b->host_ctls, b->ctls, sizeof(fl::StepCtl) * b->nprob,
                        hipMemcpyDeviceToHost, st));
  FL_HIP(hipStreamSynchronize(st));
  int bad = -1;
  double bad_ws = 0.0;
  for (int p = 0; p < b->nprob; ++p) {
    const fl::StepCtl& h = b->host_ctls[p];
    const double ws = bits_to_double(h.ws_bits[h.iters & 1]);
    if (results) {
      results[p].run_time = h.t;
      results[p].iters = h.iters;
      results[p].last_dt = h.dt;
      results[p].max_wavespeed = ws;
    }
    if (bad < 0 && (!(ws > 0.0) || isnan(ws))) {
      bad = p;
      bad_ws = ws;
    }
  }
  if (bad >= 0)
    return fail(FL_ERR_NUMERIC, "problem %d: max wavespeed %g (NaN or vacuum)", bad, bad_ws);
  return FL_OK;
}

int fl_batch_conservation_sums(fl_batch* b, double* sums, void* stream) {
  if (!b || !sums) return fail(FL_ERR_ARGUMENT, "NULL argument");
  DeviceGuard g(b->device);
  hipStream_t st = (hipStream_t)stream;
  if (b->dtype == FL_F64)
    hipLaunchKernelGGL(bt::sums_kernel<double>, dim3(b->nprob), dim3(256), 0, st,
                       (const bt::Prob<double>*)b->probs, b->nprob, b->sums);
  else
    hipLaunchKernelGGL(bt::sums_kernel<float>, dim3(b->nprob), dim3(256), 0, st,
                       (const bt::Prob<float>*)b->probs, b->nprob, b->sums);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(b->host_sums, b->sums, sizeof(double) * 4 * b->nprob,
                        hipMemcpyDeviceToHost, st));
  FL_HIP(hipStreamSynchronize(st));
  memcpy(sums, b->host_sums, sizeof(double) * 4 * b->nprob);
  return FL_OK;
}

}  // extern "C"
