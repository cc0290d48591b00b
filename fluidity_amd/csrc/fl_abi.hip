// fl_abi.hip -- host side of the C ABI declared in include/fluidity_amd.h.
//
// Validates arguments, builds the kernel argument blocks and launches the
// gfx950 kernels of fl_sweep.hip on the caller's stream.  No allocation
// happens outside fl_ctx_create.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "fl_abi_internal.hpp"

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace fl {
namespace abi {

int fail(int status, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

bool valid_block(int b) { return b >= 64 && b <= fl::kMaxBlock && (b % 64) == 0; }

int resolve_tuning(const fl_tuning* t, fl_tuning* out) {
  fl_tuning n;
  n.x_block = t && t->x_block ? t->x_block : 256;
  n.x_rows = t && t->x_rows ? t->x_rows : 32;
  n.y_block = t && t->y_block ? t->y_block : 256;
  n.y_rows = t && t->y_rows ? t->y_rows : 64;
  if (!valid_block(n.x_block) || !valid_block(n.y_block))
    return fail(FL_ERR_ARGUMENT, "block widths must be multiples of 64 in [64, %d]", fl::kMaxBlock);
  if (n.x_rows < 1 || n.x_rows > 4096 || n.y_rows < 1 || n.y_rows > 4096)
    return fail(FL_ERR_ARGUMENT, "rows per workgroup must be in [1, 4096]");
  *out = n;
  return FL_OK;
}

int check_field(const fl_field* f, const char* name) {
  if (!f) return fail(FL_ERR_ARGUMENT, "%s is NULL", name);
  if (!f->data) return fail(FL_ERR_ARGUMENT, "%s->data is NULL", name);
  if (f->dtype != FL_F64 && f->dtype != FL_F32)
    return fail(FL_ERR_ARGUMENT, "%s->dtype %d is not FL_F64/FL_F32", name, f->dtype);
  if (f->layout != FL_LAYOUT_AOS && f->layout != FL_LAYOUT_SOA)
    return fail(FL_ERR_ARGUMENT, "%s->layout %d is not AoS/SoA", name, f->layout);
  if (f->ndim < 1 || f->ndim > 3)
    return fail(FL_ERR_ARGUMENT, "%s->ndim %d out of range", name, f->ndim);
  if (f->ndim == 3) return fail(FL_ERR_UNSUPPORTED, "%s: 3D grids are not supported", name);
  for (int d = 0; d < f->ndim; ++d)
    if (f->n[d] < 4)
      return fail(FL_ERR_ARGUMENT, "%s->n[%d] = %lld: need at least 4 cells (ghost pad)",
                  name, d, (long long)f->n[d]);
  if (f->ndim == 2 && f->n[1] > (int64_t)65535 * 4096)
    return fail(FL_ERR_ARGUMENT, "%s->n[1] too large", name);
  // the 2D sweeps address a row with 32-bit byte offsets (4 FP64 components)
  if (f->ndim == 2 && f->n[0] > ((int64_t)1 << 27))
    return fail(FL_ERR_ARGUMENT, "%s->n[0] = %lld: 2D rows hold at most 2^27 cells", name,
                (long long)f->n[0]);
  return FL_OK;
}

int check_pair(const fl_field* in, const fl_field* out) {
  int rc = check_field(in, "in");
  if (rc) return rc;
  rc = check_field(out, "out");
  if (rc) return rc;
  if (in->dtype != out->dtype || in->layout != out->layout || in->ndim != out->ndim)
    return fail(FL_ERR_ARGUMENT, "in and out differ in dtype/layout/ndim");
  for (int d = 0; d < in->ndim; ++d)
    if (in->n[d] != out->n[d]) return fail(FL_ERR_ARGUMENT, "in and out differ in shape");
  if (in->data == out->data)
    return fail(FL_ERR_ARGUMENT, "in and out alias: a sweep cannot update in place");
  return FL_OK;
}

int check_scheme(const fl_scheme* s, int ndim) {
  if (!s) return fail(FL_ERR_ARGUMENT, "scheme is NULL");
  if (s->limiter != FL_LIMITER_VAN_LEER && s->limiter != FL_LIMITER_SUPERBEE)
    return fail(FL_ERR_ARGUMENT, "scheme->limiter %d invalid", s->limiter);
  if (s->form != FL_FORM_CONSERVATIVE && s->form != FL_FORM_PRIMITIVE)
    return fail(FL_ERR_ARGUMENT, "scheme->form %d invalid", s->form);
  if (s->flux < FL_FLUX_HLLC || s->flux > FL_FLUX_RICHTMYER)
    return fail(FL_ERR_ARGUMENT, "scheme->flux %d invalid", s->flux);
  if (s->eos != FL_EOS_IDEAL && s->eos != FL_EOS_STIFF && s->eos != FL_EOS_STIFF_CORRECTED)
    return fail(FL_ERR_ARGUMENT, "scheme->eos %d invalid", s->eos);
  if (!(s->gamma > 1.0) || !isfinite(s->gamma))
    return fail(FL_ERR_ARGUMENT, "scheme->gamma %g must be finite and > 1", s->gamma);
  if (!isfinite(s->p_inf)) return fail(FL_ERR_ARGUMENT, "scheme->p_inf not finite");
  for (int d = 0; d < ndim; ++d)
    if (s->bc[d] < FL_BC_TRANSMISSIVE || s->bc[d] > FL_BC_PERIODIC)
      return fail(FL_ERR_ARGUMENT, "scheme->bc[%d] %d invalid", d, s->bc[d]);
  if (s->numerics != FL_NUMERICS_EXACT && s->numerics != FL_NUMERICS_FAST)
    return fail(FL_ERR_ARGUMENT, "scheme->numerics %d invalid", s->numerics);
  return FL_OK;
}

int require_exact(const fl_scheme* s, const char* what) {
  if (s->numerics != FL_NUMERICS_EXACT)
    return fail(FL_ERR_UNSUPPORTED, "%s runs the exact FP64 numerics only", what);
  return FL_OK;
}

void strides(const fl_field* f, int64_t& es, int64_t& cs) {
  const int S = 2 + f->ndim;
  const int64_t cells = f->n[0] * (f->ndim > 1 ? f->n[1] : 1);
  if (f->layout == FL_LAYOUT_AOS) {
    es = S;
    cs = 1;
  } else {
    es = 1;
    cs = cells;
  }
}

int64_t cell_count(const fl_field* f) { return f->n[0] * (f->ndim > 1 ? f->n[1] : 1); }

template <typename T>
fl::SweepArgs<T> make_args(const fl_context* ctx, const fl_field* in, const fl_field* out,
                           const fl_scheme* s, double dtdh, const fl::StepCtl* ctl,
                           unsigned long long* ws_out, int rows) {
  fl::SweepArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.in = (const T*)in->data;
  a.out = (T*)out->data;
  strides(in, a.es, a.cs);
  a.nx = in->n[0];
  a.ny = in->ndim > 1 ? in->n[1] : 1;
  a.mat = make_material<T>(s);
  a.mat64 = make_material<double>(s);
  a.limiter = s->limiter;
  a.form = s->form;
  a.bcx = s->bc[0];
  a.bcy = in->ndim > 1 ? s->bc[1] : 0;
  a.rows_per_block = rows;
  a.dtdh_host = (T)dtdh;
  a.ctl = ctl;
  a.ws_out = ws_out;
  (void)ctx;
  return a;
}

// Rows per workgroup, halved (down to 4) while the launch would have fewer
// than kMinWaves waves (1.5 per SIMD of the 1024 on the device): small grids
// keep the machine filled whatever the tuning says.
constexpr int64_t kMinWaves = 1536;
int fit_rows(int rows, int64_t bands, int64_t ny, int block) {
  const int64_t waves = block / 64;
  while (rows > 4 && bands * ((ny + rows - 1) / rows) * waves < kMinWaves) rows /= 2;
  return rows;
}

// Kernels with van Leer in conservative form fixed at compile time exist for
// HLLC and FORCE.
template <int FLUX>
constexpr bool has_vlc() { return FLUX == fl::kFluxHllc || FLUX == fl::kFluxForce; }
inline bool is_vlc(int limiter, int form) {
  return FL_VLC_KERNELS && limiter == fl::kLimVanLeer && form == fl::kFormCons;
}

template <typename T, int FLUX, bool FAST, bool VLC>
void launch_flux_v(int ndim, int dim, dim3 grid, int B, const fl::SweepArgs<T>& a,
                   hipStream_t stream) {
  if (ndim == 1)
    hipLaunchKernelGGL((fl::sweep_x_kernel<T, 1, FLUX, FAST, VLC>), grid, dim3(B),
                       fl::x_tile_lds_bytes(B, 3, sizeof(T)), stream, a);
  else if (dim == 0)
    hipLaunchKernelGGL((fl::sweep_x_kernel<T, 2, FLUX, FAST, VLC>), grid, dim3(B),
                       fl::x_tile_lds_bytes(B, 4, sizeof(T)), stream, a);
  else
    hipLaunchKernelGGL((fl::sweep_y_kernel<T, FLUX, FAST, VLC>), grid, dim3(B),
                       fl::y_tile_lds_bytes(B, sizeof(T)), stream, a);
}

// One sweep kernel of flux FLUX: the 1D sweep, or the x or y sweep of a 2D grid.
template <typename T, int FLUX, bool FAST>
void launch_flux(int ndim, int dim, dim3 grid, int B, const fl::SweepArgs<T>& a,
                 hipStream_t stream) {
  if constexpr (has_vlc<FLUX>()) {
    if (is_vlc(a.limiter, a.form)) {
      launch_flux_v<T, FLUX, FAST, true>(ndim, dim, grid, B, a, stream);
      return;
    }
  }
  launch_flux_v<T, FLUX, FAST, false>(ndim, dim, grid, B, a, stream);
}

// The fast FP64 numerics were requested (FP32 sweeps use them always).
inline bool fast_numerics(const fl_scheme* s) { return s->numerics == FL_NUMERICS_FAST; }

// Workgroups of B lanes of one sweep kernel that a CU holds at once.
template <typename T, int FLUX, bool FAST, bool VLC>
const void* sweep_kernel_2d(int dim) {
  return dim == 0 ? reinterpret_cast<const void*>(&fl::sweep_x_kernel<T, 2, FLUX, FAST, VLC>)
                  : reinterpret_cast<const void*>(&fl::sweep_y_kernel<T, FLUX, FAST, VLC>);
}

template <typename T, int FLUX, bool FAST>
int sweep_blocks_per_cu(int dim, int B, bool vlc) {
  const void* k = sweep_kernel_2d<T, FLUX, FAST, false>(dim);
  if constexpr (has_vlc<FLUX>())
    if (vlc) k = sweep_kernel_2d<T, FLUX, FAST, true>(dim);
  int n = 0;
  const size_t lds =
      dim == 0 ? fl::x_tile_lds_bytes(B, 4, sizeof(T)) : fl::y_tile_lds_bytes(B, sizeof(T));
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, B, lds) != hipSuccess) return 0;
  return n;
}

template <typename T, bool FAST>
int sweep_blocks_per_cu_f(int dim, int flux, int B, bool vlc) {
  switch (flux) {
    case FL_FLUX_HLLC: return sweep_blocks_per_cu<T, fl::kFluxHllc, FAST>(dim, B, vlc);
    case FL_FLUX_FORCE: return sweep_blocks_per_cu<T, fl::kFluxForce, FAST>(dim, B, vlc);
    case FL_FLUX_LAX_FRIEDRICHS:
      return sweep_blocks_per_cu<T, fl::kFluxLaxFriedrichs, FAST>(dim, B, vlc);
    default: return sweep_blocks_per_cu<T, fl::kFluxRichtmyer, FAST>(dim, B, vlc);
  }
}

template <typename T>
int sweep_blocks_per_cu_t(int dim, const fl_scheme* s, int B) {
  const bool vlc = is_vlc(s->limiter, s->form);
  if (std::is_same<T, double>::value && fast_numerics(s))
    return sweep_blocks_per_cu_f<T, true>(dim, s->flux, B, vlc);
  return sweep_blocks_per_cu_f<T, false>(dim, s->flux, B, vlc);
}

// Tile heights to time for one 2D sweep: the powers of two in `base`, and,
// for R = 1..16, the smallest height whose launch fits in R rounds of the
// workgroups the device holds at once (`slots`).  On a grid whose launch is
// a few rounds, one a little over a whole number of rounds leaves most of
// the machine idle in its last round.
std::vector<int> row_candidates(const int* base, int nbase, int64_t bands, int64_t ny,
                                int64_t slots) {
  std::vector<int> rows(base, base + nbase);
  for (int64_t R = 1; slots > 0 && R <= 16; ++R) {
    const int64_t chunks = R * slots / bands;
    if (chunks < 1) continue;
    const int64_t h = (ny + chunks - 1) / chunks;
    if (h >= 4 && h <= 128 && std::find(rows.begin(), rows.end(), (int)h) == rows.end())
      rows.push_back((int)h);
  }
  return rows;
}

template <typename T, bool FAST>
void launch_flux_t(int flux, int ndim, int dim, dim3 grid, int B, const fl::SweepArgs<T>& a,
                   hipStream_t stream) {
  switch (flux) {
    case FL_FLUX_HLLC: launch_flux<T, fl::kFluxHllc, FAST>(ndim, dim, grid, B, a, stream); break;
    case FL_FLUX_FORCE: launch_flux<T, fl::kFluxForce, FAST>(ndim, dim, grid, B, a, stream); break;
    case FL_FLUX_LAX_FRIEDRICHS:
      launch_flux<T, fl::kFluxLaxFriedrichs, FAST>(ndim, dim, grid, B, a, stream);
      break;
    default: launch_flux<T, fl::kFluxRichtmyer, FAST>(ndim, dim, grid, B, a, stream); break;
  }
}

template <typename T>
int launch_sweep_t(const fl_context* ctx, const fl_field* in, const fl_field* out, int dim,
                   const fl_scheme* s, double dtdh, const fl::StepCtl* ctl,
                   unsigned long long* ws_out, hipStream_t stream) {
  fl_tuning tu = ctx->tuning;
  if (in->ndim == 2) {
    const int64_t nx = in->n[0], ny = in->n[1];
    tu.x_rows = fit_rows(tu.x_rows, (nx + tu.x_block - 3) / (tu.x_block - 2), ny, tu.x_block);
    tu.y_rows = fit_rows(tu.y_rows, (nx + tu.y_block - 1) / tu.y_block, ny, tu.y_block);
  }
  dim3 grid;
  int B;
  fl::SweepArgs<T> a;
  if (in->ndim == 1) {
    B = tu.x_block;
    a = make_args<T>(ctx, in, out, s, dtdh, ctl, ws_out, 1);
    grid = dim3((unsigned)((a.nx + (B - 3)) / (B - 2)), 1, 1);
  } else if (dim == 0) {
    B = tu.x_block;
    a = make_args<T>(ctx, in, out, s, dtdh, ctl, ws_out, tu.x_rows);
    grid = dim3((unsigned)((a.nx + (B - 3)) / (B - 2)),
                (unsigned)((a.ny + tu.x_rows - 1) / tu.x_rows), 1);
  } else {
    B = tu.y_block;
    a = make_args<T>(ctx, in, out, s, dtdh, ctl, ws_out, tu.y_rows);
    grid = dim3((unsigned)((a.nx + B - 1) / B), (unsigned)((a.ny + tu.y_rows - 1) / tu.y_rows), 1);
  }
  if (std::is_same<T, double>::value && fast_numerics(s))
    launch_flux_t<T, true>(s->flux, in->ndim, dim, grid, B, a, stream);
  else
    launch_flux_t<T, false>(s->flux, in->ndim, dim, grid, B, a, stream);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int launch_sweep(const fl_context* ctx, const fl_field* in, const fl_field* out, int dim,
                 const fl_scheme* s, double dtdh, const fl::StepCtl* ctl,
                 unsigned long long* ws_out, hipStream_t stream) {
  if (in->dtype == FL_F64)
    return launch_sweep_t<double>(ctx, in, out, dim, s, dtdh, ctl, ws_out, stream);
  return launch_sweep_t<float>(ctx, in, out, dim, s, dtdh, ctl, ws_out, stream);
}

// Both materials' sweeps of dimension `dim` in one launch: in[m] -> out[m]
// with scheme s[m].  The two grids share shape, layout and dtype; the flux
// methods must be equal (the launch instantiates one), else each grid is
// swept on its own.
template <typename T, int FLUX, bool FAST, bool VLC>
void launch_flux2_v(int ndim, int dim, dim3 grid, int B, const fl::SweepArgs<T>& a0,
                    const fl::SweepArgs<T>& a1, hipStream_t stream) {
  if (ndim == 1)
    hipLaunchKernelGGL((fl::sweep_x2_kernel<T, 1, FLUX, FAST, VLC>), grid, dim3(B),
                       fl::x_tile_lds_bytes(B, 3, sizeof(T)), stream, a0, a1);
  else if (dim == 0)
    hipLaunchKernelGGL((fl::sweep_x2_kernel<T, 2, FLUX, FAST, VLC>), grid, dim3(B),
                       fl::x_tile_lds_bytes(B, 4, sizeof(T)), stream, a0, a1);
  else
    hipLaunchKernelGGL((fl::sweep_y2_kernel<T, FLUX, FAST, VLC>), grid, dim3(B),
                       fl::y_tile_lds_bytes(B, sizeof(T)), stream, a0, a1);
}

// fast: the fast FP64 numerics (FP64 only; the FP32 kernels have their own)
template <typename T, int FLUX>
void launch_flux2(int ndim, int dim, dim3 grid, int B, const fl::SweepArgs<T>& a0,
                  const fl::SweepArgs<T>& a1, bool fast, hipStream_t stream) {
  const bool vlc = has_vlc<FLUX>() && is_vlc(a0.limiter, a0.form) && is_vlc(a1.limiter, a1.form);
  if constexpr (std::is_same<T, double>::value) {
    if (fast) {
      if constexpr (has_vlc<FLUX>())
        if (vlc) return launch_flux2_v<T, FLUX, true, true>(ndim, dim, grid, B, a0, a1, stream);
      return launch_flux2_v<T, FLUX, true, false>(ndim, dim, grid, B, a0, a1, stream);
    }
  }
  if constexpr (has_vlc<FLUX>())
    if (vlc) return launch_flux2_v<T, FLUX, false, true>(ndim, dim, grid, B, a0, a1, stream);
  launch_flux2_v<T, FLUX, false, false>(ndim, dim, grid, B, a0, a1, stream);
}

template <typename T>
int launch_sweep2_t(const fl_context* ctx, const fl_field* const in[2],
                    const fl_field* const out[2], int dim, const fl_scheme* const s[2],
                    double dtdh, hipStream_t stream, const fl::StepCtl* ctl) {
  fl_tuning tu = ctx->tuning;
  const fl_field* f = in[0];
  int rows = 0, B;
  dim3 grid;
  if (f->ndim == 1) {
    B = tu.x_block;
    grid = dim3((unsigned)((f->n[0] + (B - 3)) / (B - 2)), 1, 2);
  } else {
    const int64_t nx = f->n[0], ny = f->n[1];
    // the waves of both grids count towards filling the machine
    if (dim == 0) {
      B = tu.x_block;
      rows = fit_rows(tu.x_rows, 2 * ((nx + B - 3) / (B - 2)), ny, B);
      grid = dim3((unsigned)((nx + B - 3) / (B - 2)), (unsigned)((ny + rows - 1) / rows), 2);
    } else {
      B = tu.y_block;
      rows = fit_rows(tu.y_rows, 2 * ((nx + B - 1) / B), ny, B);
      grid = dim3((unsigned)((nx + B - 1) / B), (unsigned)((ny + rows - 1) / rows), 2);
    }
  }
  const fl::SweepArgs<T> a0 = make_args<T>(ctx, in[0], out[0], s[0], dtdh, ctl, nullptr,
                                           f->ndim == 1 ? 1 : rows);
  const fl::SweepArgs<T> a1 = make_args<T>(ctx, in[1], out[1], s[1], dtdh, ctl, nullptr,
                                           f->ndim == 1 ? 1 : rows);
  const bool fast = fast_numerics(s[0]);
  switch (s[0]->flux) {
    case FL_FLUX_HLLC:
      launch_flux2<T, fl::kFluxHllc>(f->ndim, dim, grid, B, a0, a1, fast, stream);
      break;
    case FL_FLUX_FORCE:
      launch_flux2<T, fl::kFluxForce>(f->ndim, dim, grid, B, a0, a1, fast, stream);
      break;
    case FL_FLUX_LAX_FRIEDRICHS:
      launch_flux2<T, fl::kFluxLaxFriedrichs>(f->ndim, dim, grid, B, a0, a1, fast, stream);
      break;
    default:
      launch_flux2<T, fl::kFluxRichtmyer>(f->ndim, dim, grid, B, a0, a1, fast, stream);
      break;
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int launch_sweep2(const fl_context* ctx, const fl_field* const in[2], const fl_field* const out[2],
                  int dim, const fl_scheme* const s[2], double dtdh, hipStream_t stream,
                  const fl::StepCtl* ctl) {
  // one launch instantiates one flux method and one form of the numerics
  if (s[0]->flux != s[1]->flux ||
      (in[0]->dtype == FL_F64 && fast_numerics(s[0]) != fast_numerics(s[1]))) {
    for (int m = 0; m < 2; ++m) {
      const int rc = launch_sweep(ctx, in[m], out[m], dim, s[m], dtdh, ctl, nullptr, stream);
      if (rc) return rc;
    }
    return FL_OK;
  }
  if (in[0]->dtype == FL_F64)
    return launch_sweep2_t<double>(ctx, in, out, dim, s, dtdh, stream, ctl);
  return launch_sweep2_t<float>(ctx, in, out, dim, s, dtdh, stream, ctl);
}

int launch_max_ws2(const fl_field* f0, const fl_field* f1, const fl_scheme* s0,
                   const fl_scheme* s1, unsigned long long* dst, hipStream_t stream) {
  int64_t es0, cs0, es1, cs1;
  strides(f0, es0, cs0);
  strides(f1, es1, cs1);
  const int64_t n = cell_count(f0);
  const int B = fl::kMaxBlock;
  int64_t blocks = (n + B - 1) / B;
  if (blocks > fl::kMaxWsBlocks) blocks = fl::kMaxWsBlocks;
  if (blocks < 1) blocks = 1;
  const fl::Material<double> m0 = make_material<double>(s0), m1 = make_material<double>(s1);
  const dim3 grid((unsigned)blocks);
  if (f0->dtype == FL_F64) {
    // the fast forms when both materials ask for the fast numerics
    const bool fast = fast_numerics(s0) && fast_numerics(s1);
    auto k = f0->ndim == 1 ? (fast ? fl::max_wavespeed2_kernel<double, 1, true>
                                   : fl::max_wavespeed2_kernel<double, 1, false>)
                           : (fast ? fl::max_wavespeed2_kernel<double, 2, true>
                                   : fl::max_wavespeed2_kernel<double, 2, false>);
    hipLaunchKernelGGL(k, grid, dim3(B), 0, stream, (const double*)f0->data,
                       (const double*)f1->data, es0, cs0, es1, cs1, n, m0, m1, dst);
  } else {
    if (f0->ndim == 1)
      hipLaunchKernelGGL((fl::max_wavespeed2_kernel<float, 1, false>), grid, dim3(B), 0, stream,
                         (const float*)f0->data, (const float*)f1->data, es0, cs0, es1, cs1,
                         n, m0, m1, dst);
    else
      hipLaunchKernelGGL((fl::max_wavespeed2_kernel<float, 2, false>), grid, dim3(B), 0, stream,
                         (const float*)f0->data, (const float*)f1->data, es0, cs0, es1, cs1,
                         n, m0, m1, dst);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int launch_max_ws(const fl_field* f, const fl_scheme* s, unsigned long long* dst,
                  hipStream_t stream) {
  int64_t es, cs;
  strides(f, es, cs);
  const int64_t n = cell_count(f);
  const fl::Material<double> m64 = make_material<double>(s);
  const int B = fl::kMaxBlock;
  int64_t blocks = (n + B - 1) / B;
  if (blocks > fl::kMaxWsBlocks) blocks = fl::kMaxWsBlocks;
  if (blocks < 1) blocks = 1;
  dim3 grid((unsigned)blocks);
  if (f->dtype == FL_F64) {
    if (f->ndim == 1)
      hipLaunchKernelGGL((fl::max_wavespeed_kernel<double, 1>), grid, dim3(B), 0, stream,
                         (const double*)f->data, es, cs, n, m64, dst);
    else
      hipLaunchKernelGGL((fl::max_wavespeed_kernel<double, 2>), grid, dim3(B), 0, stream,
                         (const double*)f->data, es, cs, n, m64, dst);
  } else {
    if (f->ndim == 1)
      hipLaunchKernelGGL((fl::max_wavespeed_kernel<float, 1>), grid, dim3(B), 0, stream,
                         (const float*)f->data, es, cs, n, m64, dst);
    else
      hipLaunchKernelGGL((fl::max_wavespeed_kernel<float, 2>), grid, dim3(B), 0, stream,
                         (const float*)f->data, es, cs, n, m64, dst);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

double bits_to_double(unsigned long long b) {
  double d;
  memcpy(&d, &b, sizeof(d));
  return d;
}

}  // namespace abi
}  // namespace fl

using namespace fl::abi;

// ---- persistent 1D step loop -------------------------------------------------

namespace {

constexpr int kPersistBlock = 64;              // lanes per workgroup: one 62-cell segment
constexpr int64_t kPersistMaxCells = 1 << 16;  // larger lines: steps outlast a launch
constexpr int64_t kPersistChunk = 1 << 16;     // steps per launch (then the host looks)
// Polls of a grid barrier before the loop gives up (each poll sleeps about
// 128 cycles after an uncached load: a fraction of a second, against a few
// microseconds per step).
constexpr unsigned int kPersistSpin = 1u << 18;

template <typename T, int FLUX>
const void* persist_kernel_ptr() {
  return reinterpret_cast<const void*>(&fl::persist1d_kernel<T, FLUX>);
}

template <typename T>
const void* persist_kernel_for(int flux) {
  switch (flux) {
    case FL_FLUX_HLLC: return persist_kernel_ptr<T, fl::kFluxHllc>();
    case FL_FLUX_FORCE: return persist_kernel_ptr<T, fl::kFluxForce>();
    case FL_FLUX_LAX_FRIEDRICHS: return persist_kernel_ptr<T, fl::kFluxLaxFriedrichs>();
    default: return persist_kernel_ptr<T, fl::kFluxRichtmyer>();
  }
}

// Workgroups of the persistent loop for a line of n cells, or 0 when it does
// not apply (long lines, fast FP64 numerics, or more segments than the
// device holds at once: a cooperative launch needs them all resident).
int persist_blocks(const fl_field* f, const fl_scheme* s) {
  if (f->ndim != 1 || f->n[0] > kPersistMaxCells) return 0;
  if (f->dtype == FL_F64 && fast_numerics(s)) return 0;
  const int64_t nseg = (f->n[0] + kPersistBlock - 3) / (kPersistBlock - 2);
  const void* k = f->dtype == FL_F64 ? persist_kernel_for<double>(s->flux)
                                     : persist_kernel_for<float>(s->flux);
  int per_cu = 0, dev = 0, cus = 0, coop = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
      !coop)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const size_t lds = fl::x_tile_lds_bytes(kPersistBlock, 3, f->dtype == FL_F64 ? 8 : 4);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kPersistBlock, lds) != hipSuccess)
    return 0;
  return (nseg <= (int64_t)per_cu * cus) ? (int)nseg : 0;
}

// fl_simulate on a short 1D line: the initial maximum, then the step loop in
// cooperative launches of up to kPersistChunk steps each.
template <typename T>
int simulate_persist1d(fl_context* c, fl_field* state, fl_field* work, const fl_scheme* s,
                       const fl_sim_params* p, int nblocks, fl_sim_result* result,
                       hipStream_t st) {
  hipLaunchKernelGGL(fl::zero_u64_kernel, dim3(1), dim3(1), 0, st, c->ws_tmp);
  int rc = launch_max_ws(state, s, c->ws_tmp, st);
  if (rc) return rc;
  hipLaunchKernelGGL(fl::persist_init_kernel, dim3(1), dim3(1), 0, st, c->pctl,
                     (const unsigned long long*)c->ws_tmp);
  FL_HIP(hipGetLastError());
  fl::SweepArgs<T> a_ab = make_args<T>(c, state, work, s, 0.0, nullptr, nullptr, 1);
  fl::SweepArgs<T> a_ba = make_args<T>(c, work, state, s, 0.0, nullptr, nullptr, 1);
  const void* k = persist_kernel_for<T>(s->flux);
  const size_t lds = fl::x_tile_lds_bytes(kPersistBlock, 3, sizeof(T));
  double cfl = p->cfl, dh = p->dh, sim_time = p->sim_time;
  long long max_iters = (long long)p->max_iters;
  unsigned int spin = kPersistSpin;
  long long k0 = 0, iters = 0;
  fl::PersistCtl h;
  memset(&h, 0, sizeof(h));
  for (;;) {
    long long nsteps = max_iters - k0;
    if (nsteps > kPersistChunk) nsteps = kPersistChunk;
    if (nsteps > 0) {
      hipLaunchKernelGGL(fl::persist_bar_reset_kernel, dim3(1), dim3(1), 0, st, c->pctl);
      void* args[] = {&a_ab, &a_ba, &c->pctl, &cfl, &dh, &sim_time, &max_iters, &k0, &nsteps,
                      &spin};
      FL_HIP(hipLaunchCooperativeKernel(k, dim3(nblocks), dim3(kPersistBlock), args, lds, st));
    }
    FL_HIP(hipMemcpyAsync(c->host_pctl, c->pctl, sizeof(fl::PersistCtl), hipMemcpyDeviceToHost,
                          st));
    FL_HIP(hipStreamSynchronize(st));
    h = *c->host_pctl;
    if (h.err)
      return fail(FL_ERR_HIP, "persistent step loop: a grid barrier timed out at step %lld",
                  (long long)(h.iters[0] > h.iters[1] ? h.iters[0] : h.iters[1]));
    iters = h.iters[0] > h.iters[1] ? h.iters[0] : h.iters[1];
    // the loop stopped early (condition false) or ran the chunk
    if (nsteps <= 0 || iters < k0 + nsteps) break;
    k0 += nsteps;
  }
  if (iters & 1) {
    const int64_t bytes = cell_count(state) * 3 * (int64_t)sizeof(T);
    FL_HIP(hipMemcpyAsync(state->data, work->data, (size_t)bytes, hipMemcpyDeviceToDevice, st));
    FL_HIP(hipStreamSynchronize(st));
  }
  const double t = __builtin_bit_cast(double, h.t[iters & 1]);
  const double final_ws = bits_to_double(h.ws[iters % 3]);
  if (result) {
    result->run_time = t;
    result->iters = iters;
    result->last_dt = iters > 0 ? h.dt : 0.0;
    result->max_wavespeed = final_ws;
  }
  if (!(final_ws > 0.0) || isnan(final_ws))
    return fail(FL_ERR_NUMERIC, "max wavespeed %g after %lld steps (NaN or vacuum)", final_ws,
                iters);
  return FL_OK;
}

}  // namespace

extern "C" {

const char* fl_status_string(int status) {
  switch (status) {
    case FL_OK: return "ok";
    case FL_ERR_ARGUMENT: return "invalid argument";
    case FL_ERR_UNSUPPORTED: return "unsupported";
    case FL_ERR_HIP: return "HIP runtime error";
    case FL_ERR_NUMERIC: return "numeric failure";
    default: return "unknown status";
  }
}

const char* fl_last_error(void) { return g_last_error.c_str(); }

const char* fl_version(void) { return "fluidity_amd 0.1 (gfx950)"; }

int fl_ctx_create(int device, fl_context** out) {
  if (!out) return fail(FL_ERR_ARGUMENT, "ctx out-pointer is NULL");
  *out = nullptr;
  int count = 0;
  FL_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(FL_ERR_ARGUMENT, "device %d out of range (%d devices)", device, count);
  DeviceGuard g(device);
  fl_context* c = new fl_context();
  memset(c, 0, sizeof(*c));
  c->device = device;
  (void)resolve_tuning(nullptr, &c->tuning);
  hipError_t e = hipMalloc((void**)&c->ctl, sizeof(fl::StepCtl));
  if (e == hipSuccess) e = hipMalloc((void**)&c->ws_tmp, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void**)&c->partial, sizeof(double) * kSumBlocks * 4);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_ctl, sizeof(fl::StepCtl), 0);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_partial, sizeof(double) * kSumBlocks * 4, 0);
  if (e == hipSuccess) e = hipStreamCreate(&c->cap_stream);
  c->use_graphs = 1;
  if (e == hipSuccess) e = hipMalloc((void**)&c->eik, 2 * sizeof(fl::ls::EikCtl));
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_eik, 2 * sizeof(fl::ls::EikCtl), 0);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->step_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc((void**)&c->pctl, sizeof(fl::PersistCtl));
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_pctl, sizeof(fl::PersistCtl), 0);
  if (e != hipSuccess) {
    fl_ctx_destroy(c);
    return fail(FL_ERR_HIP, "context allocation failed: %s", hipGetErrorString(e));
  }
  *out = c;
  return FL_OK;
}

int fl_ctx_destroy(fl_context* c) {
  if (!c) return FL_OK;
  DeviceGuard g(c->device);
  if (c->ctl) (void)hipFree(c->ctl);
  if (c->ws_tmp) (void)hipFree(c->ws_tmp);
  if (c->partial) (void)hipFree(c->partial);
  if (c->host_ctl) (void)hipHostFree(c->host_ctl);
  if (c->host_partial) (void)hipHostFree(c->host_partial);
  if (c->sim_exec) (void)hipGraphExecDestroy(c->sim_exec);
  if (c->sim_graph) (void)hipGraphDestroy(c->sim_graph);
  if (c->sim_exec_run) (void)hipGraphExecDestroy(c->sim_exec_run);
  if (c->sim_graph_run) (void)hipGraphDestroy(c->sim_graph_run);
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  if (c->eik) (void)hipFree(c->eik);
  if (c->host_eik) (void)hipHostFree(c->host_eik);
  if (c->step_ev) (void)hipEventDestroy(c->step_ev);
  if (c->pctl) (void)hipFree(c->pctl);
  if (c->host_pctl) (void)hipHostFree(c->host_pctl);
  delete c;
  return FL_OK;
}

int fl_ctx_set_tuning(fl_context* c, const fl_tuning* t) {
  if (!c || !t) return fail(FL_ERR_ARGUMENT, "NULL context or tuning");
  fl_tuning n = c->tuning;
  if (t->x_block) n.x_block = t->x_block;
  if (t->x_rows) n.x_rows = t->x_rows;
  if (t->y_block) n.y_block = t->y_block;
  if (t->y_rows) n.y_rows = t->y_rows;
  int rc = resolve_tuning(&n, &n);
  if (rc) return rc;
  c->tuning = n;
  return FL_OK;
}

int fl_ctx_get_tuning(const fl_context* c, fl_tuning* t) {
  if (!c || !t) return fail(FL_ERR_ARGUMENT, "NULL context or tuning");
  *t = c->tuning;
  return FL_OK;
}

int fl_ctx_autotune(fl_context* c, const fl_field* in, const fl_field* work,
                    const fl_scheme* s, int reps, fl_tuning* chosen, void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  int rc = check_pair(in, work);
  if (rc) return rc;
  if ((rc = check_scheme(s, in->ndim))) return rc;
  if (reps < 1 || reps > 100) return fail(FL_ERR_ARGUMENT, "reps %d out of [1, 100]", reps);
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  static const int widths[] = {64, 128, 256};
  static const int xrows[] = {8, 16, 32, 64};
  static const int yrows[] = {16, 32, 64, 128};
  int cus = 0;
  FL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  hipEvent_t e0, e1;
  FL_HIP(hipEventCreate(&e0));
  FL_HIP(hipEventCreate(&e1));
  const fl_tuning saved = c->tuning;
  fl_tuning best = c->tuning;
  for (int dim = 0; dim < in->ndim && !rc; ++dim) {
    float best_ms = 1e30f;
    for (int w : widths) {
      std::vector<int> rows(1, 0);  // 1D: the height is unused
      if (in->ndim == 2) {
        const int64_t nx = in->n[0], ny = in->n[1];
        const int64_t bands = dim ? (nx + w - 1) / w : (nx + w - 3) / (w - 2);
        const int per_cu = in->dtype == FL_F64 ? sweep_blocks_per_cu_t<double>(dim, s, w)
                                               : sweep_blocks_per_cu_t<float>(dim, s, w);
        rows = row_candidates(dim ? yrows : xrows, 4, bands, ny, (int64_t)per_cu * cus);
      }
      for (size_t k = 0; k < rows.size() && !rc; ++k) {
        fl_tuning t = best;
        if (dim == 0) {
          t.x_block = w;
          if (rows[k]) t.x_rows = rows[k];
        } else {
          t.y_block = w;
          t.y_rows = rows[k];
        }
        c->tuning = t;
        float total = 0.0f;
        for (int r = 0; r <= reps && !rc; ++r) {  // r = 0 warms up
          if (hipEventRecord(e0, st) != hipSuccess) rc = fail(FL_ERR_HIP, "hipEventRecord");
          if (!rc) rc = launch_sweep(c, in, work, dim, s, 1e-3, nullptr, nullptr, st);
          if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = fail(FL_ERR_HIP, "hipEventRecord");
          if (!rc && hipEventSynchronize(e1) != hipSuccess)
            rc = fail(FL_ERR_HIP, "hipEventSynchronize");
          float ms = 0.0f;
          if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
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
  c->tuning = rc ? saved : best;
  if (!rc && chosen) *chosen = best;
  return rc;
}

int fl_split_sweep(fl_context* c, const fl_field* in, const fl_field* out, int dim,
                   const fl_scheme* s, double dtdh, void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  int rc = check_pair(in, out);
  if (rc) return rc;
  rc = check_scheme(s, in->ndim);
  if (rc) return rc;
  if (dim < 0 || dim >= in->ndim) return fail(FL_ERR_ARGUMENT, "dim %d out of range", dim);
  if (!(dtdh > 0.0) || !isfinite(dtdh)) return fail(FL_ERR_ARGUMENT, "dtdh %g must be > 0", dtdh);
  DeviceGuard g(c->device);
  return launch_sweep(c, in, out, dim, s, dtdh, nullptr, nullptr, (hipStream_t)stream);
}

int fl_split_solve(fl_context* c, fl_field* in, fl_field* out, const fl_scheme* s, double dtdh,
                   void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  int rc = check_pair(in, out);
  if (rc) return rc;
  rc = check_scheme(s, in->ndim);
  if (rc) return rc;
  if (!(dtdh > 0.0) || !isfinite(dtdh)) return fail(FL_ERR_ARGUMENT, "dtdh %g must be > 0", dtdh);
  DeviceGuard g(c->device);
  for (int d = 0; d < in->ndim; ++d) {
    rc = launch_sweep(c, in, out, d, s, dtdh, nullptr, nullptr, (hipStream_t)stream);
    if (rc) return rc;
    if (in->ndim != 1 && d < in->ndim - 1) {  // solver/cuda/split_solver.cuh:75-78
      void* tmp = in->data;
      in->data = out->data;
      out->data = tmp;
    }
  }
  return FL_OK;
}

int fl_split_step(fl_context* c, fl_field* in, fl_field* out, const fl_scheme* s, double dtdh,
                  double* max_wavespeed, void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  int rc = check_pair(in, out);
  if (rc) return rc;
  rc = check_scheme(s, in->ndim);
  if (rc) return rc;
  if (!(dtdh > 0.0) || !isfinite(dtdh)) return fail(FL_ERR_ARGUMENT, "dtdh %g must be > 0", dtdh);
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  if (max_wavespeed) hipLaunchKernelGGL(fl::zero_u64_kernel, dim3(1), dim3(1), 0, st, c->ws_tmp);
  for (int d = 0; d < in->ndim; ++d) {
    // the last sweep writes the final states and reduces their wavespeeds
    unsigned long long* ws = (max_wavespeed && d == in->ndim - 1) ? c->ws_tmp : nullptr;
    rc = launch_sweep(c, in, out, d, s, dtdh, nullptr, ws, st);
    if (rc) return rc;
    if (in->ndim != 1 && d < in->ndim - 1) {  // solver/cuda/split_solver.cuh:75-78
      void* tmp = in->data;
      in->data = out->data;
      out->data = tmp;
    }
  }
  if (max_wavespeed) {
    FL_HIP(hipMemcpyAsync(&c->host_ctl->ws_bits[0], c->ws_tmp, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    *max_wavespeed = bits_to_double(c->host_ctl->ws_bits[0]);
  }
  return FL_OK;
}

int fl_max_wavespeed(fl_context* c, const fl_field* in, const fl_scheme* s, double* result,
                     void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  if (!result) return fail(FL_ERR_ARGUMENT, "result is NULL");
  int rc = check_field(in, "in");
  if (rc) return rc;
  rc = check_scheme(s, in->ndim);
  if (rc) return rc;
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fl::zero_u64_kernel, dim3(1), dim3(1), 0, st, c->ws_tmp);
  rc = launch_max_ws(in, s, c->ws_tmp, st);
  if (rc) return rc;
  FL_HIP(hipMemcpyAsync(&c->host_ctl->ws_bits[0], c->ws_tmp, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, st));
  FL_HIP(hipStreamSynchronize(st));
  *result = bits_to_double(c->host_ctl->ws_bits[0]);
  return FL_OK;
}

int fl_set_wavespeeds(fl_context* c, const fl_field* in, const fl_scheme* s, double* ws,
                      void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  if (!ws) return fail(FL_ERR_ARGUMENT, "wavespeeds is NULL");
  int rc = check_field(in, "in");
  if (rc) return rc;
  rc = check_scheme(s, in->ndim);
  if (rc) return rc;
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  int64_t es, cs;
  strides(in, es, cs);
  const int64_t n = cell_count(in);
  const fl::Material<double> m64 = make_material<double>(s);
  const int B = fl::kMaxBlock;
  int64_t blocks = (n + B - 1) / B;
  if (blocks > 8192) blocks = 8192;
  const dim3 grid((unsigned)blocks);
  if (in->dtype == FL_F64) {
    if (in->ndim == 1)
      hipLaunchKernelGGL((fl::wavespeeds_kernel<double, 1>), grid, dim3(B), 0, st,
                         (const double*)in->data, es, cs, n, m64, ws);
    else
      hipLaunchKernelGGL((fl::wavespeeds_kernel<double, 2>), grid, dim3(B), 0, st,
                         (const double*)in->data, es, cs, n, m64, ws);
  } else {
    if (in->ndim == 1)
      hipLaunchKernelGGL((fl::wavespeeds_kernel<float, 1>), grid, dim3(B), 0, st,
                         (const float*)in->data, es, cs, n, m64, ws);
    else
      hipLaunchKernelGGL((fl::wavespeeds_kernel<float, 2>), grid, dim3(B), 0, st,
                         (const float*)in->data, es, cs, n, m64, ws);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int fl_max_element(fl_context* c, const double* first, int64_t count, double* result,
                   void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  if (!first || !result) return fail(FL_ERR_ARGUMENT, "first or result is NULL");
  if (count <= 0) return fail(FL_ERR_ARGUMENT, "count must be > 0 (got %lld)", (long long)count);
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  const int B = fl::kMaxBlock;
  int64_t blocks = (count + B - 1) / B;
  if (blocks > fl::kMaxWsBlocks) blocks = fl::kMaxWsBlocks;
  hipLaunchKernelGGL(fl::zero_u64_kernel, dim3(1), dim3(1), 0, st, c->ws_tmp);
  hipLaunchKernelGGL(fl::max_element_kernel, dim3((unsigned)blocks), dim3(B), 0, st, first, count,
                     c->ws_tmp);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(&c->host_ctl->ws_bits[0], c->ws_tmp, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, st));
  FL_HIP(hipStreamSynchronize(st));
  const unsigned long long k = c->host_ctl->ws_bits[0];
  if (k == ~0ull) {
    *result = NAN;
  } else {
    // invert the order-preserving key of the max_element kernel
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    *result = bits_to_double(b);
  }
  return FL_OK;
}

int fl_ctx_set_graphs(fl_context* c, int enable) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  c->use_graphs = enable ? 1 : 0;
  return FL_OK;
}

int fl_simulate(fl_context* c, fl_field* state, fl_field* work, const fl_scheme* s,
                const fl_sim_params* p, fl_sim_result* result, void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  int rc = check_pair(state, work);
  if (rc) return rc;
  rc = check_scheme(s, state->ndim);
  if (rc) return rc;
  if (!p) return fail(FL_ERR_ARGUMENT, "params is NULL");
  if (!(p->cfl > 0.0) || !isfinite(p->cfl)) return fail(FL_ERR_ARGUMENT, "cfl must be > 0");
  if (!(p->dh > 0.0) || !isfinite(p->dh)) return fail(FL_ERR_ARGUMENT, "dh must be > 0");
  if (p->max_iters < 0) return fail(FL_ERR_ARGUMENT, "max_iters must be >= 0");
  if (isnan(p->sim_time)) return fail(FL_ERR_ARGUMENT, "sim_time is NaN");
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;

  if (c->use_graphs) {
    const int nb = persist_blocks(state, s);
    if (nb > 0) {
      if (state->dtype == FL_F64)
        return simulate_persist1d<double>(c, state, work, s, p, nb, result, st);
      return simulate_persist1d<float>(c, state, work, s, p, nb, result, st);
    }
  }

  hipLaunchKernelGGL(fl::init_control_kernel, dim3(1), dim3(1), 0, st, c->ctl);
  rc = launch_max_ws(state, s, &c->ctl->ws_bits[0], st);
  if (rc) return rc;

  // Steps are launched in batches; between batches the host reads the control
  // block once to stop when the loop condition fails.  With graphs on, runs
  // of kSimGraphSteps steps and pairs of steps (both parities) are captured
  // once and replayed, the runs while a batch has that many steps left (one
  // host launch per run instead of per pair: small grids' steps are shorter
  // than a graph launch); a step past the stop condition does nothing, so an
  // odd batch may run one step more.
  const int64_t kBatch = 256;
  constexpr int kSimGraphSteps = 16;
  int64_t launched = 0;
  fl_field a = *state, b = *work;
  auto one_step = [&](int parity, hipStream_t sx) -> int {
    hipLaunchKernelGGL(fl::step_control_kernel, dim3(1), dim3(1), 0, sx, c->ctl, parity,
                       p->cfl, p->dh, p->sim_time, (long long)p->max_iters);
    unsigned long long* ws_next = &c->ctl->ws_bits[parity ^ 1];
    if (state->ndim == 1) {
      const fl_field* src = parity ? &b : &a;
      const fl_field* dst = parity ? &a : &b;
      return launch_sweep(c, src, dst, 0, s, 1.0, c->ctl, ws_next, sx);
    }
    int r = launch_sweep(c, &a, &b, 0, s, 1.0, c->ctl, nullptr, sx);
    if (!r) r = launch_sweep(c, &b, &a, 1, s, 1.0, c->ctl, ws_next, sx);
    return r;
  };
  if (c->use_graphs) {
    SimGraphKey key;
    memset(&key, 0, sizeof(key));
    key.a = state->data;
    key.b = work->data;
    key.dtype = state->dtype;
    key.layout = state->layout;
    key.ndim = state->ndim;
    for (int d = 0; d < 3; ++d) key.n[d] = state->n[d];
    key.scheme = *s;
    key.cfl = p->cfl;
    key.dh = p->dh;
    key.sim_time = p->sim_time;
    key.max_iters = p->max_iters;
    key.tuning = c->tuning;
    if (!c->sim_exec || !c->sim_exec_run || memcmp(&key, &c->sim_key, sizeof(key)) != 0) {
      if (c->sim_exec) (void)hipGraphExecDestroy(c->sim_exec);
      if (c->sim_graph) (void)hipGraphDestroy(c->sim_graph);
      if (c->sim_exec_run) (void)hipGraphExecDestroy(c->sim_exec_run);
      if (c->sim_graph_run) (void)hipGraphDestroy(c->sim_graph_run);
      c->sim_exec = c->sim_exec_run = nullptr;
      c->sim_graph = c->sim_graph_run = nullptr;
      // captures `steps` steps (even) into *graph and instantiates *exec
      auto capture = [&](int steps, hipGraph_t* graph, hipGraphExec_t* exec) -> int {
        FL_HIP(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
        int r = FL_OK;
        for (int k = 0; k < steps && !r; ++k) r = one_step(k & 1, c->cap_stream);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->cap_stream, &gr);
        if (r) {
          if (gr) (void)hipGraphDestroy(gr);
          return r;
        }
        if (e != hipSuccess)
          return fail(FL_ERR_HIP, "step graph capture failed: %s", hipGetErrorString(e));
        *graph = gr;
        FL_HIP(hipGraphInstantiate(exec, gr, nullptr, nullptr, 0));
        return FL_OK;
      };
      if ((rc = capture(2, &c->sim_graph, &c->sim_exec))) return rc;
      if ((rc = capture(kSimGraphSteps, &c->sim_graph_run, &c->sim_exec_run))) return rc;
      c->sim_key = key;
    }
  }
  while (launched < p->max_iters) {
    const int64_t nb = (p->max_iters - launched) < kBatch ? (p->max_iters - launched) : kBatch;
    if (c->use_graphs) {
      int64_t k = 0;
      for (; k + kSimGraphSteps <= nb; k += kSimGraphSteps)
        FL_HIP(hipGraphLaunch(c->sim_exec_run, st));
      for (; k < nb; k += 2) FL_HIP(hipGraphLaunch(c->sim_exec, st));
    } else {
      for (int64_t k = 0; k < nb; ++k) {
        rc = one_step((int)((launched + k) & 1), st);
        if (rc) return rc;
      }
    }
    launched += nb;
    FL_HIP(hipMemcpyAsync(c->host_ctl, c->ctl, sizeof(fl::StepCtl), hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    if (!c->host_ctl->active) break;
  }
  if (launched == 0) {
    // max_iters == 0: no step ran; the result is the freshly initialised
    // control block (t = 0, no iterations) with the initial max wavespeed.
    FL_HIP(hipMemcpyAsync(c->host_ctl, c->ctl, sizeof(fl::StepCtl), hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
  }
  const fl::StepCtl hc = *c->host_ctl;
  if (state->ndim == 1 && (hc.iters & 1)) {
    const int64_t bytes = cell_count(state) * 3 * (state->dtype == FL_F64 ? 8 : 4);
    FL_HIP(hipMemcpyAsync(state->data, work->data, (size_t)bytes, hipMemcpyDeviceToDevice, st));
    FL_HIP(hipStreamSynchronize(st));
  }
  const double final_ws = bits_to_double(hc.ws_bits[hc.iters & 1]);
  if (result) {
    result->run_time = hc.t;
    result->iters = hc.iters;
    result->last_dt = hc.dt;
    result->max_wavespeed = final_ws;
  }
  if (!(final_ws > 0.0) || isnan(final_ws))
    return fail(FL_ERR_NUMERIC, "max wavespeed %g after %lld steps (NaN or vacuum)", final_ws,
                (long long)hc.iters);
  return FL_OK;
}

int fl_conservation_sums(fl_context* c, const fl_field* in, double* sums, void* stream) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  if (!sums) return fail(FL_ERR_ARGUMENT, "sums is NULL");
  int rc = check_field(in, "in");
  if (rc) return rc;
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  int64_t es, cs;
  strides(in, es, cs);
  const int64_t n = cell_count(in);
  const int S = 2 + in->ndim;
  int64_t blocks = (n + 255) / 256;
  if (blocks > kSumBlocks) blocks = kSumBlocks;
  dim3 grid((unsigned)blocks);
  if (in->dtype == FL_F64) {
    if (S == 3)
      hipLaunchKernelGGL((fl::component_sums_kernel<double, 3>), grid, dim3(256), 0, st,
                         (const double*)in->data, es, cs, n, c->partial);
    else
      hipLaunchKernelGGL((fl::component_sums_kernel<double, 4>), grid, dim3(256), 0, st,
                         (const double*)in->data, es, cs, n, c->partial);
  } else {
    if (S == 3)
      hipLaunchKernelGGL((fl::component_sums_kernel<float, 3>), grid, dim3(256), 0, st,
                         (const float*)in->data, es, cs, n, c->partial);
    else
      hipLaunchKernelGGL((fl::component_sums_kernel<float, 4>), grid, dim3(256), 0, st,
                         (const float*)in->data, es, cs, n, c->partial);
  }
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(c->host_partial, c->partial, sizeof(double) * blocks * S,
                        hipMemcpyDeviceToHost, st));
  FL_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < S; ++k) {
    double v = 0.0;
    for (int64_t b = 0; b < blocks; ++b) v += c->host_partial[b * S + k];
    sums[k] = v;
  }
  return FL_OK;
}

}  // extern "C"
