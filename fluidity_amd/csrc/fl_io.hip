// fl_io.hip -- initial conditions and result output of a simulator: the
// per-cell primitive <-> conservative conversions on the device, the
// level-set-masked state setter of the two-material simulator, and the
// reference's ASCII result formats written on the host.
#include <hip/hip_runtime.h>

#include <stdio.h>

#include <string>
#include <vector>

#include "fl_abi_internal.hpp"
#include "fl_numerics.hpp"

namespace fl {
namespace io {

// Primitive planes of conservative states: rho, p, v_0.., and the material's
// equation of state of that primitive state.  primitive()
// (state/state_impl.hpp:339-357: p from the conservative pressure, which uses
// the adiabatic index only, v_d = m_d / rho) and IdealGas::eos
// (material/ideal_gas.hpp:60-63) / StiffFluid::eos
// (material/stiff_fluid.hpp:77-80).
template <typename T, int ND>
__global__ void __launch_bounds__(256) primitive_planes_kernel(const T* __restrict__ in,
                                                              int64_t es, int64_t cs,
                                                              int64_t n, Material<T> m,
                                                              double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T u[2 + ND];
#pragma unroll
  for (int k = 0; k < 2 + ND; ++k) u[k] = in[i * es + k * cs];
  const Prim<T, ND> w = prim_of<T, ND>(u, m);
  out[i] = (double)u[0];
  out[n + i] = (double)w.p;
#pragma unroll
  for (int d = 0; d < ND; ++d) out[(2 + d) * n + i] = (double)w.vel[d];
  const T num = m.stiff ? w.p + m.gamma * m.p_inf : w.p;
  out[(2 + ND) * n + i] = (double)dv(num, m.gm1 * u[0]);
}

// Conservative states from primitive planes (rho, p, v_0..): conservative()
// of a primitive state (state/state_impl.hpp:388-406) with energy
// rho (1/2 |v|^2 + eos) (:135-142).
template <typename T, int ND>
__global__ void __launch_bounds__(256) from_primitive_kernel(const double* __restrict__ prim,
                                                            int64_t n, Material<T> m,
                                                            T* __restrict__ out, int64_t es,
                                                            int64_t cs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T rho = (T)prim[i];
  const T p = (T)prim[n + i];
  T v[ND];
  T vss = T(0);
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    v[d] = (T)prim[(2 + d) * n + i];
    vss += v[d] * v[d];
  }
  const T num = m.stiff ? p + m.gamma * m.p_inf : p;
  const T eos = dv(num, m.gm1 * rho);
  out[i * es] = rho;
  out[i * es + cs] = rho * (T(0.5) * vss + eos);
#pragma unroll
  for (int d = 0; d < ND; ++d) out[i * es + (2 + d) * cs] = v[d] * rho;
}

// set_states_impl (state/cuda/state_initialization.cuh:72-87): every cell
// inside the level set (phi <= 0, levelset/levelset.hpp:166-170) takes the
// conservative form of one primitive state; the others are left as they are.
template <typename T, int ND>
__global__ void __launch_bounds__(256) set_states_kernel(const double* __restrict__ phi,
                                                        int64_t n, Cell<T, ND> c,
                                                        T* __restrict__ out, int64_t es,
                                                        int64_t cs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!(phi[i] <= 0.0)) return;
#pragma unroll
  for (int k = 0; k < 2 + ND; ++k) out[i * es + k * cs] = c.u[k];
}

}  // namespace io
}  // namespace fl

using namespace fl::abi;

namespace {

int64_t blocks_for(int64_t n) { return (n + 255) / 256; }

int check_io_field(const fl_field* f, const char* name) {
  int rc = check_field(f, name);
  if (rc) return rc;
  if (cell_count(f) > (int64_t)0x7fffffff * 256)
    return fail(FL_ERR_ARGUMENT, "%s has too many cells", name);
  return FL_OK;
}

template <typename T>
int launch_primitive(const fl_field* f, const fl_scheme* s, double* out, hipStream_t st) {
  int64_t es, cs;
  strides(f, es, cs);
  const int64_t n = cell_count(f);
  const fl::Material<T> m = make_material<T>(s);
  const dim3 grid((unsigned)blocks_for(n));
  if (f->ndim == 1)
    hipLaunchKernelGGL((fl::io::primitive_planes_kernel<T, 1>), grid, dim3(256), 0, st,
                       (const T*)f->data, es, cs, n, m, out);
  else
    hipLaunchKernelGGL((fl::io::primitive_planes_kernel<T, 2>), grid, dim3(256), 0, st,
                       (const T*)f->data, es, cs, n, m, out);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

template <typename T>
int launch_from_primitive(const double* prim, const fl_scheme* s, const fl_field* f,
                          hipStream_t st) {
  int64_t es, cs;
  strides(f, es, cs);
  const int64_t n = cell_count(f);
  const fl::Material<T> m = make_material<T>(s);
  const dim3 grid((unsigned)blocks_for(n));
  if (f->ndim == 1)
    hipLaunchKernelGGL((fl::io::from_primitive_kernel<T, 1>), grid, dim3(256), 0, st, prim, n,
                       m, (T*)f->data, es, cs);
  else
    hipLaunchKernelGGL((fl::io::from_primitive_kernel<T, 2>), grid, dim3(256), 0, st, prim, n,
                       m, (T*)f->data, es, cs);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// The conservative state of one primitive state, computed on the host with
// the device kernel's expressions (the same IEEE operations in T).
template <typename T, int ND>
fl::Cell<T, ND> cons_of_prim(const double* w, const fl_scheme* s) {
  const fl::Material<T> m = make_material<T>(s);
  const T rho = (T)w[0], p = (T)w[1];
  T v[ND];
  T vss = T(0);
  for (int d = 0; d < ND; ++d) {
    v[d] = (T)w[2 + d];
    vss += v[d] * v[d];
  }
  const T num = m.stiff ? p + m.gamma * m.p_inf : p;
  const T eos = num / (m.gm1 * rho);
  fl::Cell<T, ND> c;
  c.u[0] = rho;
  c.u[1] = rho * (T(0.5) * vss + eos);
  for (int d = 0; d < ND; ++d) c.u[2 + d] = v[d] * rho;
  return c;
}

template <typename T>
int launch_set_states(const double* phi, const double* w, const fl_scheme* s, const fl_field* f,
                      hipStream_t st) {
  int64_t es, cs;
  strides(f, es, cs);
  const int64_t n = cell_count(f);
  const dim3 grid((unsigned)blocks_for(n));
  if (f->ndim == 1)
    hipLaunchKernelGGL((fl::io::set_states_kernel<T, 1>), grid, dim3(256), 0, st, phi, n,
                       cons_of_prim<T, 1>(w, s), (T*)f->data, es, cs);
  else
    hipLaunchKernelGGL((fl::io::set_states_kernel<T, 2>), grid, dim3(256), 0, st, phi, n,
                       cons_of_prim<T, 2>(w, s), (T*)f->data, es, cs);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Host copy of the primitive planes (and of phi when given) of a field.
int fetch_planes(const fl_field* f, const fl_scheme* s, const double* phi, hipStream_t st,
                 std::vector<double>& planes, std::vector<double>& ls) {
  const int64_t n = cell_count(f);
  const int P = 3 + f->ndim;
  double* dev = nullptr;
  FL_HIP(hipMalloc((void**)&dev, sizeof(double) * P * n));
  int rc = f->dtype == FL_F64 ? launch_primitive<double>(f, s, dev, st)
                              : launch_primitive<float>(f, s, dev, st);
  planes.resize((size_t)P * n);
  if (!rc && hipStreamSynchronize(st) != hipSuccess)
    rc = fail(FL_ERR_HIP, "hipStreamSynchronize failed");
  if (!rc && hipMemcpy(planes.data(), dev, sizeof(double) * P * n, hipMemcpyDeviceToHost) !=
                 hipSuccess)
    rc = fail(FL_ERR_HIP, "hipMemcpy of the primitive planes failed");
  (void)hipFree(dev);
  if (!rc && phi) {
    ls.resize((size_t)n);
    if (hipMemcpy(ls.data(), phi, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(FL_ERR_HIP, "hipMemcpy of phi failed");
  }
  return rc;
}

const char* const kElementNames[] = {"rho", "p", "v_x", "v_y"};

// One value as the reference streams it: std::setw(12) << std::left <<
// std::fixed << std::showpoint << std::setprecision(8) << value << " ".
void put_fixed(FILE* out, double v) { fprintf(out, "%-12.8f ", v); }

// stream_output_ascii (simulator/simulator_impl.hpp:435-508); with phi, the
// two-material variant (simulator/multimaterial_simulator.hpp:882-971),
// which adds one level-set column per dimension.
void write_ascii(FILE* out, const fl_field* f, const std::vector<double>& planes,
                 const std::vector<double>& ls, bool with_ls, double run_time,
                 double resolution) {
  const int nd = f->ndim;
  const int64_t nx = f->n[0], ny = nd > 1 ? f->n[1] : 1, n = nx * ny;
  const int S = 2 + nd;
  fprintf(out, "# t = %f\n", run_time);
  int column = 1;
  for (int d = 0; d < nd; ++d) fprintf(out, "# Column %d: %c coordinate\n", column++, 'x' + d);
  if (with_ls)
    for (int d = 0; d < nd; ++d)
      fprintf(out, "# Column %d: %c ls coordinate\n", column++, 'x' + d);
  for (int k = 0; k < S; ++k) fprintf(out, "# Column %d: %s\n", column++, kElementNames[k]);
  fprintf(out, with_ls ? "# Column %d: internal energy (e)\n"
                       : "# Column %d : internal energy (e)\n",
          column);
  for (int64_t y = 0; y < ny; ++y) {
    for (int64_t x = 0; x < nx; ++x) {
      const int64_t i = y * nx + x;
      put_fixed(out, ((double)x + 0.5) * resolution);
      if (nd > 1) put_fixed(out, ((double)y + 0.5) * resolution);
      if (with_ls)
        for (int d = 0; d < nd; ++d) put_fixed(out, ls[i]);
      for (int k = 0; k <= S; ++k) put_fixed(out, planes[(size_t)k * n + i]);
      fputc('\n', out);
    }
  }
}

}  // namespace

extern "C" {

int fl_primitive_planes(fl_context* ctx, const fl_field* state, const fl_scheme* s, double* out,
                        void* stream) {
  if (!ctx) return fail(FL_ERR_ARGUMENT, "ctx is NULL");
  int rc = check_io_field(state, "state");
  if (rc) return rc;
  if ((rc = check_scheme(s, state->ndim))) return rc;
  if (!out) return fail(FL_ERR_ARGUMENT, "out is NULL");
  DeviceGuard g(ctx->device);
  const hipStream_t st = (hipStream_t)stream;
  return state->dtype == FL_F64 ? launch_primitive<double>(state, s, out, st)
                                : launch_primitive<float>(state, s, out, st);
}

int fl_from_primitive(fl_context* ctx, const double* prim, const fl_scheme* s, fl_field* state,
                      void* stream) {
  if (!ctx) return fail(FL_ERR_ARGUMENT, "ctx is NULL");
  int rc = check_io_field(state, "state");
  if (rc) return rc;
  if ((rc = check_scheme(s, state->ndim))) return rc;
  if (!prim) return fail(FL_ERR_ARGUMENT, "prim is NULL");
  DeviceGuard g(ctx->device);
  const hipStream_t st = (hipStream_t)stream;
  return state->dtype == FL_F64 ? launch_from_primitive<double>(prim, s, state, st)
                                : launch_from_primitive<float>(prim, s, state, st);
}

int fl_set_states(fl_context* ctx, const double* phi, const double* prim_state,
                  const fl_scheme* s, fl_field* state, void* stream) {
  if (!ctx) return fail(FL_ERR_ARGUMENT, "ctx is NULL");
  int rc = check_io_field(state, "state");
  if (rc) return rc;
  if ((rc = check_scheme(s, state->ndim))) return rc;
  if (!phi || !prim_state) return fail(FL_ERR_ARGUMENT, "phi or prim_state is NULL");
  DeviceGuard g(ctx->device);
  const hipStream_t st = (hipStream_t)stream;
  return state->dtype == FL_F64 ? launch_set_states<double>(phi, prim_state, s, state, st)
                                : launch_set_states<float>(phi, prim_state, s, state, st);
}

int fl_write_results(fl_context* ctx, const fl_field* state, const fl_scheme* s,
                     const double* phi, double run_time, double resolution, const char* path,
                     void* stream) {
  if (!ctx) return fail(FL_ERR_ARGUMENT, "ctx is NULL");
  int rc = check_io_field(state, "state");
  if (rc) return rc;
  if ((rc = check_scheme(s, state->ndim))) return rc;
  DeviceGuard g(ctx->device);
  std::vector<double> planes, ls;
  if ((rc = fetch_planes(state, s, phi, (hipStream_t)stream, planes, ls))) return rc;
  FILE* out = path ? fopen(path, "w") : stdout;
  if (!out) return fail(FL_ERR_ARGUMENT, "cannot open %s for writing", path);
  write_ascii(out, state, planes, ls, phi != nullptr, run_time, resolution);
  const bool bad = ferror(out) != 0;
  if (path) {
    if (fclose(out) != 0 || bad) return fail(FL_ERR_ARGUMENT, "writing %s failed", path);
  } else {
    fflush(out);
    if (bad) return fail(FL_ERR_ARGUMENT, "writing to stdout failed");
  }
  return FL_OK;
}

int fl_write_results_separate_raw(fl_context* ctx, const fl_field* state, const fl_scheme* s,
                                  double sim_time, const char* prefix, void* stream) {
  if (!ctx) return fail(FL_ERR_ARGUMENT, "ctx is NULL");
  if (!prefix) return fail(FL_ERR_ARGUMENT, "prefix is NULL");
  int rc = check_io_field(state, "state");
  if (rc) return rc;
  if ((rc = check_scheme(s, state->ndim))) return rc;
  DeviceGuard g(ctx->device);
  std::vector<double> planes, ls;
  if ((rc = fetch_planes(state, s, nullptr, (hipStream_t)stream, planes, ls))) return rc;
  const int nd = state->ndim;
  const int64_t nx = state->n[0], ny = nd > 1 ? state->n[1] : 1, n = nx * ny;
  char stamp[64];
  snprintf(stamp, sizeof(stamp), "%f", sim_time);  // std::to_string(double)
  for (int k = 0; k < 2 + nd; ++k) {
    const std::string name =
        std::string(prefix) + "_" + kElementNames[k] + "_" + stamp + ".dat";
    FILE* out = fopen(name.c_str(), "w");
    if (!out) return fail(FL_ERR_ARGUMENT, "cannot open %s for writing", name.c_str());
    // write_blob (simulator/simulator_impl.hpp:360-405): default stream
    // formatting (%g), one line per row; a 1D grid gets one more newline.
    for (int64_t y = 0; y < ny; ++y) {
      for (int64_t x = 0; x < nx; ++x) fprintf(out, "%g ", planes[(size_t)k * n + y * nx + x]);
      fputc('\n', out);
    }
    if (nd == 1) fputc('\n', out);
    const bool bad = ferror(out) != 0;
    if (fclose(out) != 0 || bad) return fail(FL_ERR_ARGUMENT, "writing %s failed", name.c_str());
  }
  return FL_OK;
}

}  // extern "C"
