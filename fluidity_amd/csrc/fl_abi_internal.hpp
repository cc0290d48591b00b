// fl_abi_internal.hpp -- pieces of the C ABI host code shared by fl_abi.hip
// (split-solver entry points) and fl_ls_abi.hip (level-set / ghost-fluid entry
// points): the context, error reporting and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "../../include/fluidity_amd.h"
#include "fl_kernels.hpp"
#include "fl_levelset.hpp"

// The sweep kernels with van Leer in conservative form fixed at compile time
// run where a scheme asks for them (tuning builds: FL_VLC_KERNELS=0 always
// launches the generic kernels).
#ifndef FL_VLC_KERNELS
#define FL_VLC_KERNELS 1
#endif

// What a captured fl_simulate step pair depends on (compared bytewise).
struct SimGraphKey {
  void* a;
  void* b;
  int32_t dtype, layout, ndim, pad;
  int64_t n[3];
  fl_scheme scheme;
  double cfl, dh, sim_time;
  int64_t max_iters;
  fl_tuning tuning;
};

struct fl_context {
  int device;
  fl::StepCtl* ctl;            // device control block of fl_simulate
  unsigned long long* ws_tmp;  // device scratch for fl_max_wavespeed
  double* partial;             // device partial sums, kSumBlocks * 4
  fl::StepCtl* host_ctl;       // pinned host mirror of ctl
  double* host_partial;        // pinned host mirror of partial
  fl::ls::EikCtl* eik;         // device control words of the eikonal loop, 2 jobs
  fl::ls::EikCtl* host_eik;    // pinned host mirror of eik
  fl_tuning tuning;
  hipEvent_t step_ev;          // fl_mm_simulate: the step's control block is on the host
  int64_t eik_sweeps;          // sweeps the last eikonal solve took (first batch size)
  int use_graphs;              // fl_simulate replays a captured pair of steps
                               // (small 1D grids: runs the persistent loop)
  fl::PersistCtl* pctl;        // device control words of the persistent 1D loop
  fl::PersistCtl* host_pctl;   // pinned host mirror of pctl
  hipStream_t cap_stream;      // capture stream of that graph
  hipGraph_t sim_graph;        // a pair of steps (both parities)
  hipGraphExec_t sim_exec;
  hipGraph_t sim_graph_run;    // kSimGraphSteps steps
  hipGraphExec_t sim_exec_run;
  SimGraphKey sim_key;
};

namespace fl {
namespace abi {

constexpr int kSumBlocks = 1024;

// Records the message for fl_last_error() and returns `status`.
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define FL_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::fl::abi::fail(FL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Material constants of a scheme, derived on the host with the expressions the
// reference evaluates per call (gamma - 1, (gamma + 1) / (2 gamma)).
template <typename T>
inline fl::Material<T> make_material(const fl_scheme* s) {
  fl::Material<T> m;
  m.gamma = (T)s->gamma;
  m.p_inf = (T)s->p_inf;
  m.gm1 = m.gamma - T(1);
  m.q = (m.gamma + T(1)) / (T(2) * m.gamma);
  m.stiff = (s->eos == FL_EOS_STIFF || s->eos == FL_EOS_STIFF_CORRECTED) ? 1 : 0;
  m.corr = s->eos == FL_EOS_STIFF_CORRECTED ? 1 : 0;
  m.gpinf = m.corr ? m.gamma * m.p_inf : T(0);
  return m;
}

double bits_to_double(unsigned long long b);
// `t` (nullable: all defaults) with zeros replaced by the built-in launch
// shapes, validated, into *out.
int resolve_tuning(const fl_tuning* t, fl_tuning* out);
int check_pair(const fl_field* in, const fl_field* out);
int check_field(const fl_field* f, const char* name);
int check_scheme(const fl_scheme* s, int ndim);
// FL_ERR_UNSUPPORTED unless s asks for the exact numerics.
int require_exact(const fl_scheme* s, const char* what);
void strides(const fl_field* f, int64_t& es, int64_t& cs);
int64_t cell_count(const fl_field* f);
// One sweep; ctl/ws_out as in fl_simulate (NULL: dtdh from the argument, no
// fused wavespeed).
int launch_sweep(const fl_context* ctx, const fl_field* in, const fl_field* out, int dim,
                 const fl_scheme* s, double dtdh, const fl::StepCtl* ctl,
                 unsigned long long* ws_out, hipStream_t stream);
// Both materials' sweeps of dimension `dim` in one launch; ctl (nullable)
// as in launch_sweep.
int launch_sweep2(const fl_context* ctx, const fl_field* const in[2], const fl_field* const out[2],
                  int dim, const fl_scheme* const s[2], double dtdh, hipStream_t stream,
                  const fl::StepCtl* ctl = nullptr);
int launch_max_ws2(const fl_field* f0, const fl_field* f1, const fl_scheme* s0,
                   const fl_scheme* s1, unsigned long long* dst, hipStream_t stream);
// atomicMax of the FP64 wavespeed bits of every cell into *dst.
int launch_max_ws(const fl_field* f, const fl_scheme* s, unsigned long long* dst,
                  hipStream_t stream);

}  // namespace abi
}  // namespace fl
