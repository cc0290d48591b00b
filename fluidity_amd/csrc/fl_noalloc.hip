// fl_noalloc.hip -- fl_simulate_noalloc: the device step loop of fl_simulate
// for callers that cannot allocate device or pinned memory (serving loops,
// harnesses that forbid hipMalloc).  It takes no context: its only scratch,
// the step-control block that the kernels read dt/dh from and reduce the
// fused CFL maximum into, lives in a file-scope device array, one copy per
// GPU, held until the process ends.
#include <hip/hip_runtime.h>

#include <math.h>
#include <mutex>
#include <string.h>

#include "fl_abi_internal.hpp"

using namespace fl::abi;

namespace {

// Room for one step-control block (fl::StepCtl, 48 bytes).  The loop needs
// exactly one whatever the grid size, so no grid is too large for it; every
// call on a device shares that device's copy.
constexpr size_t kScratchBytes = 256;
static_assert(sizeof(fl::StepCtl) <= kScratchBytes, "step-control block must fit the scratch");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
fl::StepCtl* scratch_ptr[kMaxDevices];  // device address of `scratch` per device

// The device address of this device's `scratch` (looked up once per device,
// with that device current).  In host code the array's name is only a
// host-side handle: kernels and copies take this address.
int scratch_of(int dev, fl::StepCtl** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(FL_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FL_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                  hipGetErrorString(e));
    scratch_ptr[dev] = (fl::StepCtl*)p;
  }
  *out = scratch_ptr[dev];
  return FL_OK;
}

}  // namespace

extern "C" {

int fl_simulate_noalloc(const fl_tuning* tuning, fl_field* state, fl_field* work,
                        const fl_scheme* s, const fl_sim_params* p, fl_sim_result* result,
                        void* stream) {
  int rc = check_pair(state, work);
  if (rc) return rc;
  rc = check_scheme(s, state->ndim);
  if (rc) return rc;
  if (!p) return fail(FL_ERR_ARGUMENT, "params is NULL");
  if (!(p->cfl > 0.0) || !isfinite(p->cfl)) return fail(FL_ERR_ARGUMENT, "cfl must be > 0");
  if (!(p->dh > 0.0) || !isfinite(p->dh)) return fail(FL_ERR_ARGUMENT, "dh must be > 0");
  if (p->max_iters < 0) return fail(FL_ERR_ARGUMENT, "max_iters must be >= 0");
  if (isnan(p->sim_time)) return fail(FL_ERR_ARGUMENT, "sim_time is NaN");
  // A context on the stack that carries only the launch shapes: launch_sweep
  // reads nothing else from it, and nothing here allocates.
  fl_context c;
  memset(&c, 0, sizeof(c));
  FL_HIP(hipGetDevice(&c.device));
  if ((rc = resolve_tuning(tuning, &c.tuning))) return rc;
  fl::StepCtl* ctl = nullptr;
  if ((rc = scratch_of(c.device, &ctl))) return rc;
  hipStream_t st = (hipStream_t)stream;

  hipLaunchKernelGGL(fl::init_control_kernel, dim3(1), dim3(1), 0, st, ctl);
  if ((rc = launch_max_ws(state, s, &ctl->ws_bits[0], st))) return rc;
  // Steps in batches; the host reads the control block once per batch into a
  // stack copy and stops when the loop condition fails.  Steps past it do
  // nothing (their kernels read the inactive flag).
  const int64_t kBatch = 256;
  fl_field a = *state, b = *work;
  fl::StepCtl hc;
  memset(&hc, 0, sizeof(hc));
  int64_t launched = 0;
  do {
    const int64_t left = p->max_iters - launched;
    const int64_t nb = left < kBatch ? left : kBatch;
    for (int64_t k = 0; k < nb; ++k) {
      const int parity = (int)((launched + k) & 1);
      hipLaunchKernelGGL(fl::step_control_kernel, dim3(1), dim3(1), 0, st, ctl, parity, p->cfl,
                         p->dh, p->sim_time, (long long)p->max_iters);
      unsigned long long* ws_next = &ctl->ws_bits[parity ^ 1];
      if (state->ndim == 1) {
        rc = launch_sweep(&c, parity ? &b : &a, parity ? &a : &b, 0, s, 1.0, ctl, ws_next, st);
      } else {
        rc = launch_sweep(&c, &a, &b, 0, s, 1.0, ctl, nullptr, st);
        if (!rc) rc = launch_sweep(&c, &b, &a, 1, s, 1.0, ctl, ws_next, st);
      }
      if (rc) return rc;
    }
    launched += nb;
    FL_HIP(hipMemcpyAsync(&hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
  } while (launched < p->max_iters && hc.active);
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

}  // extern "C"
