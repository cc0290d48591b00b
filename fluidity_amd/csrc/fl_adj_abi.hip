// fl_adj_abi.hip -- host side of the adjoint entry points of
// include/fluidity_amd.h: the checkpointed forward run, the backward run with
// per-segment recompute, and the single-step vector-Jacobian product.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include "fl_abi_internal.hpp"
#include "fl_adjoint_kernels.hpp"

using namespace fl::abi;
namespace aj = fl::adj;

namespace {

int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }
int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int check_adj_field(const fl_field* f, const char* name) {
  int rc = check_field(f, name);
  if (rc) return rc;
  if (f->dtype != FL_F64) return fail(FL_ERR_UNSUPPORTED, "%s: the adjoint runs in FP64", name);
  return FL_OK;
}

int check_adj_scheme(const fl_scheme* s, int ndim) {
  int rc = check_scheme(s, ndim);
  if (rc) return rc;
  if ((rc = require_exact(s, "the adjoint"))) return rc;
  if (s->eos == FL_EOS_STIFF_CORRECTED)
    return fail(FL_ERR_UNSUPPORTED, "the adjoint covers the reference's equations of state");
  if (s->flux != FL_FLUX_HLLC && s->flux != FL_FLUX_FORCE)
    return fail(FL_ERR_UNSUPPORTED, "the adjoint covers the HLLC and FORCE fluxes");
  return FL_OK;
}

bool same_shape(const fl_field* a, const fl_field* b) {
  return a->ndim == b->ndim && a->layout == b->layout && a->dtype == b->dtype &&
         a->n[0] == b->n[0] && (a->ndim < 2 || a->n[1] == b->n[1]);
}

fl::Material<double> mat64(const fl_scheme* s) {
  fl::Material<double> m;
  m.gamma = s->gamma;
  m.p_inf = s->p_inf;
  m.gm1 = m.gamma - 1.0;
  m.q = (m.gamma + 1.0) / (2.0 * m.gamma);
  m.stiff = s->eos == FL_EOS_STIFF ? 1 : 0;
  m.corr = 0;
  m.gpinf = 0.0;
  return m;
}

int64_t cells_of(const fl_field* f) { return f->n[0] * (f->ndim > 1 ? f->n[1] : 1); }
int64_t grid_bytes(const fl_field* f) { return cells_of(f) * (2 + f->ndim) * 8; }

int64_t max_faces(const fl_field* f) {
  const int64_t nx = f->n[0], ny = f->ndim > 1 ? f->n[1] : 1;
  const int64_t fx = (nx + 1) * ny;
  const int64_t fy = f->ndim > 1 ? nx * (ny + 1) : 0;
  return fx > fy ? fx : fy;
}

// Scratch of the backward pass for segments of K steps.
struct AdjLayout {
  double* U;         // K+1 grids: states of the segment
  double* X;         // K grids: x-sweep outputs (2D)
  double* lam[3];    // adjoint ping-pong + adjoint after the y sweep
  double* contrib;   // faces * 4 * S
  double* face_g;    // faces * 2
  double* P;         // (nx+8)(ny+8): pressure within 4 ghost cells
  double* V0;        // same: x velocity
  double* V1;        // same: y velocity
  double* chi;       // (nx+2)(ny+2): flattening factor within 1 ghost cell
  double* chi_g;     // faces * 2: chi adjoints of each face's two cells
  int* rec_k;        // (nx+2)(ny+2): the factor each chi minimum selected
  double* rec_v;     // (nx+2)(ny+2) * 4: its pressure adjoints
  int* any;          // 1: some chi position has a record
  double* partials;  // 4 slots * 2 * kPartialBlocks: x, y faces; x, y chi
  aj::StepRec* rec;  // K
  double* acc;       // 2
  fl::StepCtl* ctl;  // 1
  int64_t bytes;
};

AdjLayout adj_layout(void* base, const fl_field* f, int K) {
  char* p = (char*)base;
  int64_t off = 0;
  auto take = [&](int64_t b) {
    char* r = p ? p + off : nullptr;
    off += align256(b);
    return r;
  };
  const int64_t gb = grid_bytes(f);
  const int S = 2 + f->ndim;
  const int64_t nf = max_faces(f);
  AdjLayout L;
  L.U = (double*)take(gb * (K + 1));
  L.X = (double*)take(f->ndim > 1 ? gb * K : 8);
  for (int i = 0; i < 3; ++i) L.lam[i] = (double*)take(gb);
  L.contrib = (double*)take(nf * 4 * S * 8);
  L.face_g = (double*)take(nf * 2 * 8);
  const int64_t nx = f->n[0], ny = f->ndim > 1 ? f->n[1] : 1;
  const int64_t padded = (nx + 8) * (f->ndim > 1 ? ny + 8 : 1);
  L.P = (double*)take(padded * 8);
  L.V0 = (double*)take(padded * 8);
  L.V1 = (double*)take(padded * 8);
  const int64_t cpad = (nx + 2) * (f->ndim > 1 ? ny + 2 : 1);
  L.chi = (double*)take(cpad * 8);
  L.chi_g = (double*)take(nf * 2 * 8);
  L.rec_k = (int*)take(cpad * 4);
  L.rec_v = (double*)take(cpad * 4 * 8);
  L.any = (int*)take(4);
  L.partials = (double*)take(4 * 2 * aj::kPartialBlocks * 8);
  L.rec = (aj::StepRec*)take(sizeof(aj::StepRec) * K);
  L.acc = (double*)take(2 * 8);
  L.ctl = (fl::StepCtl*)take(sizeof(fl::StepCtl));
  L.bytes = off;
  return L;
}

fl_field with_data(const fl_field* f, void* data) {
  fl_field g = *f;
  g.data = data;
  return g;
}

aj::AGrid agrid(const fl_field* f, const double* data, const fl_scheme* s) {
  aj::AGrid g;
  g.U = data;
  strides(f, g.es, g.cs);
  g.nx = f->n[0];
  g.ny = f->ndim > 1 ? f->n[1] : 1;
  g.bcx = s->bc[0];
  g.bcy = f->ndim > 1 ? s->bc[1] : 0;
  return g;
}

// Forward steps a..b-1 from the state in U[0] with device-side dt: U[k+1] =
// step(U[k]); X[k] = the x-sweep output; rec[k] = the step's max wavespeed,
// dt/dh and argmax cell.  If keep_all is false, U[0] and U[1] ping-pong and
// only the final state is kept (in U[(b-a) & 1]).
int forward_steps(const fl_context* c, const fl_field* f, const fl_scheme* s, double cfl,
                  double dh, int64_t nsteps, double* Ubase, double* Xbase, aj::StepRec* rec,
                  fl::StepCtl* ctl, bool keep_all, hipStream_t st) {
  const int64_t cells = cells_of(f);
  const int S = 2 + f->ndim;
  const fl::Material<double> m = mat64(s);
  hipLaunchKernelGGL(fl::init_control_kernel, dim3(1), dim3(1), 0, st, ctl);
  fl_field u0 = with_data(f, Ubase);
  int rc = launch_max_ws(&u0, s, &ctl->ws_bits[0], st);
  if (rc) return rc;
  for (int64_t k = 0; k < nsteps; ++k) {
    const int parity = (int)(k & 1);
    double* Uk = keep_all ? Ubase + k * cells * S : Ubase + (k & 1) * cells * S;
    double* Un = keep_all ? Ubase + (k + 1) * cells * S : Ubase + ((k + 1) & 1) * cells * S;
    hipLaunchKernelGGL(fl::step_control_kernel, dim3(1), dim3(1), 0, st, ctl, parity, cfl, dh,
                       (double)INFINITY, (long long)1 << 62);
    if (rec) {
      hipLaunchKernelGGL(aj::record_kernel, dim3(1), dim3(1), 0, st, (const fl::StepCtl*)ctl,
                         parity, rec + k);
      const aj::AGrid g = agrid(f, Uk, s);
      int64_t blocks = ceil_div(cells, 256);
      if (blocks > 2048) blocks = 2048;
      if (f->ndim == 1)
        hipLaunchKernelGGL(aj::argmax_kernel<1>, dim3((unsigned)blocks), dim3(256), 0, st, g, m,
                           rec + k);
      else
        hipLaunchKernelGGL(aj::argmax_kernel<2>, dim3((unsigned)blocks), dim3(256), 0, st, g, m,
                           rec + k);
    }
    const fl_field fin = with_data(f, Uk);
    const fl_field fout = with_data(f, Un);
    unsigned long long* ws_next = &ctl->ws_bits[parity ^ 1];
    if (f->ndim == 1) {
      rc = launch_sweep(c, &fin, &fout, 0, s, 1.0, ctl, ws_next, st);
    } else {
      double* Xk = keep_all ? Xbase + k * cells * S : Xbase;
      const fl_field fx = with_data(f, Xk);
      rc = launch_sweep(c, &fin, &fx, 0, s, 1.0, ctl, nullptr, st);
      if (!rc) rc = launch_sweep(c, &fx, &fout, 1, s, 1.0, ctl, ws_next, st);
    }
    if (rc) return rc;
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Adjoint of one sweep: lam_out -> lam_in; the faces' (dt/dh, gamma) sums
// into `partials`, the flattening factor's gamma sums into `chi_partials`.
template <int ND, int D>
void sweep_adj(const fl_field* f, const fl_scheme* s, const double* Uin, const double* lam_out,
               double* lam_in, const aj::StepRec* rec, const AdjLayout& L, double* partials,
               double* chi_partials, hipStream_t st) {
  const aj::AGrid g = agrid(f, Uin, s);
  const fl::Material<double> m = mat64(s);
  const int64_t n = D == 0 ? g.nx : g.ny;
  const int64_t lines = D == 0 ? g.ny : g.nx;
  const int64_t nf = (n + 1) * lines;
  const int64_t padded = (g.nx + 8) * (ND > 1 ? g.ny + 8 : 1);
  const int64_t cpad = (g.nx + 2) * (ND > 1 ? g.ny + 2 : 1);
  hipLaunchKernelGGL(aj::prim_pad_kernel<ND>, dim3((unsigned)ceil_div(padded, 256)), dim3(256), 0,
                     st, g, m, L.P, L.V0, L.V1);
  hipLaunchKernelGGL(aj::chi_pad_kernel<ND>, dim3((unsigned)ceil_div(cpad, 256)), dim3(256), 0,
                     st, g, (const double*)L.P, (const double*)L.V0, (const double*)L.V1, L.chi,
                     L.any);
  const bool force = s->flux == FL_FLUX_FORCE, prim = s->form == FL_FORM_PRIMITIVE;
  auto face = force ? (prim ? aj::face_adj_kernel<ND, D, fl::kFluxForce, true>
                            : aj::face_adj_kernel<ND, D, fl::kFluxForce, false>)
                    : (prim ? aj::face_adj_kernel<ND, D, fl::kFluxHllc, true>
                            : aj::face_adj_kernel<ND, D, fl::kFluxHllc, false>);
  hipLaunchKernelGGL(face, dim3((unsigned)ceil_div(nf, 256)), dim3(256), 0, st, g, lam_out, m,
                     s->limiter, rec, (const double*)L.chi, L.contrib, L.face_g, L.chi_g);
  hipLaunchKernelGGL(aj::partial_sum_kernel, dim3(aj::kPartialBlocks), dim3(256), 0, st,
                     (const double*)L.face_g, nf, partials);
  const int64_t cells = g.nx * g.ny;
  hipLaunchKernelGGL((aj::gather_adj_kernel<ND, D>), dim3((unsigned)ceil_div(cells, 256)),
                     dim3(256), 0, st, g, lam_out, lam_in, (const double*)L.contrib);
  // the flattening factor: the selected factors' pressure adjoints where
  // 0 < chi < 1, then into lambda and the gamma sum
  hipLaunchKernelGGL((aj::chi_adj_kernel<ND, D>), dim3((unsigned)ceil_div(cpad, 256)), dim3(256),
                     0, st, g, (const double*)L.P, (const double*)L.V0, (const double*)L.V1,
                     (const double*)L.chi, (const double*)L.chi_g, L.rec_k, L.rec_v, L.any);
  hipLaunchKernelGGL(aj::chi_gather_kernel<ND>, dim3(aj::kPartialBlocks), dim3(256), 0, st, g, m,
                     (const int*)L.rec_k, (const double*)L.rec_v, (const int*)L.any, lam_in,
                     chi_partials);
}

// Backward through step k of the segment: lam_next -> lam_prev.
int backward_step(const fl_field* f, const fl_scheme* s, const AdjLayout& L, int64_t k,
                  const double* lam_next, double* lam_prev, hipStream_t st) {
  const int64_t cells = cells_of(f);
  const int S = 2 + f->ndim;
  const double* Uk = L.U + k * cells * S;
  const aj::StepRec* rec = L.rec + k;
  const fl::Material<double> m = mat64(s);
  // partial slots (2 * kPartialBlocks each): faces of the sweeps, then their
  // flattening-factor terms
  auto slot = [&](int i) { return L.partials + 2 * aj::kPartialBlocks * i; };
  if (f->ndim == 1) {
    sweep_adj<1, 0>(f, s, Uk, lam_next, lam_prev, rec, L, slot(0), slot(1), st);
    hipLaunchKernelGGL(aj::finalize_kernel<1>, dim3(1), dim3(256), 0, st, agrid(f, Uk, s), m, rec,
                       (const double*)L.partials, 2, lam_prev, L.acc);
  } else {
    const double* Xk = L.X + k * cells * S;
    sweep_adj<2, 1>(f, s, Xk, lam_next, L.lam[2], rec, L, slot(1), slot(3), st);
    sweep_adj<2, 0>(f, s, Uk, L.lam[2], lam_prev, rec, L, slot(0), slot(2), st);
    hipLaunchKernelGGL(aj::finalize_kernel<2>, dim3(1), dim3(256), 0, st, agrid(f, Uk, s), m, rec,
                       (const double*)L.partials, 4, lam_prev, L.acc);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int check_common(const fl_context* c, const fl_field* f, const fl_scheme* s, double cfl,
                 double dh) {
  if (!c) return fail(FL_ERR_ARGUMENT, "context is NULL");
  int rc = check_adj_field(f, "state");
  if (rc) return rc;
  if ((rc = check_adj_scheme(s, f->ndim))) return rc;
  if (!(cfl > 0.0) || !isfinite(cfl) || !(dh > 0.0) || !isfinite(dh))
    return fail(FL_ERR_ARGUMENT, "cfl and dh must be finite and > 0");
  return FL_OK;
}

}  // namespace

extern "C" {

int64_t fl_adj_checkpoint_bytes(const fl_field* state, int64_t steps, int checkpoint_every) {
  if (!state || checkpoint_every < 1 || steps < 0) return -1;
  return grid_bytes(state) * (steps / checkpoint_every + 1);
}

int64_t fl_adj_work_bytes(const fl_field* state, int checkpoint_every) {
  if (!state || checkpoint_every < 1) return -1;
  return adj_layout(nullptr, state, checkpoint_every).bytes;
}

int fl_adj_forward(fl_context* c, fl_field* state, const fl_scheme* s, double cfl, double dh,
                   int64_t steps, int checkpoint_every, void* checkpoints, int64_t ckpt_bytes,
                   void* work, int64_t work_bytes, fl_sim_result* result, void* stream) {
  int rc = check_common(c, state, s, cfl, dh);
  if (rc) return rc;
  const int K = checkpoint_every;
  if (K < 1 || steps < 0) return fail(FL_ERR_ARGUMENT, "steps >= 0 and checkpoint_every >= 1");
  if (!checkpoints || ckpt_bytes < fl_adj_checkpoint_bytes(state, steps, K))
    return fail(FL_ERR_ARGUMENT, "checkpoint buffer too small");
  const AdjLayout L = adj_layout(work, state, K);
  if (!work || work_bytes < L.bytes || ((uintptr_t)work & 255))
    return fail(FL_ERR_ARGUMENT, "work must be %lld bytes, 256-byte aligned", (long long)L.bytes);
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t gb = grid_bytes(state);
  char* ck = (char*)checkpoints;
  double t = 0.0, last_dt = 0.0;
  FL_HIP(hipMemcpyAsync(ck, state->data, gb, hipMemcpyDeviceToDevice, st));
  // segments of K steps from each checkpoint, through the work grids
  for (int64_t a = 0; a < steps; a += K) {
    const int64_t ns = (steps - a) < K ? (steps - a) : K;
    FL_HIP(hipMemcpyAsync(L.U, ck + (a / K) * gb, gb, hipMemcpyDeviceToDevice, st));
    if ((rc = forward_steps(c, state, s, cfl, dh, ns, L.U, L.X, nullptr, L.ctl, false, st)))
      return rc;
    const double* fin = L.U + (ns & 1) * cells_of(state) * (2 + state->ndim);
    if ((a + ns) % K == 0)
      FL_HIP(hipMemcpyAsync(ck + ((a + ns) / K) * gb, fin, gb, hipMemcpyDeviceToDevice, st));
    if (a + ns == steps) FL_HIP(hipMemcpyAsync(state->data, fin, gb, hipMemcpyDeviceToDevice, st));
    FL_HIP(hipMemcpyAsync(c->host_ctl, L.ctl, sizeof(fl::StepCtl), hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    t += c->host_ctl->t;
    last_dt = c->host_ctl->dt;
  }
  FL_HIP(hipStreamSynchronize(st));
  if (result) {
    result->iters = steps;
    result->run_time = t;
    result->last_dt = last_dt;
    result->max_wavespeed = NAN;
  }
  return FL_OK;
}

int fl_adj_backward(fl_context* c, fl_field* grad, const fl_scheme* s, double cfl, double dh,
                    int64_t steps, int checkpoint_every, const void* checkpoints,
                    int64_t ckpt_bytes, void* work, int64_t work_bytes, double* dgamma,
                    void* stream) {
  int rc = check_common(c, grad, s, cfl, dh);
  if (rc) return rc;
  const int K = checkpoint_every;
  if (K < 1 || steps < 0) return fail(FL_ERR_ARGUMENT, "steps >= 0 and checkpoint_every >= 1");
  if (!checkpoints || ckpt_bytes < fl_adj_checkpoint_bytes(grad, steps, K))
    return fail(FL_ERR_ARGUMENT, "checkpoint buffer too small");
  const AdjLayout L = adj_layout(work, grad, K);
  if (!work || work_bytes < L.bytes || ((uintptr_t)work & 255))
    return fail(FL_ERR_ARGUMENT, "work must be %lld bytes, 256-byte aligned", (long long)L.bytes);
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t gb = grid_bytes(grad);
  const char* ck = (const char*)checkpoints;
  hipLaunchKernelGGL(aj::acc_clear_kernel, dim3(1), dim3(1), 0, st, L.acc);
  FL_HIP(hipMemcpyAsync(L.lam[0], grad->data, gb, hipMemcpyDeviceToDevice, st));
  int cur = 0;
  const int64_t nseg = (steps + K - 1) / K;
  for (int64_t sg = nseg - 1; sg >= 0; --sg) {
    const int64_t a = sg * K;
    const int64_t ns = (steps - a) < K ? (steps - a) : K;
    FL_HIP(hipMemcpyAsync(L.U, ck + sg * gb, gb, hipMemcpyDeviceToDevice, st));
    if ((rc = forward_steps(c, grad, s, cfl, dh, ns, L.U, L.X, L.rec, L.ctl, true, st)))
      return rc;
    for (int64_t k = ns - 1; k >= 0; --k) {
      if ((rc = backward_step(grad, s, L, k, L.lam[cur], L.lam[cur ^ 1], st))) return rc;
      cur ^= 1;
    }
  }
  FL_HIP(hipMemcpyAsync(grad->data, L.lam[cur], gb, hipMemcpyDeviceToDevice, st));
  double acc[2];
  FL_HIP(hipMemcpyAsync(acc, L.acc, sizeof(acc), hipMemcpyDeviceToHost, st));
  FL_HIP(hipStreamSynchronize(st));
  if (dgamma) *dgamma = acc[0];
  return FL_OK;
}

int fl_split_step_vjp(fl_context* c, const fl_field* state, const fl_field* lambda_next,
                      fl_field* lambda, const fl_scheme* s, double cfl, double dh,
                      double* dgamma, void* work, int64_t work_bytes, void* stream) {
  int rc = check_common(c, state, s, cfl, dh);
  if (rc) return rc;
  if ((rc = check_adj_field(lambda_next, "lambda_next"))) return rc;
  if ((rc = check_adj_field(lambda, "lambda"))) return rc;
  if (!same_shape(state, lambda_next) || !same_shape(state, lambda))
    return fail(FL_ERR_ARGUMENT, "state and adjoint fields differ in shape or layout");
  const AdjLayout L = adj_layout(work, state, 1);
  if (!work || work_bytes < L.bytes || ((uintptr_t)work & 255))
    return fail(FL_ERR_ARGUMENT, "work must be %lld bytes, 256-byte aligned", (long long)L.bytes);
  DeviceGuard g(c->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t gb = grid_bytes(state);
  hipLaunchKernelGGL(aj::acc_clear_kernel, dim3(1), dim3(1), 0, st, L.acc);
  FL_HIP(hipMemcpyAsync(L.U, state->data, gb, hipMemcpyDeviceToDevice, st));
  if ((rc = forward_steps(c, state, s, cfl, dh, 1, L.U, L.X, L.rec, L.ctl, true, st))) return rc;
  FL_HIP(hipMemcpyAsync(L.lam[0], lambda_next->data, gb, hipMemcpyDeviceToDevice, st));
  if ((rc = backward_step(state, s, L, 0, L.lam[0], L.lam[1], st))) return rc;
  FL_HIP(hipMemcpyAsync(lambda->data, L.lam[1], gb, hipMemcpyDeviceToDevice, st));
  double acc[2];
  FL_HIP(hipMemcpyAsync(acc, L.acc, sizeof(acc), hipMemcpyDeviceToHost, st));
  FL_HIP(hipStreamSynchronize(st));
  if (dgamma) *dgamma = acc[0];
  return FL_OK;
}

}  // extern "C"
