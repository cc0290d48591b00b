// fl_adjoint_kernels.hpp -- argument blocks and declarations of the adjoint
// kernels (fl_adjoint.hip), shared with the C ABI host code.
#pragma once

#include "fl_kernels.hpp"

namespace fl {
namespace adj {

// An FP64 grid (state or adjoint field; both use the same strides).
struct AGrid {
  const double* U;
  int64_t es, cs;   // element and component strides
  int64_t nx, ny;
  int bcx, bcy;
};

// What the backward pass needs of one forward step.
struct StepRec {
  unsigned long long maxws_bits;  // max wavespeed of the step's input state
  double dtdh;                    // dt/dh the step's sweeps used
  unsigned long long argmax;      // lowest cell index attaining the maximum
  double pad;
};

constexpr int kPartialBlocks = 256;

template <int ND>
__global__ void prim_pad_kernel(AGrid g, Material<double> m, double* P, double* V0, double* V1);
template <int ND>
__global__ void chi_pad_kernel(AGrid g, const double* P, const double* V0, const double* V1,
                               double* chi, int* any);
template <int ND, int D, int FLUX, bool PRIM>
__global__ void face_adj_kernel(AGrid g, const double* lam, Material<double> m, int limiter,
                                const StepRec* rec, const double* chi_pad, double* contrib,
                                double* face_g, double* chi_g);
template <int ND, int D>
__global__ void chi_adj_kernel(AGrid g, const double* P, const double* V0, const double* V1,
                               const double* chi, const double* chi_g, int* rec_k,
                               double* rec_v, int* any);
template <int ND>
__global__ void chi_gather_kernel(AGrid g, Material<double> m, const int* rec_k,
                                  const double* rec_v, const int* any, double* lam,
                                  double* partials);
template <int ND, int D>
__global__ void gather_adj_kernel(AGrid g, const double* lam_out, double* lam_in,
                                  const double* contrib);
__global__ void partial_sum_kernel(const double* face_g, int64_t nfaces, double* partials);
__global__ void record_kernel(const StepCtl* ctl, int parity, StepRec* rec);
template <int ND>
__global__ void argmax_kernel(AGrid g, Material<double> m, StepRec* rec);
template <int ND>
__global__ void finalize_kernel(AGrid g, Material<double> m, const StepRec* rec,
                                const double* partials, int nsweeps, double* lam_n, double* acc);
__global__ void acc_clear_kernel(double* acc);

}  // namespace adj
}  // namespace fl
