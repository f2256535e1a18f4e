// dmd_solver.h -- per-instance descriptors of the batched Newton-Krylov solver.
#pragma once
#include <hip/hip_runtime.h>

#include "dmd_kernels.h"

namespace dmd {

// Update flags, applied in this order by k_solver_step.
enum : int {
  kUpdAcceptY = 1,    // X = Y (accept the line-search trial)
  kUpdCgStep = 2,     // Ap = (Rv - Rbase)/eps_prev; D += alpha P; Rr -= alpha Ap; P = Rr + beta P
  kUpdR0FromRv = 4,   // R0 = Rv (FP64 residual at X)
  kUpdR0fFromRv = 8,  // R0f = Rv (split-precision residual at X, for FP32 probes)
  kUpdCgInit = 16,    // D = 0; Rr = P = -R0
  kUpdSdDir = 32      // D = c_sd P
};
enum : int { kMakeNone = 0, kMakeX = 1, kMakeXP = 2, kMakeXD = 3 };
enum : int { kDotNone = 0, kDotNorm = 1, kDotCg = 2, kDotLs = 3 };

// Relaxed degrees of freedom (dmdgpu_relax_opts.dof_mask): Mg positions,
// Mg widths, H positions, H widths.
enum : int { kDofMgQ = 1, kDofMgS = 2, kDofHQ = 4, kDofHS = 8 };

struct SolveOp {
  int upd, make, dot, mixed;
  int dof;          // kDof* mask; residual components outside it are zero
  double eps_prev;  // probe step of the previous tick (kUpdCgStep)
  double alpha, beta, c_sd;
  double eps;       // probe step of this tick (kMakeXP, kDotCg)
  double alpha_ls;  // line-search step (kMakeXD)
};

struct SolveVecs {
  double4* X;    // iterate
  double4* R0;   // FP64 residual at X
  double4* R0f;  // split-precision residual at X
  double4* D;    // Newton step
  double4* Rr;   // CG residual
  double4* Pp;   // CG direction
};

hipError_t launch_solver_step(hipStream_t st, int n, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const SolveVecs& v,
                              double4* qs, const Outputs& o);
hipError_t launch_solver_dots(hipStream_t st, int n, int ninst, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const SolveVecs& v,
                              const Outputs& o, double* partials, const int* chunk_lo,
                              const int* chunk_hi, double* out);

}  // namespace dmd
