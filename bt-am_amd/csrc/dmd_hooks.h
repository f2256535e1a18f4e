// dmd_hooks.h -- per-(u, i) derivatives of V0_i (reference granularity).
#pragma once
#include "dmd_kernels.h"

namespace dmd {

// For queries q < nq with arena sites u = qu[q], i = qi[q] of one instance
// whose FP64 phase-1 records `rec` are current: out[5 q .. 5 q + 4] =
// dV0_i/dq_u (3), dV0_i/dsigma_u, dV0_i/dx_u (the reference's form of
// evaluate_DV0_i_Dq_u_MgHx and evaluate_DV0_i_Dxi_u_adp_MgHx); zeros when u is
// neither i nor a neighbour of i, or when u or i is inactive.
hipError_t launch_site_dv0(hipStream_t st, int rule, const State& s, const Tables& T,
                           const double* rec, int nq, const int* qu, const int* qi, double* out);

}  // namespace dmd
