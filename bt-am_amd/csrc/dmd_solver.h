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
  kUpdSdDir = 32,     // D = c_sd P
  // NGMRES (D and Rr hold the line-search point x^M and its residual F^M)
  kUpdNgKeepM = 64,     // D = Y, Rr = Rv
  kUpdNgAcceptM = 128,  // X = D, R0 = Rr
  kUpdNgPush = 256      // history slot `slot` = (X, R0)
};
enum : int { kMakeNone = 0, kMakeX = 1, kMakeXP = 2, kMakeXD = 3, kMakeXR = 4, kMakeNgA = 5 };
enum : int { kDotNone = 0, kDotNorm = 1, kDotCg = 2, kDotLs = 3, kDotDiff = 4, kDotPc = 5,
             kDotNg = 6 };
// NGMRES history depth (iterates and residuals kept per site).
constexpr int kNgHist = 3;
// Dots per instance: [0..2] sums, [3..4] maxima, [5..13] sums (preconditioned CG),
// [14] sites of the tick's evaluation whose density left the embedding
// table's domain, [15] its free energy sum_i S0_i.
constexpr int kSolveDots = 16;
constexpr int kDotDomain = 14;
constexpr int kDotEnergy = 15;

// Diffusion-step modes (SolveOp.diff): one Picard sweep of the implicit
// occupancy update, or the bare rates dx/dt written into Xt.
enum : int { kDiffNone = 0, kDiffStep = 1, kDiffRates = 2 };

// Diffusive neighbours: H sites within 2.6 A of an H site (src/Macros.hpp:860),
// at most 50 of them (src/Macros.hpp:863).
constexpr int kDCap = 50;
constexpr double kRcEb = 2.6;

// Relaxed degrees of freedom (dmdgpu_relax_opts.dof_mask): Mg positions,
// Mg widths, H positions, H widths.
enum : int { kDofMgQ = 1, kDofMgS = 2, kDofHQ = 4, kDofHS = 8 };

struct SolveOp {
  int upd, make, dot, mixed;
  int dof;          // kDof* mask; residual components outside it are zero
  int pc;           // 1: block-Jacobi preconditioned CG (SolveVecs::minv)
  double eps_prev;  // probe step of the previous tick (kUpdCgStep)
  double alpha, beta, c_sd;
  double eps;       // probe step of this tick (kMakeXP, kDotCg)
  double alpha_ls;  // line-search step (kMakeXD; kMakeXR: Y = X - alpha_ls R0)
  int slot;         // NGMRES history slot written by kUpdNgPush
  int nh;           // NGMRES history entries used by kMakeNgA
  double b[kNgHist];  // kMakeNgA: Y = D + sum_s b[s] (HX_s - D)
  int diff;         // kDiff* mode of a diffusion-step solve
  double dt;        // time step (s)
  double d0;        // diffusivity prefactor (A^2/s)
  double eb;        // migration barrier (eV)
};

struct SolveVecs {
  double4* X;    // iterate
  double4* R0;   // FP64 residual at X
  double4* R0f;  // split-precision residual at X
  double4* D;    // Newton step
  double4* Rr;   // CG residual
  double4* Pp;   // CG direction
  double* Xn;    // occupancy at the start of a diffusion step
  double* Xt;    // occupancy of the next Picard iterate (or rates)
  const double* minv;  // n x 10: inverse of the site's q block (3x3 row-major), sigma factor
  double4* hx;   // n x kNgHist: NGMRES history of iterates (site-major)
  double4* hf;   // n x kNgHist: their residuals
};

// Block-Jacobi preconditioner of an instance's sites [base, base + cap) from
// the diagonal Hessian blocks D (n x 9) of its n sites: minv = D_u^-1 on
// relaxed q of active sites (a scaled identity where D_u is not positive
// definite), 1 / (3/sigma^2 + max(tr D_u, 0)) on relaxed sigma, identity
// elsewhere (also on the padding sites [base + n, base + cap)).
hipError_t launch_pc_invert(hipStream_t st, int base, int n, int cap, int dof, const int* spc,
                            const double* ne, const double4* qs, const double* D, double* minv);
hipError_t launch_solver_step(hipStream_t st, int n, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const SolveVecs& v,
                              double4* qs, const Outputs& o);
// Filters each H site's neighbour row down to its diffusive neighbours.
// flags[0] is set on overflow (> kDCap), flags[1] when an H site has none.
hipError_t launch_diff_neighbors(hipStream_t st, int base, int n, const Box& box, const int* spc,
                                 const double4* qs, const int* nnb, const int* nbr, int* dnn,
                                 int* dnbr, int* flags);
// Picard sweep of every instance in a diffusion step: Xt = Xn + dt * xdot(x).
hipError_t launch_diff_flux(hipStream_t st, const State& s, const SolveOp* ops, const SolveVecs& v,
                            const Outputs& o, const int* dnn, const int* dnbr);
// Accepts the sweep: x = Xt (raw and effective occupancy) for kDiffStep.
hipError_t launch_diff_apply(hipStream_t st, int n, const int* inst, const int* spc,
                             const SolveOp* ops, const SolveVecs& v, double* xr, double* ne);
hipError_t launch_solver_dots(hipStream_t st, int n, int ninst, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const int* prec,
                              const SolveVecs& v, const Outputs& o, const double* xr,
                              const double4* qs, double* partials,
                              const int* chunk_lo, const int* chunk_hi, double* out);

}  // namespace dmd
