// dmd_solver.hip -- batched vector kernels of the Newton-Krylov relaxation.
//
// Every resident instance runs its own inexact Newton iteration on the
// unknowns (q, sigma) of its active sites; one "tick" evaluates the residual of
// every instance at that instance's evaluation point Y (its iterate, a
// finite-difference probe X + eps P, or a line-search trial X + alpha D).
// These kernels apply the per-instance vector updates that the host's state
// machine requests (descriptor per instance) and form the per-instance dot
// products and maxima the host needs, in a fixed reduction order.
//
// Per-site vectors are double4 (q_x, q_y, q_z, sigma).  The residual of the
// evaluation point is Rv = (R_q, R_sigma) from the phase-2 kernel.
#include "dmd_device.h"
#include "dmd_solver.h"

namespace dmd {

// Residual of the evaluation point restricted to the relaxed degrees of
// freedom of site i (inactive sites and frozen components give zero).
DEV double4 ld_rv(const Outputs& o, int i, int dof, int slot, bool active) {
  const int mq = slot == 0 ? kDofMgQ : kDofHQ, ms = slot == 0 ? kDofMgS : kDofHS;
  const double fq = active && (dof & mq) ? 1.0 : 0.0;
  const double fs = active && (dof & ms) ? 1.0 : 0.0;
  return make_double4(fq * o.Rq[3 * i], fq * o.Rq[3 * i + 1], fq * o.Rq[3 * i + 2], fs * o.Rs[i]);
}
DEV double4 axpy4(double a, const double4& x, const double4& y) {
  return make_double4(y.x + a * x.x, y.y + a * x.y, y.z + a * x.z, y.w + a * x.w);
}
DEV double dot4(const double4& a, const double4& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

// Applies the requested updates of each instance, then writes its evaluation
// point into the residual kernels' input qs.
__global__ __launch_bounds__(256) void k_solver_step(int n, const int* inst, const int* spc,
                                                    const double* ne, const SolveOp* ops,
                                                    SolveVecs v, double4* qs, Outputs o) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int k = inst[i];
  if (k < 0) return;
  const SolveOp op = ops[k];
  if (op.upd == 0 && op.make == kMakeNone) return;
  double4 X = v.X[i];
  const double4 Y = qs[i];
  if (op.upd & kUpdAcceptY) { X = Y; v.X[i] = X; }
  if (op.upd & (kUpdR0FromRv | kUpdR0fFromRv | kUpdCgStep)) {
    const double4 Rv = ld_rv(o, i, op.dof, spc[i], ne[i] != 0.0);
    if (op.upd & kUpdCgStep) {
      // Ap = (R(X + eps P) - R(X)) / eps, the finite-difference Hessian action
      const double4 Rb = op.mixed ? v.R0f[i] : v.R0[i];
      const double ie = 1.0 / op.eps_prev;
      const double4 Ap = make_double4((Rv.x - Rb.x) * ie, (Rv.y - Rb.y) * ie, (Rv.z - Rb.z) * ie,
                                      (Rv.w - Rb.w) * ie);
      const double4 Pp = v.Pp[i];
      const double4 D = axpy4(op.alpha, Pp, v.D[i]);
      const double4 Rr = axpy4(-op.alpha, Ap, v.Rr[i]);
      v.D[i] = D;
      v.Rr[i] = Rr;
      v.Pp[i] = axpy4(op.beta, Pp, Rr);
    }
    if (op.upd & kUpdR0FromRv) v.R0[i] = Rv;
    if (op.upd & kUpdR0fFromRv) v.R0f[i] = Rv;
  }
  if (op.upd & kUpdCgInit) {
    const double4 R0 = v.R0[i];
    const double4 m = make_double4(-R0.x, -R0.y, -R0.z, -R0.w);
    v.D[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    v.Rr[i] = m;
    v.Pp[i] = m;
  }
  if (op.upd & kUpdSdDir) v.D[i] = make_double4(op.c_sd * v.Pp[i].x, op.c_sd * v.Pp[i].y,
                                                op.c_sd * v.Pp[i].z, op.c_sd * v.Pp[i].w);
  double4 Yn = Y;
  if (op.make == kMakeX) Yn = X;
  else if (op.make == kMakeXP) Yn = axpy4(op.eps, v.Pp[i], X);
  else if (op.make == kMakeXD) Yn = axpy4(op.alpha_ls, v.D[i], X);
  if (op.make != kMakeNone) qs[i] = Yn;
}

// Per-chunk partials: [0] sum a, [1] sum b, [2] sum c, [3] max d, [4] max e.
//   kDotNorm : a = |Rv|^2
//   kDotCg   : Ap = (Rv - Rbase)/eps; a = P.Ap, b = Rr.Ap, c = Ap.Ap
//   kDotLs   : d = max |D_q| (component), e = max(-D_sigma / X_sigma)
//   kDotDiff : over H sites, a = sum Xt (hydrogen content of the new iterate),
//              b = sum x (of the current one), d = max |Xt - x|,
//              e = max(-Xt, Xt - 1) (> 0 when an occupancy leaves [0, 1])
__global__ __launch_bounds__(256) void k_solver_dots(int n, const int* inst, const int* spc,
                                                    const double* ne, const SolveOp* ops,
                                                    SolveVecs v, Outputs o, const double* xr,
                                                    double* partials) {
  __shared__ double s[5][256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int k = inst[blockIdx.x * 256];
  double a = 0, b = 0, c = 0, d = 0, e = -1e300;
  if (i < n && k >= 0 && ops[k].dot == kDotDiff) {
    if (spc[i] == 1) {
      const double xt = v.Xt[i], x = xr[i];
      a = xt;
      b = x;
      d = fabs(xt - x);
      e = fmax(-xt, xt - 1.0);
    }
  } else if (i < n && k >= 0 && ne[i] != 0.0) {
    const SolveOp op = ops[k];
    if (op.dot == kDotNorm) {
      const double4 Rv = ld_rv(o, i, op.dof, spc[i], true);
      a = dot4(Rv, Rv);
    } else if (op.dot == kDotCg) {
      const double4 Rv = ld_rv(o, i, op.dof, spc[i], true);
      const double4 Rb = op.mixed ? v.R0f[i] : v.R0[i];
      const double ie = 1.0 / op.eps;
      const double4 Ap = make_double4((Rv.x - Rb.x) * ie, (Rv.y - Rb.y) * ie, (Rv.z - Rb.z) * ie,
                                      (Rv.w - Rb.w) * ie);
      a = dot4(v.Pp[i], Ap);
      b = dot4(v.Rr[i], Ap);
      c = dot4(Ap, Ap);
    } else if (op.dot == kDotLs) {
      const double4 D = v.D[i];
      d = fmax(fabs(D.x), fmax(fabs(D.y), fabs(D.z)));
      e = -D.w / v.X[i].w;
    }
  }
  s[0][threadIdx.x] = a; s[1][threadIdx.x] = b; s[2][threadIdx.x] = c;
  s[3][threadIdx.x] = d; s[4][threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
      s[2][threadIdx.x] += s[2][threadIdx.x + w];
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + w]);
      s[4][threadIdx.x] = fmax(s[4][threadIdx.x], s[4][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 5) partials[5 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// One block per instance over its chunk range, fixed order.
__global__ __launch_bounds__(64) void k_solver_reduce(const int* chunk_lo, const int* chunk_hi,
                                                     const SolveOp* ops, const double* partials,
                                                     double* out) {
  const int k = blockIdx.x;
  if (threadIdx.x != 0) return;
  if (ops[k].dot == kDotNone) return;
  double a = 0, b = 0, c = 0, d = 0, e = -1e300;
  for (int ch = chunk_lo[k]; ch < chunk_hi[k]; ch++) {
    a += partials[5 * ch];
    b += partials[5 * ch + 1];
    c += partials[5 * ch + 2];
    d = fmax(d, partials[5 * ch + 3]);
    e = fmax(e, partials[5 * ch + 4]);
  }
  out[5 * k] = a; out[5 * k + 1] = b; out[5 * k + 2] = c; out[5 * k + 3] = d; out[5 * k + 4] = e;
}

hipError_t launch_solver_step(hipStream_t st, int n, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const SolveVecs& v,
                              double4* qs, const Outputs& o) {
  k_solver_step<<<(n + 255) / 256, 256, 0, st>>>(n, inst, spc, ne, ops, v, qs, o);
  return hipGetLastError();
}

hipError_t launch_solver_dots(hipStream_t st, int n, int ninst, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const SolveVecs& v,
                              const Outputs& o, const double* xr, double* partials,
                              const int* chunk_lo, const int* chunk_hi, double* out) {
  k_solver_dots<<<(n + 255) / 256, 256, 0, st>>>(n, inst, spc, ne, ops, v, o, xr, partials);
  k_solver_reduce<<<ninst, 64, 0, st>>>(chunk_lo, chunk_hi, ops, partials, out);
  return hipGetLastError();
}

}  // namespace dmd
