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
// z = M^-1 r of one site (3x3 block on q, scalar on sigma).
DEV double4 apply_minv(const double* m, const double4& r) {
  return make_double4(m[0] * r.x + m[1] * r.y + m[2] * r.z, m[3] * r.x + m[4] * r.y + m[5] * r.z,
                      m[6] * r.x + m[7] * r.y + m[8] * r.z, m[9] * r.w);
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
  double4 R0 = v.R0[i];
  if (op.upd & (kUpdR0FromRv | kUpdR0fFromRv | kUpdCgStep | kUpdNgKeepM)) {
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
      v.Pp[i] = axpy4(op.beta, Pp, op.pc ? apply_minv(v.minv + 10 * (size_t)i, Rr) : Rr);
    }
    if (op.upd & kUpdR0FromRv) { R0 = Rv; v.R0[i] = Rv; }
    if (op.upd & kUpdR0fFromRv) v.R0f[i] = Rv;
    if (op.upd & kUpdNgKeepM) { v.D[i] = Y; v.Rr[i] = Rv; }
  }
  if (op.upd & kUpdNgAcceptM) {
    X = v.D[i];
    R0 = v.Rr[i];
    v.X[i] = X;
    v.R0[i] = R0;
  }
  if (op.upd & kUpdNgPush) {
    v.hx[(size_t)kNgHist * i + op.slot] = X;
    v.hf[(size_t)kNgHist * i + op.slot] = R0;
  }
  if (op.upd & kUpdCgInit) {
    const double4 R0 = v.R0[i];
    const double4 m = make_double4(-R0.x, -R0.y, -R0.z, -R0.w);
    v.D[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    v.Rr[i] = m;
    v.Pp[i] = op.pc ? apply_minv(v.minv + 10 * (size_t)i, m) : m;
  }
  if (op.upd & kUpdSdDir) v.D[i] = make_double4(op.c_sd * v.Pp[i].x, op.c_sd * v.Pp[i].y,
                                                op.c_sd * v.Pp[i].z, op.c_sd * v.Pp[i].w);
  double4 Yn = Y;
  if (op.make == kMakeX) Yn = X;
  else if (op.make == kMakeXP) Yn = axpy4(op.eps, v.Pp[i], X);
  else if (op.make == kMakeXD) Yn = axpy4(op.alpha_ls, v.D[i], X);
  else if (op.make == kMakeXR) Yn = axpy4(-op.alpha_ls, R0, X);
  else if (op.make == kMakeNgA) {
    const double4 M = v.D[i];
    Yn = M;
    for (int t = 0; t < op.nh; t++) {
      const double4 H = v.hx[(size_t)kNgHist * i + t];
      Yn = axpy4(op.b[t], make_double4(H.x - M.x, H.y - M.y, H.z - M.z, H.w - M.w), Yn);
    }
  }
  if (op.make != kMakeNone) qs[i] = Yn;
}

// Per-chunk partials: [0] sum a, [1] sum b, [2] sum c, [3] max d, [4] max e,
// [5..13] sums g[0..8].
//   kDotNorm : a = |Rv|^2
//   kDotCg   : Ap = (Rv - Rbase)/eps; a = P.Ap, b = Rr.Ap, c = Ap.Ap; with the
//              preconditioner (z = M^-1 Rr, y = M^-1 Ap) also g = z.Ap, Ap.y,
//              Rr.z, z.z, z.y, y.y, z.P, y.P, P.P -- what the host needs to
//              advance r.z and |P|^2 of the next direction without a new pass
//   kDotPc   : after CgInit (P = M^-1 Rr): g[2] = Rr.P, g[8] = P.P
//   kDotLs   : a = R0 . D (the slope of the energy along D), d = max |D_q|
//              (component), e = max(-D_sigma / X_sigma)
//   kDotNg   : NGMRES, at the evaluation point Y: a = |Rv|^2, b = Rv . R0 (the
//              line-search direction), c = relaxed widths <= 0, d = max |Rv_q|
//              (component), e = max
//              Rv_sigma / Y_sigma over relaxed widths (the step bounds if Y
//              becomes the iterate), g[0..2] = Rv . HF_s, g[3] = |Y - D|^2,
//              g[4..6] = |HX_s - Y|^2
//   kDotDiff : over H sites, a = sum Xt (hydrogen content of the new iterate),
//              b = sum x (of the current one), d = max |Xt - x|,
//              e = max(-Xt, Xt - 1) (> 0 when an occupancy leaves [0, 1])
__global__ __launch_bounds__(256) void k_solver_dots(int n, const int* inst, const int* spc,
                                                    const double* ne, const SolveOp* ops,
                                                    SolveVecs v, Outputs o, const double* xr,
                                                    const double4* qs, double* partials) {
  __shared__ double s[kDotDomain][256];  // the reduced slots 0..13
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int k = inst[blockIdx.x * 256];
  double a = 0, b = 0, c = 0, d = 0, e = -1e300;
  double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
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
      const double4 P = v.Pp[i], Rr = v.Rr[i];
      a = dot4(P, Ap);
      b = dot4(Rr, Ap);
      c = dot4(Ap, Ap);
      if (op.pc) {
        const double* m = v.minv + 10 * (size_t)i;
        const double4 z = apply_minv(m, Rr), y = apply_minv(m, Ap);
        g[0] = dot4(z, Ap); g[1] = dot4(Ap, y); g[2] = dot4(Rr, z);
        g[3] = dot4(z, z); g[4] = dot4(z, y); g[5] = dot4(y, y);
        g[6] = dot4(z, P); g[7] = dot4(y, P); g[8] = dot4(P, P);
      }
    } else if (op.dot == kDotNg) {
      const int sl = spc[i];
      const double4 Rv = ld_rv(o, i, op.dof, sl, true);
      const double4 Y = qs[i], M = v.D[i];
      a = dot4(Rv, Rv);
      b = dot4(Rv, v.R0[i]);
      d = fmax(fabs(Rv.x), fmax(fabs(Rv.y), fabs(Rv.z)));
      if (op.dof & (sl == 0 ? kDofMgS : kDofHS)) {
        e = Rv.w / Y.w;
        c = Y.w > 0.0 ? 0.0 : 1.0;  // a relaxed width that is not positive
      }
      const double4 dm = make_double4(Y.x - M.x, Y.y - M.y, Y.z - M.z, Y.w - M.w);
      g[3] = dot4(dm, dm);
      for (int t = 0; t < op.nh; t++) {  // the history entries in use
        const double4 H = v.hx[(size_t)kNgHist * i + t];
        const double4 dh = make_double4(H.x - Y.x, H.y - Y.y, H.z - Y.z, H.w - Y.w);
        g[t] = dot4(Rv, v.hf[(size_t)kNgHist * i + t]);
        g[4 + t] = dot4(dh, dh);
      }
    } else if (op.dot == kDotPc) {
      const double4 P = v.Pp[i];
      g[2] = dot4(v.Rr[i], P);
      g[8] = dot4(P, P);
    } else if (op.dot == kDotLs) {
      const double4 D = v.D[i];
      a = dot4(v.R0[i], D);
      d = fmax(fabs(D.x), fmax(fabs(D.y), fabs(D.z)));
      e = -D.w / v.X[i].w;
    }
  }
  s[0][threadIdx.x] = a; s[1][threadIdx.x] = b; s[2][threadIdx.x] = c;
  s[3][threadIdx.x] = d; s[4][threadIdx.x] = e;
#pragma unroll
  for (int t = 0; t < 9; t++) s[5 + t][threadIdx.x] = g[t];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
      s[2][threadIdx.x] += s[2][threadIdx.x + w];
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + w]);
      s[4][threadIdx.x] = fmax(s[4][threadIdx.x], s[4][threadIdx.x + w]);
#pragma unroll
      for (int t = 5; t < kDotDomain; t++) s[t][threadIdx.x] += s[t][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < kDotDomain)
    partials[kSolveDots * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// One block per instance over its chunk range, fixed order.  Slots 14 and 15
// carry the table-domain count and the free energy of the instance's
// evaluation in this tick (per-instance totals; 0 when it was not evaluated).
__global__ __launch_bounds__(64) void k_solver_reduce(const int* chunk_lo, const int* chunk_hi,
                                                     const SolveOp* ops, const int* prec,
                                                     const double* totals, const double* partials,
                                                     double* out) {
  const int k = blockIdx.x;
  if (threadIdx.x != 0) return;
  double* o = out + (size_t)kSolveDots * k;
  o[kDotDomain] = prec[k] >= 0 ? totals[3 * k + 2] : 0.0;
  o[kDotEnergy] = prec[k] >= 0 ? totals[3 * k] : 0.0;
  if (ops[k].dot == kDotNone) return;
  double a = 0, b = 0, c = 0, d = 0, e = -1e300;
  double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int ch = chunk_lo[k]; ch < chunk_hi[k]; ch++) {
    const double* p = partials + (size_t)kSolveDots * ch;
    a += p[0];
    b += p[1];
    c += p[2];
    d = fmax(d, p[3]);
    e = fmax(e, p[4]);
#pragma unroll
    for (int t = 0; t < 9; t++) g[t] += p[5 + t];
  }
  o[0] = a; o[1] = b; o[2] = c; o[3] = d; o[4] = e;
#pragma unroll
  for (int t = 0; t < 9; t++) o[5 + t] = g[t];
}

__global__ __launch_bounds__(256) void k_pc_invert(int base, int n, int cap, int dof,
                                                  const int* spc, const double* ne,
                                                  const double4* qs, const double* D,
                                                  double* minv) {
  const int li = blockIdx.x * 256 + threadIdx.x;
  if (li >= cap) return;
  const int i = base + li;
  double* m = minv + 10 * (size_t)i;
  if (li >= n) {  // padding sites of the instance's range: identity
#pragma unroll
    for (int t = 0; t < 10; t++) m[t] = (t % 4 == 0 || t == 9) ? 1.0 : 0.0;
    return;
  }
  const double* d = D + 9 * (size_t)li;
  const bool act = ne[i] != 0.0;
  const int slot = spc[i];
  const bool rq = act && (dof & (slot == 0 ? kDofMgQ : kDofHQ));
  const bool rs = act && (dof & (slot == 0 ? kDofMgS : kDofHS));
  const double tr = d[0] + d[4] + d[8];
  // q block: inverse by cofactors when all leading minors are positive
  const double c00 = d[4] * d[8] - d[5] * d[7], c01 = d[5] * d[6] - d[3] * d[8],
               c02 = d[3] * d[7] - d[4] * d[6];
  const double det = d[0] * c00 + d[1] * c01 + d[2] * c02;
  const bool spd = d[0] > 0.0 && d[0] * d[4] - d[1] * d[3] > 0.0 && det > 0.0;
  if (rq && spd) {
    const double id = 1.0 / det;
    m[0] = c00 * id;
    m[1] = (d[2] * d[7] - d[1] * d[8]) * id;
    m[2] = (d[1] * d[5] - d[2] * d[4]) * id;
    m[3] = c01 * id;
    m[4] = (d[0] * d[8] - d[2] * d[6]) * id;
    m[5] = (d[2] * d[3] - d[0] * d[5]) * id;
    m[6] = c02 * id;
    m[7] = (d[1] * d[6] - d[0] * d[7]) * id;
    m[8] = (d[0] * d[4] - d[1] * d[3]) * id;
  } else {
    const double sc = rq && tr > 0.0 ? 3.0 / tr : 1.0;
#pragma unroll
    for (int t = 0; t < 9; t++) m[t] = (t % 4 == 0) ? sc : 0.0;
  }
  const double sg = qs[i].w;
  m[9] = rs ? 1.0 / (3.0 / (sg * sg) + fmax(tr, 0.0)) : 1.0;
}

hipError_t launch_pc_invert(hipStream_t st, int base, int n, int cap, int dof, const int* spc,
                            const double* ne, const double4* qs, const double* D, double* minv) {
  if (cap <= 0) return hipSuccess;
  k_pc_invert<<<(cap + 255) / 256, 256, 0, st>>>(base, n, cap, dof, spc, ne, qs, D, minv);
  return hipGetLastError();
}

hipError_t launch_solver_step(hipStream_t st, int n, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const SolveVecs& v,
                              double4* qs, const Outputs& o) {
  k_solver_step<<<(n + 255) / 256, 256, 0, st>>>(n, inst, spc, ne, ops, v, qs, o);
  return hipGetLastError();
}

hipError_t launch_solver_dots(hipStream_t st, int n, int ninst, const int* inst, const int* spc,
                              const double* ne, const SolveOp* ops, const int* prec,
                              const SolveVecs& v, const Outputs& o, const double* xr,
                              const double4* qs, double* partials,
                              const int* chunk_lo, const int* chunk_hi, double* out) {
  k_solver_dots<<<(n + 255) / 256, 256, 0, st>>>(n, inst, spc, ne, ops, v, o, xr, qs, partials);
  k_solver_reduce<<<ninst, 64, 0, st>>>(chunk_lo, chunk_hi, ops, prec, o.totals, partials, out);
  return hipGetLastError();
}

}  // namespace dmd
