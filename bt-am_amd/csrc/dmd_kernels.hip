// dmd_kernels.hip -- gfx950 kernels of the mean-field DMD force path.
//
// Work decomposition: K consecutive lanes of a wavefront own one site; each
// lane walks every K-th neighbour of the site and the K partial sums are
// combined with a butterfly of cross-lane shuffles.  All reductions run in a
// fixed order, so a site's result does not depend on what else is in flight.
//
// Mathematics (FP64).  The per-point quantities are the reference's
// Gaussian-averaged ADP terms
//   rho_ij  = n_j rho(r)                 (src/ADP/MgHx-ADP.cpp:526-553)
//   V_pair  = 1/2 n_i n_j phi(r)         (:869-900)
//   V_dip   = 1/2 a_j1 . a_j2, a = n_i n_j u(r) r                 (:1233-1281)
//   V_quad  = 1/2 B_j1:B_j2 - 1/6 trB_j1 trB_j2, B = n_i n_j w r r^T (:1781-1831)
// under the 3rd-order Stroud rule (src/Numerical/Quadrature-Hermitian-3th.cpp):
// sqrt(6) sigma shifts with weight 1/12 for pairs and coincident triplets,
// 3 sigma shifts with weight 1/18 for distinct triplets.  The reference sums
// the triplets over ordered pairs j1 <= j2 (MgHx-mf-V-bulk.cpp:879-936); the
// angular terms are bilinear in psi_j(r) = n_j (u r, w r r^T), and the 9D rule
// moves one site at a time, so with
//   G0 = sum_j psi_j(r_j), G_ls = sum_j psi_j(r_j + 3 s sigma_i e_l),
//   T_j = sum_ls psi_j(r_j + 3 s sigma_j e_l), U = sum_j T_j
// the distinct-triplet sum is
//   1/18 [ sum_ls (<G_ls,G_ls> - sum_j <psi_j^ls,psi_j^ls>) + 2<U,G0>
//          - 2 sum_j <T_j, psi_j(r_j)> ]
// -- O(nb) per site instead of O(nb^2).  Phase 1 forms these sums; phase 2
// forms the gradient from their adjoints.
#include "dmd_kernels.h"

namespace dmd {

#define DEV __device__ __forceinline__

constexpr double kSqrt6 = 2.449489742783178;  // src/Macros.hpp:42
constexpr double kUma = 103.4993333E-6;        // src/Macros.hpp:851
constexpr double kHbar = 6.5821192815E-4;      // src/Macros.hpp:847

// ---------------------------------------------------------------------------
// Splines (reference clamp, src/Numerical/cubic-spline.cpp:84-128)
// ---------------------------------------------------------------------------
DEV int seg_index(double x, double dx, double inv_dx, int n, double& p) {
  int m = (int)(x * inv_dx);
  m = m > n - 1 ? n - 1 : m;
  m = m < 0 ? 0 : m;
  double pp = x - (double)m * dx;
  p = pp > dx ? dx : pp;
  return m;
}

DEV double cval(const double* c, double p) { return c[0] + (c[1] + (c[2] + c[3] * p) * p) * p; }
DEV double cder(const double* c, double p) { return c[4] + (c[5] + c[6] * p) * p; }

// u and w (values only) at distance r for pair slot t.
DEV void uw_val(const Tables& T, int t, double r, double& u, double& w) {
  double p;
  int m = seg_index(r, T.dxr, T.inv_dxr, T.nr, p);
  const double* c = T.uw + ((size_t)(t * T.nr + m) << 4);
  u = cval(c, p);
  w = cval(c + 8, p);
}
// u, u', w, w' at distance r.
DEV void uw_vald(const Tables& T, int t, double r, double& u, double& du, double& w, double& dw) {
  double p;
  int m = seg_index(r, T.dxr, T.inv_dxr, T.nr, p);
  const double* c = T.uw + ((size_t)(t * T.nr + m) << 4);
  u = cval(c, p);
  du = cder(c, p);
  w = cval(c + 8, p);
  dw = cder(c + 8, p);
}

// ---------------------------------------------------------------------------
// Feature vectors psi = (u p, w p p^T) stored as 9 doubles
// (a_x a_y a_z B_xx B_yy B_zz B_xy B_xz B_yz) and the bilinear form
// <X,Y> = 1/2 a.a' + 1/2 B:B' - 1/6 trB trB'.
// ---------------------------------------------------------------------------
struct V9 {
  double v[9];
};

DEV void psi_from(double u, double w, const double* p, double* o) {
  o[0] = u * p[0]; o[1] = u * p[1]; o[2] = u * p[2];
  o[3] = w * p[0] * p[0]; o[4] = w * p[1] * p[1]; o[5] = w * p[2] * p[2];
  o[6] = w * p[0] * p[1]; o[7] = w * p[0] * p[2]; o[8] = w * p[1] * p[2];
}

DEV double dotM(const double* x, const double* y) {
  double trx = x[3] + x[4] + x[5], tr_y = y[3] + y[4] + y[5];
  return 0.5 * (x[0] * y[0] + x[1] * y[1] + x[2] * y[2] + x[3] * y[3] + x[4] * y[4] + x[5] * y[5]) +
         (x[6] * y[6] + x[7] * y[7] + x[8] * y[8]) - (1.0 / 6.0) * trx * tr_y;
}

// o = c * M x
DEV void applyM(double c, const double* x, double* o) {
  double tr = (x[3] + x[4] + x[5]) * (1.0 / 6.0);
  double h = 0.5 * c;
  o[0] = h * x[0]; o[1] = h * x[1]; o[2] = h * x[2];
  o[3] = c * (0.5 * x[3] - tr); o[4] = c * (0.5 * x[4] - tr); o[5] = c * (0.5 * x[5] - tr);
  o[6] = c * x[6]; o[7] = c * x[7]; o[8] = c * x[8];
}

DEV double dot9(const double* x, const double* y) {
  double a = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) a += x[k] * y[k];
  return a;
}

// Gradient in p of L . psi(p) given u, u', w, w' at r = |p| (inv_r = 1/r).
DEV void grad_psi(const double* L, const double* p, double inv_r, double u, double du, double w,
                  double dw, double* g) {
  double la = L[0] * p[0] + L[1] * p[1] + L[2] * p[2];
  double g0 = 2.0 * L[3] * p[0] + L[6] * p[1] + L[7] * p[2];
  double g1 = L[6] * p[0] + 2.0 * L[4] * p[1] + L[8] * p[2];
  double g2 = L[7] * p[0] + L[8] * p[1] + 2.0 * L[5] * p[2];
  double Q = 0.5 * (p[0] * g0 + p[1] * g1 + p[2] * g2);
  double c = (du * la + dw * Q) * inv_r;
  g[0] = c * p[0] + u * L[0] + w * g0;
  g[1] = c * p[1] + u * L[1] + w * g1;
  g[2] = c * p[2] + u * L[2] + w * g2;
}

DEV void flip9(const double* v, double* o) {
  o[0] = -v[0]; o[1] = -v[1]; o[2] = -v[2];
#pragma unroll
  for (int c = 3; c < 9; c++) o[c] = v[c];
}

DEV void min_image(const Box& b, const double4& qi, const double4& qj, double* r) {
  double v0 = qi.x - qj.x, v1 = qi.y - qj.y, v2 = qi.z - qj.z;
  if (b.per[0]) v0 -= b.L[0] * rint(v0 * b.invL[0]);
  if (b.per[1]) v1 -= b.L[1] * rint(v1 * b.invL[1]);
  if (b.per[2]) v2 -= b.L[2] * rint(v2 * b.invL[2]);
  r[0] = v0; r[1] = v1; r[2] = v2;
}

DEV int pair_slot(int si, int sj) { return si == sj ? si : 2; }

template <int K>
DEV double group_sum(double v) {
#pragma unroll
  for (int o = K / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, K);
  return v;
}

// XCD-aware block order: the hardware deals consecutive workgroups to the 8
// XCDs round-robin; remap so each XCD walks one contiguous range of sites and
// the neighbour records it gathers stay in its own L2.
DEV int xcd_block(int b, int nb) {
  const int nx = 8;
  const int q = nb / nx, rem = nb % nx;
  const int x = b % nx, k = b / nx;
  // XCD x owns q (+1 when x < rem) consecutive logical blocks.
  return x * q + (x < rem ? x : rem) + k;
}

// ---------------------------------------------------------------------------
// Phase 1: per-site sums.  For site i:
//   G0 = sum_j n_j psi(r_ij),  G_ls = sum_j n_j psi(r_ij + 3 s sigma_i e_l),
//   U = sum_j n_j T_j with T_j = sum_ls psi(r_ij + 3 s sigma_j e_l),
//   rho_bar, pair sum, coincident and self terms on the sqrt(6) sigma points.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void k_phase1(State s, Box box, Tables T, Outputs o) {
  const int nblk = gridDim.x;
  const int blk = xcd_block(blockIdx.x, nblk);
  const int gt = blk * blockDim.x + threadIdx.x;
  const int i = gt / K;
  const int g = gt % K;
  if (i >= s.n) return;
  double* rec = o.rec + (size_t)i * kRec;
  const double ni = s.ne[i];
  if (ni == 0.0) {
    for (int c = g; c < kRec; c += K) rec[c] = 0.0;
    if (g == 0) { o.V0[i] = 0.0; o.rho[i] = 0.0; }
    return;
  }
  const double4 qi = s.qs[i];
  const double si = qi.w;
  const int spi = s.spc[i];
  double rho = 0, Ep = 0, Cc = 0, selfA = 0, TPsi = 0;
  double G0[9], U[9], GA[6][9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    G0[c] = 0; U[c] = 0;
#pragma unroll
    for (int ls = 0; ls < 6; ls++) GA[ls][c] = 0;
  }
  const int nn = s.nnb[i];
  const int* row = s.nbr + (size_t)i * kCap;
  for (int kk = g; kk < nn; kk += K) {
    const int j = row[kk];
    const double nj = s.ne[j];
    if (nj == 0.0) continue;
    const double4 qj = s.qs[j];
    const double sj = qj.w;
    const int spj = s.spc[j];
    const int t = pair_slot(spi, spj);
    double r[3];
    min_image(box, qi, qj, r);
    double psi0[9], ps[9], Tj[9];
    {
      double u, w;
      uw_val(T, t, sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]), u, w);
      psi_from(u, w, r, psi0);
    }
#pragma unroll
    for (int c = 0; c < 9; c++) { G0[c] += nj * psi0[c]; Tj[c] = 0.0; }
    const double nj2 = nj * nj;
#pragma unroll
    for (int l = 0; l < 3; l++) {
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        const double sh = sg ? -3.0 : 3.0;
        double p[3] = {r[0], r[1], r[2]};
        p[l] += sh * si;
        double u, w;
        uw_val(T, t, sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]), u, w);
        psi_from(u, w, p, ps);
#pragma unroll
        for (int c = 0; c < 9; c++) GA[2 * l + sg][c] += nj * ps[c];
        selfA += nj2 * dotM(ps, ps);
        double pb[3] = {r[0], r[1], r[2]};
        pb[l] += sh * sj;
        uw_val(T, t, sqrt(pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2]), u, w);
        psi_from(u, w, pb, ps);
#pragma unroll
        for (int c = 0; c < 9; c++) Tj[c] += ps[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 9; c++) U[c] += nj * Tj[c];
    TPsi += nj2 * dotM(Tj, psi0);
    double Phi = 0, R = 0;
    const double* phc = T.ph;
    const double* rhc = T.rh + (size_t)spj * T.nr * 8;
#pragma unroll
    for (int wsel = 0; wsel < 2; wsel++) {
      const double sd = kSqrt6 * (wsel ? sj : si);
#pragma unroll
      for (int l = 0; l < 3; l++) {
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
          double p[3] = {r[0], r[1], r[2]};
          p[l] += sg ? -sd : sd;
          const double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
          double pp;
          const int m = seg_index(pr, T.dxr, T.inv_dxr, T.nr, pp);
          const double* c = T.uw + ((size_t)(t * T.nr + m) << 4);
          psi_from(cval(c, pp), cval(c + 8, pp), p, ps);
          Cc += nj2 * dotM(ps, ps);
          Phi += cval(phc + ((size_t)(t * T.nr + m) << 3), pp);
          R += cval(rhc + ((size_t)m << 3), pp);
        }
      }
    }
    rho += nj * R * (1.0 / 12.0);
    Ep += nj * Phi * (1.0 / 12.0);
  }
  if (K > 1) {
    rho = group_sum<K>(rho); Ep = group_sum<K>(Ep); Cc = group_sum<K>(Cc);
    selfA = group_sum<K>(selfA); TPsi = group_sum<K>(TPsi);
#pragma unroll
    for (int c = 0; c < 9; c++) {
      G0[c] = group_sum<K>(G0[c]);
      U[c] = group_sum<K>(U[c]);
#pragma unroll
      for (int ls = 0; ls < 6; ls++) GA[ls][c] = group_sum<K>(GA[ls][c]);
    }
  }
  double GG = 0;
#pragma unroll
  for (int ls = 0; ls < 6; ls++) GG += dotM(GA[ls], GA[ls]);
  const double At = (GG - selfA + 2.0 * dotM(U, G0) - 2.0 * TPsi) * (1.0 / 18.0) + Cc * (1.0 / 12.0);
  double pe;
  const int me = seg_index(rho, T.dxe[spi], T.inv_dxe[spi], T.ne[spi], pe);
  const double* ce = T.em + ((size_t)spi * T.ne_stride + me) * 8;
  const double F = cval(ce, pe), Fp = cder(ce, pe);
  double tmp[9];
#pragma unroll
  for (int ls = 0; ls < 6; ls++) {
    applyM(1.0 / 9.0, GA[ls], tmp);
#pragma unroll
    for (int c = 0; c < 9; c++)
      if (((ls * 9 + c) % K) == g) rec[kRecLA + ls * 9 + c] = tmp[c];
  }
  applyM(1.0 / 9.0, U, tmp);
#pragma unroll
  for (int c = 0; c < 9; c++)
    if (((kRecLU + c) % K) == g) rec[kRecLU + c] = tmp[c];
  applyM(1.0 / 9.0, G0, tmp);
#pragma unroll
  for (int c = 0; c < 9; c++)
    if (((kRecLG0 + c) % K) == g) rec[kRecLG0 + c] = tmp[c];
  if (g == 0) {
    rec[kRecFP] = Fp;
    rec[kRecDVDN] = F + 0.5 * Ep + 2.0 * ni * At;
    rec[kRecRHO] = rho;
    o.V0[i] = ni * F + 0.5 * ni * Ep + ni * ni * At;
    o.rho[i] = rho;
  }
}

// ---------------------------------------------------------------------------
// Phase 2: gradient of sum_c beta_c V0_c with respect to (q_u, sigma_u, x_u).
// Role (a): V0_u (u is the centre); role (b): V0_k of each neighbour k (u is a
// neighbour).  Both roles are evaluated on the same 25 points
// p = r_uk + shift; their adjoints are combined before one gradient per point.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void k_phase2(State s, Box box, Tables T, Outputs o) {
  const int nblk = gridDim.x;
  const int blk = xcd_block(blockIdx.x, nblk);
  const int gt = blk * blockDim.x + threadIdx.x;
  const int u = gt / K;
  const int g = gt % K;
  if (u >= s.n) return;
  const double nu = s.ne[u];
  if (nu == 0.0) {
    if (g == 0) {
      o.Rq[3 * u] = 0.0; o.Rq[3 * u + 1] = 0.0; o.Rq[3 * u + 2] = 0.0;
      o.Rs[u] = 0.0; o.Rx[u] = 0.0; o.S0[u] = 0.0;
    }
    return;
  }
  const double4 qu = s.qs[u];
  const double su = qu.w;
  const int spu = s.spc[u];
  const double bu = s.beta[u];
  const double* Ru = o.rec + (size_t)u * kRec;
  const double FPu = Ru[kRecFP];
  double q0 = 0, q1 = 0, q2 = 0, sacc = 0, xacc = 0;
  const int nn = s.nnb[u];
  const int* row = s.nbr + (size_t)u * kCap;
  for (int kk = g; kk < nn; kk += K) {
    const int k = row[kk];
    const double nk = s.ne[k];
    if (nk == 0.0) continue;
    const double4 qk = s.qs[k];
    const double sk = qk.w;
    const int spk = s.spc[k];
    const double bk = s.beta[k];
    const double* Rk = o.rec + (size_t)k * kRec;
    const double wa = bu * nu * nu * nk, wb = bk * nk * nk * nu;
    const double xb = bk * nk * nk;
    const int t = pair_slot(spu, spk);
    double r[3];
    min_image(box, qu, qk, r);
    double psi0[9], ps[9], Tuk[9], SA[9], la[9], lb[9], L[9], tmp[9];
    const double r0 = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double u0, du0, w0, dw0;
    uw_vald(T, t, r0, u0, du0, w0, dw0);
    psi_from(u0, w0, r, psi0);
#pragma unroll
    for (int c = 0; c < 9; c++) { Tuk[c] = 0.0; SA[c] = 0.0; }
    // j-shift points of centre u (shift sigma_k) == i-shift points (l,-s) of k
#pragma unroll
    for (int l = 0; l < 3; l++) {
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        const double sh = sg ? -3.0 : 3.0;
        double p[3] = {r[0], r[1], r[2]};
        p[l] += sh * sk;
        const double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        double uu, du, ww, dw, g3[3];
        uw_vald(T, t, pr, uu, du, ww, dw);
        psi_from(uu, ww, p, ps);
#pragma unroll
        for (int c = 0; c < 9; c++) Tuk[c] += ps[c];
        applyM(-nk * (1.0 / 9.0), psi0, tmp);
#pragma unroll
        for (int c = 0; c < 9; c++) la[c] = Ru[kRecLG0 + c] + tmp[c];
        flip9(Rk + kRecLA + 9 * (2 * l + (1 - sg)), lb);
        applyM(-nu * (1.0 / 9.0), ps, tmp);
#pragma unroll
        for (int c = 0; c < 9; c++) {
          lb[c] += tmp[c];
          L[c] = wa * la[c] + wb * lb[c];
        }
        grad_psi(L, p, 1.0 / pr, uu, du, ww, dw, g3);
        q0 += g3[0]; q1 += g3[1]; q2 += g3[2];
        xacc += xb * dot9(lb, ps);
      }
    }
    // i-shift points of centre u (shift sigma_u) == j-shift points of k
#pragma unroll
    for (int l = 0; l < 3; l++) {
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        const double sh = sg ? -3.0 : 3.0;
        double p[3] = {r[0], r[1], r[2]};
        p[l] += sh * su;
        const double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        double uu, du, ww, dw, g3[3];
        uw_vald(T, t, pr, uu, du, ww, dw);
        psi_from(uu, ww, p, ps);
#pragma unroll
        for (int c = 0; c < 9; c++) SA[c] += ps[c];
        applyM(-nk * (1.0 / 9.0), ps, tmp);
#pragma unroll
        for (int c = 0; c < 9; c++) la[c] = Ru[kRecLA + 9 * (2 * l + sg) + c] + tmp[c];
        flip9(Rk + kRecLG0, lb);
        applyM(-nu * (1.0 / 9.0), psi0, tmp);
#pragma unroll
        for (int c = 0; c < 9; c++) {
          lb[c] += tmp[c];
          L[c] = wa * la[c] + wb * lb[c];
        }
        grad_psi(L, p, 1.0 / pr, uu, du, ww, dw, g3);
        q0 += g3[0]; q1 += g3[1]; q2 += g3[2];
        sacc += sh * g3[l];
        xacc += xb * dot9(lb, ps);
      }
    }
    // unshifted point
    {
      double g3[3];
      applyM(-nk * (1.0 / 9.0), Tuk, tmp);
#pragma unroll
      for (int c = 0; c < 9; c++) la[c] = Ru[kRecLU + c] + tmp[c];
      flip9(Rk + kRecLU, lb);
      applyM(-nu * (1.0 / 9.0), SA, tmp);
#pragma unroll
      for (int c = 0; c < 9; c++) {
        lb[c] += tmp[c];
        L[c] = wa * la[c] + wb * lb[c];
      }
      grad_psi(L, r, 1.0 / r0, u0, du0, w0, dw0, g3);
      q0 += g3[0]; q1 += g3[1]; q2 += g3[2];
      xacc += xb * dot9(lb, psi0);
    }
    // sqrt(6) sigma points: coincident angular term, pair and density
    {
      const double FPk = Rk[kRecFP];
      const double cphi = 0.5 * nu * nk * (bu + bk) * (1.0 / 12.0);
      const double cra = bu * nu * FPu * nk * (1.0 / 12.0);
      const double crb = bk * nk * FPk * nu * (1.0 / 12.0);
      const double cang = (wa * nk + wb * nu) * (1.0 / 6.0);
      const double xang = xb * nu * (1.0 / 6.0);
      const double xrho = bk * nk * FPk * (1.0 / 12.0);
      const double xphi = bk * 0.5 * nk * (1.0 / 12.0);
      const double* phc = T.ph + (size_t)t * T.nr * 8;
      const double* rkc = T.rh + (size_t)spk * T.nr * 8;
      const double* ruc = T.rh + (size_t)spu * T.nr * 8;
#pragma unroll
      for (int wsel = 0; wsel < 2; wsel++) {
        const double sd = kSqrt6 * (wsel ? sk : su);
#pragma unroll
        for (int l = 0; l < 3; l++) {
#pragma unroll
          for (int sg = 0; sg < 2; sg++) {
            const double sh = sg ? -sd : sd;
            double p[3] = {r[0], r[1], r[2]};
            p[l] += sh;
            const double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            const double ir = 1.0 / pr;
            double pp;
            const int m = seg_index(pr, T.dxr, T.inv_dxr, T.nr, pp);
            const double* c = T.uw + ((size_t)(t * T.nr + m) << 4);
            const double uu = cval(c, pp), du = cder(c, pp), ww = cval(c + 8, pp), dw = cder(c + 8, pp);
            const double* cp = phc + ((size_t)m << 3);
            const double* ck = rkc + ((size_t)m << 3);
            const double* cu = ruc + ((size_t)m << 3);
            psi_from(uu, ww, p, ps);
            applyM(cang, ps, L);
            double g3[3];
            grad_psi(L, p, ir, uu, du, ww, dw, g3);
            const double cs = (cphi * cder(cp, pp) + cra * cder(ck, pp) + crb * cder(cu, pp)) * ir;
            g3[0] += cs * p[0]; g3[1] += cs * p[1]; g3[2] += cs * p[2];
            q0 += g3[0]; q1 += g3[1]; q2 += g3[2];
            if (!wsel) sacc += (sg ? -kSqrt6 : kSqrt6) * g3[l];
            xacc += xang * dotM(ps, ps) + xrho * cval(cu, pp) + xphi * cval(cp, pp);
          }
        }
      }
    }
  }
  if (K > 1) {
    q0 = group_sum<K>(q0); q1 = group_sum<K>(q1); q2 = group_sum<K>(q2);
    sacc = group_sum<K>(sacc); xacc = group_sum<K>(xacc);
  }
  if (g == 0) {
    const double xu = nu;
    const double mi = kUma * xu * T.mass[spu];
    double logZ = 3.0 * log((su * sqrt(mi / bu)) / kHbar);
    double dx = -1.5 / xu;
    if (xu < 1.0) {
      logZ += log(1.0 / (1.0 - xu));
      dx -= 1.0 / (1.0 - xu);
    }
    double gx = 0.0;
    if (spu == 1) {
      gx = s.gamma[u] * (xu > 0.9999 ? 0.9999 : xu);
      if (xu < 0.9999) dx -= s.gamma[u];
    }
    const double V0 = o.V0[u];
    o.S0[u] = -logZ + bu * (1.0 / (2.0 * bu) + V0) - gx;
    o.Rq[3 * u] = q0; o.Rq[3 * u + 1] = q1; o.Rq[3 * u + 2] = q2;
    o.Rs[u] = sacc - 3.0 / su;
    o.Rx[u] = bu * Ru[kRecDVDN] + xacc + dx;
  }
}

// ---------------------------------------------------------------------------
// Deterministic totals: per-block sums in a fixed tree, then one block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_block_sums(int n, const double* S0, const double* V0,
                                                   double* partials) {
  __shared__ double sa[256], sb[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  sa[threadIdx.x] = i < n ? S0[i] : 0.0;
  sb[threadIdx.x] = i < n ? V0[i] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sa[0];
    partials[2 * blockIdx.x + 1] = sb[0];
  }
}

__global__ __launch_bounds__(256) void k_final_sums(int nb, const double* partials, double* totals) {
  __shared__ double sa[256], sb[256];
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) {
    a += partials[2 * k];
    b += partials[2 * k + 1];
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    totals[0] = sa[0];
    totals[1] = sb[0];
  }
}

// Active ordered pairs per block (for the work count).
__global__ __launch_bounds__(256) void k_active_pairs(State s, long long* out) {
  __shared__ long long sh[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  long long c = 0;
  if (i < s.n && s.ne[i] != 0.0) {
    const int* row = s.nbr + (size_t)i * kCap;
    for (int k = 0; k < s.nnb[i]; k++) c += s.ne[row[k]] != 0.0;
  }
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------------------
// Neighbour build: counting sort of sites into cells of edge >= r_c, then a
// 27-cell scan per site and a per-row sort so rows list site indices in
// increasing order, as the reference's O(N^2) scan does (Neighbors.cpp:67-117).
// The distance test is evaluated uncontracted (no FMA) so it rounds exactly as
// a plain IEEE scan of the same expression does.
// ---------------------------------------------------------------------------
DEV int cell_coord(double q, double lo, double L, int per, double inv_cs, int nc) {
  double v = q - lo;
  if (per) {
    v = v - L * floor(v / L);
  }
  int c = (int)floor(v * inv_cs);
  c = c < 0 ? 0 : (c > nc - 1 ? nc - 1 : c);
  return c;
}

__global__ __launch_bounds__(256) void k_cell_assign(int n, Box box, CellGrid g, const double4* qs,
                                                    int* cell_of, int* count) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double4 q = qs[i];
  int cx = cell_coord(q.x, g.lo[0], box.L[0], box.per[0], g.inv_cs[0], g.nc[0]);
  int cy = cell_coord(q.y, g.lo[1], box.L[1], box.per[1], g.inv_cs[1], g.nc[1]);
  int cz = cell_coord(q.z, g.lo[2], box.L[2], box.per[2], g.inv_cs[2], g.nc[2]);
  int c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
  cell_of[i] = c;
  atomicAdd(&count[c], 1);
}

// Exclusive scan of the cell counts in one block (cells <= a few 10^5).
__global__ __launch_bounds__(1024) void k_scan_cells(int nc, const int* count, int* start, int* cursor) {
  __shared__ int part[1024];
  int t = threadIdx.x;
  int chunk = (nc + 1023) / 1024;
  int b = t * chunk, e = min(nc, b + chunk);
  int s = 0;
  for (int c = b; c < e; c++) s += count[c];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int c = b; c < e; c++) {
    start[c] = run;
    cursor[c] = run;
    run += count[c];
  }
  if (t == 1023) start[nc] = part[1023];
}

__global__ __launch_bounds__(256) void k_cell_fill(int n, const int* cell_of, int* cursor, int* atoms) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int pos = atomicAdd(&cursor[cell_of[i]], 1);
  atoms[pos] = i;
}

DEV int axis_cells(int c, int nc, int per, int* out) {
  if (per) {
    if (nc == 1) { out[0] = 0; return 1; }
    if (nc == 2) { out[0] = 0; out[1] = 1; return 2; }
    out[0] = (c + nc - 1) % nc; out[1] = c; out[2] = (c + 1) % nc;
    return 3;
  }
  int k = 0;
  for (int d = -1; d <= 1; d++) {
    int v = c + d;
    if (v >= 0 && v < nc) out[k++] = v;
  }
  return k;
}

__global__ __launch_bounds__(256) void k_build(int n, Box box, CellGrid g, double rc, const double4* qs,
                                              const int* cell_of, const int* start, const int* atoms,
                                              int* nnb, int* nbr, int* flags) {
#pragma clang fp contract(off)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double4 qi = qs[i];
  int c = cell_of[i];
  int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
  int ax[3], ay[3], az[3];
  int nx = axis_cells(cx, g.nc[0], box.per[0], ax);
  int ny = axis_cells(cy, g.nc[1], box.per[1], ay);
  int nz = axis_cells(cz, g.nc[2], box.per[2], az);
  int cnt = 0;
  int* row = nbr + (size_t)i * kCap;
  bool over = false;
  for (int a = 0; a < nz; a++)
    for (int b = 0; b < ny; b++)
      for (int d = 0; d < nx; d++) {
        int cc = (az[a] * g.nc[1] + ay[b]) * g.nc[0] + ax[d];
        for (int p = start[cc]; p < start[cc + 1]; p++) {
          int j = atoms[p];
          if (j == i) continue;
          double4 qj = qs[j];
          double v[3] = {qi.x - qj.x, qi.y - qj.y, qi.z - qj.z};
#pragma unroll
          for (int e = 0; e < 3; e++)
            if (box.per[e]) {
              double sh = box.L[e] * rint(v[e] / box.L[e]);
              v[e] = v[e] - sh;
            }
          double s0 = v[0] * v[0];
          double s1 = v[1] * v[1];
          double s2 = v[2] * v[2];
          double rr = sqrt((s0 + s1) + s2);
          if (rr <= rc) {
            if (cnt < kMaxNeigh) row[cnt] = j;
            else over = true;
            cnt++;
          }
        }
      }
  nnb[i] = cnt < kMaxNeigh ? cnt : kMaxNeigh;
  if (over) atomicOr(&flags[0], 1);
  if (cnt == 0) atomicOr(&flags[1], 1);
}

// Bitonic sort of one neighbour row per block (256 keys, padded with INT_MAX).
__global__ __launch_bounds__(256) void k_sort_rows(int n, const int* nnb, int* nbr) {
  __shared__ int key[kCap];
  int i = blockIdx.x;
  int t = threadIdx.x;
  int cnt = nnb[i];
  int* row = nbr + (size_t)i * kCap;
  key[t] = t < cnt ? row[t] : 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= kCap; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      int ixj = t ^ j;
      if (ixj > t) {
        int a = key[t], b = key[ixj];
        bool up = (t & k) == 0;
        if ((a > b) == up) { key[t] = b; key[ixj] = a; }
      }
      __syncthreads();
    }
  }
  if (t < cnt) row[t] = key[t];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_neighbors(hipStream_t st, const Box& box, const CellGrid& g, double rc, int n,
                            const double4* qs, int* cell_of, int* cell_count, int* cell_start,
                            int* cell_cursor, int* cell_atoms, int* nnb, int* nbr, int* flags) {
  int nc = g.nc[0] * g.nc[1] * g.nc[2];
  hipError_t e = hipMemsetAsync(cell_count, 0, sizeof(int) * nc, st);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(flags, 0, sizeof(int) * 2, st);
  if (e != hipSuccess) return e;
  int nb = (n + 255) / 256;
  k_cell_assign<<<nb, 256, 0, st>>>(n, box, g, qs, cell_of, cell_count);
  k_scan_cells<<<1, 1024, 0, st>>>(nc, cell_count, cell_start, cell_cursor);
  k_cell_fill<<<nb, 256, 0, st>>>(n, cell_of, cell_cursor, cell_atoms);
  k_build<<<nb, 256, 0, st>>>(n, box, g, rc, qs, cell_of, cell_start, cell_atoms, nnb, nbr, flags);
  k_sort_rows<<<n, kCap, 0, st>>>(n, nnb, nbr);
  return hipGetLastError();
}

template <int K>
static void p1(hipStream_t st, const State& s, const Box& b, const Tables& T, const Outputs& o) {
  long long threads = (long long)s.n * K;
  int nb = (int)((threads + 255) / 256);
  k_phase1<K><<<nb, 256, 0, st>>>(s, b, T, o);
}
template <int K>
static void p2(hipStream_t st, const State& s, const Box& b, const Tables& T, const Outputs& o) {
  long long threads = (long long)s.n * K;
  int nb = (int)((threads + 255) / 256);
  k_phase2<K><<<nb, 256, 0, st>>>(s, b, T, o);
}

hipError_t launch_phase1(hipStream_t st, int K, const State& s, const Box& box, const Tables& T,
                         const Outputs& o) {
  switch (K) {
    case 1: p1<1>(st, s, box, T, o); break;
    case 2: p1<2>(st, s, box, T, o); break;
    case 4: p1<4>(st, s, box, T, o); break;
    case 8: p1<8>(st, s, box, T, o); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_phase2(hipStream_t st, int K, const State& s, const Box& box, const Tables& T,
                         const Outputs& o) {
  switch (K) {
    case 1: p2<1>(st, s, box, T, o); break;
    case 2: p2<2>(st, s, box, T, o); break;
    case 4: p2<4>(st, s, box, T, o); break;
    case 8: p2<8>(st, s, box, T, o); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_totals(hipStream_t st, int n, const Outputs& o) {
  int nb = (n + 255) / 256;
  k_block_sums<<<nb, 256, 0, st>>>(n, o.S0, o.V0, o.partials);
  k_final_sums<<<1, 256, 0, st>>>(nb, o.partials, o.totals);
  return hipGetLastError();
}

hipError_t launch_active_pairs(hipStream_t st, const State& s, long long* out, int nblocks) {
  k_active_pairs<<<nblocks, 256, 0, st>>>(s, out);
  return hipGetLastError();
}

}  // namespace dmd
