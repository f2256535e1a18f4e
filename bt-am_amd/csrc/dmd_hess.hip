// dmd_hess.hip -- diagonal 3x3 blocks of the force residual's Jacobian (gfx950).
//
// D_u = sum_{i in {u} U N(u)} beta_i d2V0_i/dq_u2, the per-(u, i) Hessians of
// evaluate_D2V0_i_Dmeanq2_u_MgHx (src/ADP/MgHx-mf-V-bulk.cpp:1366-1599)
// summed with the weights of the residual R_q(u) = sum_i beta_i dV0_i/dq_u.
//
// Factorised form (FP64).  The phase-1 record of every site holds the adjoint
// blocks L of its dipole / quadrupole sums (1/9 M G_ls, 1/9 M U, 1/9 M G0).
// Phase 2 forms the gradient as dpsi/dp^T L at the 25 points p = r_uk + shift
// of each pair; here the same adjoints contract the second derivatives:
// sum_a L_a d2psi_a/dp2.  What is quadratic in features of the centre u adds
// J^T M J terms, J = dpsi/dp: 1/9 (sum_ls JG_ls^T M JG_ls + JU^T M JG0 +
// JG0^T M JU) with the Jacobian sums JG_ls = sum_k n_k J(r_uk + 3 s sigma_u
// e_l), JU = sum_k n_k sum_ls J(r_uk + 3 s sigma_k e_l), JG0 = sum_k n_k J(r_uk),
// less their j1 = j2 parts, plus 2 n_k^2 J^T M J on the coincident points.
// Everything at a pair-rule point (coincident angular term, its J^T M J, pair,
// densities) and every j1 = j2 part of one point is an isotropic tensor of the
// point, a p p^T + b I, formed from u, w and their r-derivatives.
// For a neighbour centre k only psi_u moves; the reference sums codes 4 and 8
// of the coincident triplet (k, u, u) and not the mixed codes 5 and 7
// (:1540-1570), so its 2 n_k^2 n_u^2 J^T M J is added only when `exact` is
// set (then D_u is the exact diagonal block of dR_q/dq).  Pair and density:
// radial Hessians on the pair-rule points; embedding: x F''(rho) g g^T +
// x F'(rho) d2rho for u and for every neighbour centre.
//
// One wavefront per site u; the lanes take the neighbours k = lane, lane + 64,
// ... and every sum over k closes with a butterfly reduction (the same value
// on every lane, independent of the order the lanes ran).  The work is split
// into walks over the neighbours so that little state is live at once: the
// adjoint points; the pair-rule points; the i-shift Jacobian sums, three
// shifts per walk; JU and JG0 together.  Each Jacobian sum is
// held in 20 values.  Two wavefronts per SIMD, no spills.  (In one pass the
// kernel held one wavefront per SIMD and took 7.3 ms on 54,000 sites.)
#include "dmd_device.h"

namespace dmd {
namespace {

// Spline second derivative 2c + 6d p of a coefficient row a b c d.
DEV double cder2(const double* c, double p) { return 2.0 * c[2] + 6.0 * c[3] * p; }

// u, u', u'', w, w', w'' of pair table t at segment m, offset off.
DEV void uw_at(const Tables& T, int t, int m, double off, double* uv) {
  const double* c = T.uw + ((size_t)t * T.nr + m) * 8;
  uv[0] = cval(c, off); uv[1] = cder(c, off); uv[2] = cder2(c, off);
  uv[3] = cval(c + 4, off); uv[4] = cder(c + 4, off); uv[5] = cder2(c + 4, off);
}

// Symmetric 3x3 blocks are kept as (xx, yy, zz, xy, xz, yz).
// H += k d2c(|p|)/dp2 from c' = d1 and c'' = d2 (ir = 1/|p|).
DEV void rad_hess(double d1, double d2, const double* p, double ir, double k, double* H) {
  const double a = k * (d2 - d1 * ir) * ir * ir, b = k * d1 * ir;
  H[0] += a * p[0] * p[0] + b; H[1] += a * p[1] * p[1] + b; H[2] += a * p[2] * p[2] + b;
  H[3] += a * p[0] * p[1]; H[4] += a * p[0] * p[2]; H[5] += a * p[1] * p[2];
}

// H += k sum_a L_a d2psi_a/dp2.  L . psi = u(r) (La . p) + w(r) Q(p) with Q the
// quadratic form of L[3..8] (gradient gQ, constant Hessian HQ).
DEV void hess_psi(const double* L, const double* p, double ir, const double* uv, double k,
                  double* H) {
  const double la = L[0] * p[0] + L[1] * p[1] + L[2] * p[2];
  const double g0 = 2.0 * L[3] * p[0] + L[6] * p[1] + L[7] * p[2];
  const double g1 = L[6] * p[0] + 2.0 * L[4] * p[1] + L[8] * p[2];
  const double g2 = L[7] * p[0] + L[8] * p[1] + 2.0 * L[5] * p[2];
  const double Q = 0.5 * (p[0] * g0 + p[1] * g1 + p[2] * g2);
  rad_hess(la * uv[1] + Q * uv[4], la * uv[2] + Q * uv[5], p, ir, k, H);
  const double cu = k * uv[1] * ir, cw = k * uv[4] * ir, w = k * uv[3];
  H[0] += 2.0 * (cu * p[0] * L[0] + cw * p[0] * g0 + w * L[3]);
  H[1] += 2.0 * (cu * p[1] * L[1] + cw * p[1] * g1 + w * L[4]);
  H[2] += 2.0 * (cu * p[2] * L[2] + cw * p[2] * g2 + w * L[5]);
  H[3] += cu * (p[0] * L[1] + L[0] * p[1]) + cw * (p[0] * g1 + g0 * p[1]) + w * L[6];
  H[4] += cu * (p[0] * L[2] + L[0] * p[2]) + cw * (p[0] * g2 + g0 * p[2]) + w * L[7];
  H[5] += cu * (p[1] * L[2] + L[1] * p[2]) + cw * (p[1] * g2 + g1 * p[2]) + w * L[8];
}

// A^T M B for 9 x 3 Jacobians, M the form above written out: weights 1/2 on
// the dipole and diagonal-quadrupole rows, 1 on the off-diagonal rows, and
// -1/6 trA trB from the trace coupling.
DEV double mdot(const double* A, const double* B, int b, int g) {
  const double tA = A[9 + b] + A[12 + b] + A[15 + b], tB = B[9 + g] + B[12 + g] + B[15 + g];
  double h = A[b] * B[g];
#pragma unroll
  for (int a = 1; a < 6; a++) h += A[3 * a + b] * B[3 * a + g];
  double f = A[18 + b] * B[18 + g];
  f += A[21 + b] * B[21 + g];
  f += A[24 + b] * B[24 + g];
  return 0.5 * h + f - (1.0 / 6.0) * tA * tB;
}
// H += k (A^T M A) (symmetric).
DEV void jtmj_self(const double* A, double k, double* H) {
  H[0] += k * mdot(A, A, 0, 0); H[1] += k * mdot(A, A, 1, 1); H[2] += k * mdot(A, A, 2, 2);
  H[3] += k * mdot(A, A, 0, 1); H[4] += k * mdot(A, A, 0, 2); H[5] += k * mdot(A, A, 1, 2);
}
// H += k (A^T M B + B^T M A).
DEV void jtmj_sym(const double* A, const double* B, double k, double* H) {
  H[0] += 2.0 * k * mdot(A, B, 0, 0); H[1] += 2.0 * k * mdot(A, B, 1, 1);
  H[2] += 2.0 * k * mdot(A, B, 2, 2);
  H[3] += k * (mdot(A, B, 0, 1) + mdot(A, B, 1, 0));
  H[4] += k * (mdot(A, B, 0, 2) + mdot(A, B, 2, 0));
  H[5] += k * (mdot(A, B, 1, 2) + mdot(A, B, 2, 1));
}

// L = wa Ru[bu] + wb flip(Rk[bk]) + M v: the combined adjoint of one point
// (centre u's block, neighbour k's block with the dipole part negated since
// psi(-p) flips it, and the j1 = j2 corrections v).
DEV void adjoint(const double* Ru, int bu, double wa, const double* Rk, int bk, double wb,
                 const double* v, double* L) {
  applyM(1.0, v, L);
  L[0] += wa * Ru[8 * bu] - wb * Rk[8 * bk];
  L[1] += wa * Ru[8 * bu + 1] - wb * Rk[8 * bk + 1];
  L[2] += wa * Ru[8 * bu + 2] - wb * Rk[8 * bk + 2];
#pragma unroll
  for (int c = 3; c < 8; c++) L[c] += wa * Ru[8 * bu + c] + wb * Rk[8 * bk + c];
  L[8] += wa * Ru[kRecTail + bu] + wb * Rk[kRecTail + bk];
}

// At a pair-rule point the coincident angular term's adjoint is the point's own
// L = k M psi(p), and both of its Hessian terms are isotropic tensors of p
// (a p p^T + b I).  With u, w and their r-derivatives at r = |p|:
//   J^T M J = gam p p^T + alp I,
//     alp = u^2/2 + w^2 r^2,
//     gam = u u'/r + u'^2/2 + (w'^2 r^2 + 4 w w' r + w^2)/3;
//   sum_a (M psi)_a d2psi_a/dp2 = (1/2) d2<psi, M psi>/dp2 - J^T M J
//                               = ang_a p p^T + ang_b I  (<psi, M psi> = u^2 r^2/2 + w^2 r^4/3),
//     ang_a = u (u'' + u'/r)/2 + w (w + w' r + w'' r^2/3),
//     ang_b = u u' r/2 + w r^2 (w' r - w)/3.
struct SelfTerms {
  double alp, gam, anga, angb;
};
DEV SelfTerms self_terms(const double* uv, double r, double ir) {
  const double u = uv[0], du = uv[1], d2u = uv[2], w = uv[3], dw = uv[4], d2w = uv[5];
  const double wr = w * r, dwr = dw * r;
  SelfTerms o;
  o.alp = 0.5 * u * u + wr * wr;
  o.gam = u * du * ir + 0.5 * du * du + (dwr * dwr + 4.0 * w * dwr + w * w) * (1.0 / 3.0);
  o.anga = 0.5 * u * (d2u + du * ir) + w * (w + dwr + (1.0 / 3.0) * d2w * r * r);
  o.angb = 0.5 * u * du * r + (1.0 / 3.0) * wr * r * (dwr - w);
  return o;
}

// H += a p p^T + b I.
DEV void iso_add(const double* p, double a, double b, double* H) {
  H[0] += a * p[0] * p[0] + b; H[1] += a * p[1] * p[1] + b; H[2] += a * p[2] * p[2] + b;
  H[3] += a * p[0] * p[1]; H[4] += a * p[0] * p[2]; H[5] += a * p[1] * p[2];
}

// A weighted sum of Jacobians sum_k n_k J(p_k), held in 20 values instead of
// 27: J(p) = [u I + U p p^T ; W p_g p_i p_j + w (d_ig p_j + d_jg p_i)] (U = u'/r,
// W = w'/r), so the sum is fixed by sum n u, sum n U p p^T (6), the symmetric
// third-order tensor sum n W p p p (10) and the vector sum n w p (3).
struct JSum {
  double a[20];  // u | U pp: xx yy zz xy xz yz | W ppp: xxx yyy zzz xxy xxz xyy xzz yyz yzz xyz | w p
  DEV void zero() {
#pragma unroll
    for (int c = 0; c < 20; c++) a[c] = 0.0;
  }
  DEV void add(double n, const double* p, double ir, const double* uv) {
    const double nu = n * uv[0], nU = n * uv[1] * ir, nW = n * uv[4] * ir, nw = n * uv[3];
    const double xx = p[0] * p[0], yy = p[1] * p[1], zz = p[2] * p[2];
    const double xy = p[0] * p[1], xz = p[0] * p[2], yz = p[1] * p[2];
    a[0] += nu;
    a[1] += nU * xx; a[2] += nU * yy; a[3] += nU * zz;
    a[4] += nU * xy; a[5] += nU * xz; a[6] += nU * yz;
    const double wx = nW * p[0], wy = nW * p[1], wz = nW * p[2];
    a[7] += wx * xx; a[8] += wy * yy; a[9] += wz * zz;
    a[10] += wy * xx; a[11] += wz * xx; a[12] += wx * yy;
    a[13] += wx * zz; a[14] += wz * yy; a[15] += wy * zz; a[16] += wx * yz;
    a[17] += nw * p[0]; a[18] += nw * p[1]; a[19] += nw * p[2];
  }
  // The 27 values S[3 a + g] = sum n dpsi_a/dp_g.
  DEV void expand(double* S) const {
    // index of U pp (i, j) and of W ppp (i, j, g) in a[]
    constexpr int u2[3][3] = {{1, 4, 5}, {4, 2, 6}, {5, 6, 3}};
    constexpr int t3[3][3][3] = {{{7, 10, 11}, {10, 12, 16}, {11, 16, 13}},
                                 {{10, 12, 16}, {12, 8, 14}, {16, 14, 15}},
                                 {{11, 16, 13}, {16, 14, 15}, {13, 15, 9}}};
    constexpr int qi[6] = {0, 1, 2, 0, 0, 1}, qj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int g = 0; g < 3; g++) S[3 * r + g] = a[u2[r][g]] + (r == g ? a[0] : 0.0);
#pragma unroll
    for (int m = 0; m < 6; m++)
#pragma unroll
      for (int g = 0; g < 3; g++)
        S[3 * (3 + m) + g] = a[t3[qi[m]][qj[m]][g]] + (qi[m] == g ? a[17 + qj[m]] : 0.0) +
                             (qj[m] == g ? a[17 + qi[m]] : 0.0);
  }
};

DEV double emb_d2(const Tables& T, int sp, double rho) {
  double pe;
  const int me = seg_index(rho, T.dxe[sp], T.inv_dxe[sp], T.ne[sp], pe);
  return cder2(T.em + ((size_t)sp * T.ne_stride + me) * 4, pe);
}

template <int RULE>
__global__ __launch_bounds__(64, 2) void k_diag(State s, Tables T, const double* __restrict__ rec,
                                             int base, int exact, int spc_mask,
                                             double* __restrict__ D) {
  const int u = base + (int)blockIdx.x;
  const int lane = (int)threadIdx.x;
  double* Du = D + (size_t)blockIdx.x * 9;
  const double nu = s.ne[u];
  // inactive site, or a species whose blocks are not asked for (wave-uniform)
  if (!(nu > 0.0) || !((spc_mask >> s.spc[u]) & 1)) {
    if (lane < 9) Du[lane] = 0.0;
    return;
  }
  const Box box = s.boxes[s.inst[u]];
  const double4 qu = s.qs[u];
  const double su = qu.w, bu = s.beta[u];
  const int spu = s.spc[u];
  const int nn = s.nnb[u];
  const int* row = s.nbr + (size_t)u * kCap;
  const double* Ru = rec + (size_t)u * kRec;

  double H[6] = {0, 0, 0, 0, 0, 0}, HJ[6] = {0, 0, 0, 0, 0, 0}, gu[3] = {0, 0, 0};

  // ---- pass A: adjoint-contracted Hessians at the 13 shifted / unshifted points ----
#pragma unroll 1
  for (int kk = lane; kk < nn; kk += 64) {
    const int k = row[kk];
    const double nk = s.ne[k];
    if (!(nk > 0.0)) continue;
    const double4 qk = s.qs[k];
    const double sk = qk.w, bk = s.beta[k];
    const int t = pair_slot(spu, s.spc[k]);
    const double* Rk = rec + (size_t)k * kRec;
    const double wa = bu * nu * nu * nk, wb = bk * nk * nk * nu;
    const double ca = -wa * nk / 9.0, cb = -wb * nu / 9.0;
    double rd[3];
    min_image(box, qu, qk, rd);
    PairGeo<double> geo;
    geo.init(rd);
    double p[3], ir, off, uv[6], ps[9], psi0[9], L[9], v[9];
    double Tuk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, SA[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int m;
    // the unshifted point, kept for its own contraction at the end
    double p0[3], ir0, uv0[6];
    geo.point(T, 0, 0.0, p0, m, off, ir0);
    uw_at(T, t, m, off, uv0);
    psi_from(uv0[0], uv0[3], p0, psi0);
    // j-shift points of centre u (shift sigma_k) = i-shift points (l, -s) of k
#pragma unroll 1
    for (int ls = 0; ls < 6; ls++) {
      const int l = ls >> 1, sg = ls & 1;
      geo.point(T, l, (sg ? -3.0 : 3.0) * sk, p, m, off, ir);
      uw_at(T, t, m, off, uv);
      psi_from(uv[0], uv[3], p, ps);
#pragma unroll
      for (int c = 0; c < 9; c++) {
        Tuk[c] += ps[c];
        v[c] = ca * psi0[c] + cb * ps[c];
      }
      adjoint(Ru, kBlkLG0, wa, Rk, 2 * l + (1 - sg), wb, v, L);
      hess_psi(L, p, ir, uv, 1.0, H);
    }
    // i-shift points of centre u (shift sigma_u) = j-shift points of k
#pragma unroll 1
    for (int ls = 0; ls < 6; ls++) {
      const int l = ls >> 1, sg = ls & 1;
      geo.point(T, l, (sg ? -3.0 : 3.0) * su, p, m, off, ir);
      uw_at(T, t, m, off, uv);
      psi_from(uv[0], uv[3], p, ps);
#pragma unroll
      for (int c = 0; c < 9; c++) {
        SA[c] += ps[c];
        v[c] = ca * ps[c] + cb * psi0[c];
      }
      adjoint(Ru, ls, wa, Rk, kBlkLG0, wb, v, L);
      hess_psi(L, p, ir, uv, 1.0, H);
    }
    // unshifted point
#pragma unroll
    for (int c = 0; c < 9; c++) v[c] = ca * Tuk[c] + cb * SA[c];
    adjoint(Ru, kBlkLU, wa, Rk, kBlkLU, wb, v, L);
    hess_psi(L, p0, ir0, uv0, 1.0, H);
  }

  // ---- pass B: pair-rule points: coincident angular term, pair and density ----
#pragma unroll 1
  for (int kk = lane; kk < nn; kk += 64) {
    const int k = row[kk];
    const double nk = s.ne[k];
    if (!(nk > 0.0)) continue;
    const double4 qk = s.qs[k];
    const double sk = qk.w, bk = s.beta[k];
    const int spk = s.spc[k];
    const int t = pair_slot(spu, spk);
    const double* Rk = rec + (size_t)k * kRec;
    double rd[3];
    min_image(box, qu, qk, rd);
    PairGeo<double> geo;
    geo.init(rd);
    const double cphi = 0.5 * nu * nk * (bu + bk);
    const double cra = bu * nu * Ru[kRecFP] * nk;
    const double crb = bk * nk * Rk[kRecFP] * nu;
    const double cang = 2.0 * (bu * nu * nu * nk * nk + bk * nk * nk * nu * nu);
    const double cjj = exact ? 2.0 * bk * nk * nk * nu * nu : 0.0;
    double gk[3] = {0, 0, 0};
#pragma unroll 1
    for (int pt = 0; pt < PairRule<RULE>::NP; pt++) {
      double p[3], ir, off, uv[6], dsi[3], dsj[3];
      int m;
      const double wt = PairRule<RULE>::point(geo, T, pt, su, sk, p, m, off, ir, dsi, dsj);
      uw_at(T, t, m, off, uv);
      const double r = (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) * ir;
      const SelfTerms st = self_terms(uv, r, ir);
      const double* f = T.pr + ((size_t)t * T.nr + m) * 16;
      const double* fk = f + 4 + 4 * spk;
      const double* fu = f + 4 + 4 * spu;
      const double dk = cder(fk, off), du = cder(fu, off);
      // pair and both densities: radial Hessians k (c'' - c'/r)/r^2 p p^T + k c'/r I
      const double kp = wt * cphi, kk1 = wt * cra, ku = wt * crb;
      const double s1 = (kp * cder(f, off) + kk1 * dk + ku * du) * ir;
      const double s2 = kp * cder2(f, off) + kk1 * cder2(fk, off) + ku * cder2(fu, off);
      // angular: wt cang (M psi)-contracted Hessian; J^T M J with wt cjj to the
      // mixed codes 5 + 7 of (k, u, u) (cjj = 0 unless exact) and 2 wt n_k^2 to
      // the centre's terms
      const double kang = wt * cang, c2 = wt * cjj, c1 = 2.0 * wt * nk * nk;
      iso_add(p, kang * st.anga + c2 * st.gam + (s2 - s1) * ir * ir,
              kang * st.angb + c2 * st.alp + s1, H);
      iso_add(p, c1 * st.gam, c1 * st.alp, HJ);
      const double a1 = wt * nk * dk * ir, b1 = wt * nu * du * ir;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        gu[a] += a1 * p[a];
        gk[a] += b1 * p[a];
      }
    }
    const double ck = bk * nk * emb_d2(T, spk, Rk[kRecRHO]);
    H[0] += ck * gk[0] * gk[0]; H[1] += ck * gk[1] * gk[1]; H[2] += ck * gk[2] * gk[2];
    H[3] += ck * gk[0] * gk[1]; H[4] += ck * gk[0] * gk[2]; H[5] += ck * gk[1] * gk[2];
  }

  // ---- Jacobian sums of centre u ----
  // JG_ls three blocks (three of the six shifts, one walk) at a time, with
  // their j1 = j2 parts; then JU and JG0 in one walk, closing JU^T M JG0 +
  // JG0^T M JU, with the j1 = j2 part of that product from the same points.
  double Hw[6] = {0, 0, 0, 0, 0, 0};  // wave-uniform
#pragma unroll 1
  for (int w3 = 0; w3 < 2; w3++) {
    JSum Z[3];
    Z[0].zero();
    Z[1].zero();
    Z[2].zero();
#pragma unroll 1
    for (int kk = lane; kk < nn; kk += 64) {
      const int k = row[kk];
      const double nk = s.ne[k];
      if (!(nk > 0.0)) continue;
      const double4 qk = s.qs[k];
      const int t = pair_slot(spu, s.spc[k]);
      double rd[3];
      min_image(box, qu, qk, rd);
      PairGeo<double> geo;
      geo.init(rd);
      const double cj = -nk * nk / 9.0;
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const int ls = 3 * w3 + b;
        double p[3], ir, off, uv[6];
        int m;
        geo.point(T, ls >> 1, ((ls & 1) ? -3.0 : 3.0) * su, p, m, off, ir);
        uw_at(T, t, m, off, uv);
        Z[b].add(nk, p, ir, uv);
        // the j1 = j2 part: J^T M J of one point, isotropic
        const SelfTerms st = self_terms(uv, (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) * ir, ir);
        iso_add(p, cj * st.gam, cj * st.alp, HJ);
      }
    }
#pragma unroll
    for (int b = 0; b < 3; b++) {
#pragma unroll
      for (int c = 0; c < 20; c++) Z[b].a[c] = group_sum<64>(Z[b].a[c]);
      double S[27];
      Z[b].expand(S);
      jtmj_self(S, 1.0 / 9.0, Hw);
    }
  }
  {
    JSum ZU, Z0;
    ZU.zero();
    Z0.zero();
#pragma unroll 1
    for (int kk = lane; kk < nn; kk += 64) {
      const int k = row[kk];
      const double nk = s.ne[k];
      if (!(nk > 0.0)) continue;
      const double4 qk = s.qs[k];
      const int t = pair_slot(spu, s.spc[k]);
      double rd[3], p[3], ir, off, uv[6];
      int m;
      min_image(box, qu, qk, rd);
      PairGeo<double> geo;
      geo.init(rd);
      // -n_k^2/9 (J1^T M J0 + J0^T M J1) summed over this pair's six j-shift
      // points p1 against its unshifted point p0, in closed form: with
      // c = p1.p0, U = u'/r, W = w'/r, t = W r^2 + 2 w (index 1 at p1, 0 at p0),
      //   J1^T M J0 + J0^T M J1 = 2 a I + 2 b1 p1 p1^T + 2 b0 p0 p0^T + d (p1 p0^T + p0 p1^T),
      //   a = u1 u0/2 + w1 w0 c,  b1 = U1 u0/2 + W1 w0 c,  b0 = u1 U0/2 + w1 W0 c,
      //   d = U1 U0 c/2 + W1 W0 c^2/2 - t1 t0/6 + w1 w0.
      double p0[3], ir0, off0, uv0[6];
      geo.point(T, 0, 0.0, p0, m, off0, ir0);
      uw_at(T, t, m, off0, uv0);
      Z0.add(nk, p0, ir0, uv0);
      const double u0 = uv0[0], U0 = uv0[1] * ir0, w0 = uv0[3], W0 = uv0[4] * ir0;
      const double t0 = W0 * (p0[0] * p0[0] + p0[1] * p0[1] + p0[2] * p0[2]) + 2.0 * w0;
      const double k2 = -2.0 * nk * nk / 9.0;
      double b0s = 0.0, q[3] = {0, 0, 0};
#pragma unroll 1
      for (int ls = 0; ls < 6; ls++) {
        geo.point(T, ls >> 1, ((ls & 1) ? -3.0 : 3.0) * qk.w, p, m, off, ir);
        uw_at(T, t, m, off, uv);
        ZU.add(nk, p, ir, uv);
        const double u1 = uv[0], U1 = uv[1] * ir, w1 = uv[3], W1 = uv[4] * ir;
        const double c = p[0] * p0[0] + p[1] * p0[1] + p[2] * p0[2];
        const double t1 = W1 * (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) + 2.0 * w1;
        const double a = 0.5 * u1 * u0 + w1 * w0 * c;
        const double b1 = 0.5 * U1 * u0 + W1 * w0 * c;
        b0s += 0.5 * u1 * U0 + w1 * W0 * c;
        const double d = (0.5 * U1 * U0 + 0.5 * W1 * W0 * c) * c - (1.0 / 6.0) * t1 * t0 + w1 * w0;
        iso_add(p, k2 * b1, k2 * a, HJ);
        q[0] += d * p[0]; q[1] += d * p[1]; q[2] += d * p[2];
      }
      const double kh = 0.5 * k2;
      iso_add(p0, k2 * b0s, 0.0, HJ);
      HJ[0] += k2 * q[0] * p0[0]; HJ[1] += k2 * q[1] * p0[1]; HJ[2] += k2 * q[2] * p0[2];
      HJ[3] += kh * (q[0] * p0[1] + p0[0] * q[1]);
      HJ[4] += kh * (q[0] * p0[2] + p0[0] * q[2]);
      HJ[5] += kh * (q[1] * p0[2] + p0[1] * q[2]);
    }
#pragma unroll
    for (int c = 0; c < 20; c++) {
      ZU.a[c] = group_sum<64>(ZU.a[c]);
      Z0.a[c] = group_sum<64>(Z0.a[c]);
    }
    double JU[27], J0[27];
    ZU.expand(JU);
    Z0.expand(J0);
    jtmj_sym(JU, J0, 1.0 / 9.0, Hw);
  }

  // ---- close ----
#pragma unroll
  for (int c = 0; c < 6; c++) {
    H[c] = group_sum<64>(H[c]);
    HJ[c] = group_sum<64>(HJ[c]);
  }
#pragma unroll
  for (int a = 0; a < 3; a++) gu[a] = group_sum<64>(gu[a]);
  if (lane == 0) {
    const double cu = bu * nu * emb_d2(T, spu, Ru[kRecRHO]);
    const double w2 = bu * nu * nu;
    double O[6];
#pragma unroll
    for (int c = 0; c < 6; c++) O[c] = H[c] + w2 * (HJ[c] + Hw[c]);
    O[0] += cu * gu[0] * gu[0]; O[1] += cu * gu[1] * gu[1]; O[2] += cu * gu[2] * gu[2];
    O[3] += cu * gu[0] * gu[1]; O[4] += cu * gu[0] * gu[2]; O[5] += cu * gu[1] * gu[2];
    Du[0] = O[0]; Du[1] = O[3]; Du[2] = O[4];
    Du[3] = O[3]; Du[4] = O[1]; Du[5] = O[5];
    Du[6] = O[4]; Du[7] = O[5]; Du[8] = O[2];
  }
}

}  // namespace

hipError_t launch_diag(hipStream_t st, int rule, int exact, int spc_mask, const State& s,
                       const Tables& T, const double* rec, int base, int n, double* D) {
  if (n <= 0) return hipSuccess;
  if (rule == 6)
    hipLaunchKernelGGL(k_diag<6>, dim3(n), dim3(64), 0, st, s, T, rec, base, exact, spc_mask, D);
  else if (rule == 5)
    hipLaunchKernelGGL(k_diag<5>, dim3(n), dim3(64), 0, st, s, T, rec, base, exact, spc_mask, D);
  else if (rule == 44)
    hipLaunchKernelGGL(k_diag<44>, dim3(n), dim3(64), 0, st, s, T, rec, base, exact, spc_mask, D);
  else
    hipLaunchKernelGGL(k_diag<3>, dim3(n), dim3(64), 0, st, s, T, rec, base, exact, spc_mask, D);
  return hipGetLastError();
}

}  // namespace dmd
