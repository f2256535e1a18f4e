// dmd_hooks.hip -- per-(u, i) derivatives of V0_i at the reference's granularity.
//
// The reference evaluates the mean-field gradient one (u, i) pair at a time:
//   evaluate_DV0_i_Dq_u_MgHx      (src/ADP/MgHx-mf-V-bulk.cpp:1823-2098):
//       [dV0_i/dq_u (3), dV0_i/dsigma_u] as a Vector4d,
//   evaluate_DV0_i_Dxi_u_adp_MgHx (:2102-2325): dV0_i/dx_u,
// registered in the dmd_equations table (:46-79, src/Macros.hpp:540-798).
// The residual kernels never form these terms separately (each pair's
// contributions to both ends are combined, and weighted by beta); this kernel
// forms them for a list of (u, i) queries from site i's phase-1 record, so
// parity can be checked term by term.
//
// V0_i = n_i F(rho_i) + 1/2 n_i sum_j n_j Q[phi] + n_i^2 A_i, with the
// angular sum A_i in the factorised form of phase 1 (1/18 [sum_ls (<G_ls,G_ls>
// - sum_j n_j^2 <psi_j^ls,psi_j^ls>) + 2 <U,G0> - 2 sum_j n_j^2 <T_j,psi_j^0>]
// + sum_j n_j^2 Q[<psi_j,psi_j>]).  For the pair (i, j) the gradient in the
// relative position r = q_i - q_j at each point is the gradient of L . psi(p)
// with the adjoint L of that point (the record holds (1/9) M G_ls, (1/9) M U,
// (1/9) M G0), plus the radial pair and embedding terms on the pair-rule
// points; dV0_i/dq_i sums it over j, dV0_i/dq_j is minus it.  Widths enter
// through the points that move with each site's sigma.
//
// dV0_i/dx_u keeps the reference's form: for an Mg centre i the embedding term
// is F'(rho_i) drho_i/dx_u, without the factor x_i, and 0 for u = i (where the
// exact derivative is F(rho_i)) (:2300-2315).
//
// One wavefront per query, lanes over the neighbours of i; butterfly sums in a
// fixed order.  FP64 throughout.
#include "dmd_device.h"
#include "dmd_hooks.h"

namespace dmd {
namespace {

DEV void rec_blk(const double* R, int b, double* o) {
#pragma unroll
  for (int c = 0; c < 9; c++) o[c] = R[rec_slot(b, c)];
}

DEV void uw4(const Tables& T, int t, int m, double off, double& u, double& du, double& w,
             double& dw) {
  const double* c = T.uw + ((size_t)t * T.nr + m) * 8;
  u = cval(c, off);
  du = cder(c, off);
  w = cval(c + 4, off);
  dw = cder(c + 4, off);
}

DEV double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Contributions of the pair (i, j) to V0_i: g = dV0_i/dr (r = q_i - q_j),
// dsi = dV0_i/dsigma_i and dsj = dV0_i/dsigma_j through this pair, and
// dxj = dV0_i/dx_j in the reference's form (embedding factor embx).
template <int RULE>
DEV void pair_terms(const State& s, const Tables& T, const double* Ri, const Box& box, int i, int j,
                    double embx, double* g, double& dsi, double& dsj, double& dxj) {
  const double ni = s.ne[i], nj = s.ne[j];
  const double4 qi = s.qs[i], qj = s.qs[j];
  const double si = qi.w, sj = qj.w;
  const int spi = s.spc[i], spj = s.spc[j];
  const int t = pair_slot(spi, spj);
  const double ni2 = ni * ni, cn = -nj / 9.0;
  double rd[3];
  min_image(box, qi, qj, rd);
  PairGeo<double> geo;
  geo.init(rd);
  g[0] = g[1] = g[2] = 0.0;
  dsi = dsj = 0.0;
  double xt = 0.0, xs = 0.0;  // triplet (times n_i^2 below) and scalar parts of dxj
  double p0[3], ir0, off, psi0[9], L[9], tmp[9], blk[9], g3[3];
  int m;
  geo.point(T, 0, 0.0, p0, m, off, ir0);
  double u0, du0, w0, dw0;
  uw4(T, t, m, off, u0, du0, w0, dw0);
  psi_from(u0, w0, p0, psi0);
  // j-shift points (sigma_j): adjoint n_j (LG0 - n_j/9 M psi0)
  double Tj[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  rec_blk(Ri, kBlkLG0, blk);
  applyM(cn, psi0, tmp);
#pragma unroll 1
  for (int ls = 0; ls < 6; ls++) {
    const int l = ls >> 1;
    const double sh = (ls & 1) ? -3.0 : 3.0;
    double p[3], ir, o2, ps[9], uu, du, ww, dw;
    geo.point(T, l, sh * sj, p, m, o2, ir);
    uw4(T, t, m, o2, uu, du, ww, dw);
    psi_from(uu, ww, p, ps);
#pragma unroll
    for (int c = 0; c < 9; c++) {
      Tj[c] += ps[c];
      L[c] = nj * (blk[c] + tmp[c]);
    }
    grad_psi(L, p, ir, uu, du, ww, dw, g3);
    g[0] += ni2 * g3[0]; g[1] += ni2 * g3[1]; g[2] += ni2 * g3[2];
    dsj += ni2 * sh * g3[l];
    xt += dot9(blk, ps) + cn * dotM(psi0, ps);
  }
  // i-shift points (sigma_i), block ls: adjoint n_j (LA_ls - n_j/9 M psi)
#pragma unroll 1
  for (int ls = 0; ls < 6; ls++) {
    const int l = ls >> 1;
    const double sh = (ls & 1) ? -3.0 : 3.0;
    double p[3], ir, o2, ps[9], uu, du, ww, dw;
    geo.point(T, l, sh * si, p, m, o2, ir);
    uw4(T, t, m, o2, uu, du, ww, dw);
    psi_from(uu, ww, p, ps);
    rec_blk(Ri, ls, blk);
    applyM(cn, ps, tmp);
#pragma unroll
    for (int c = 0; c < 9; c++) L[c] = nj * (blk[c] + tmp[c]);
    grad_psi(L, p, ir, uu, du, ww, dw, g3);
    g[0] += ni2 * g3[0]; g[1] += ni2 * g3[1]; g[2] += ni2 * g3[2];
    dsi += ni2 * sh * g3[l];
    xt += dot9(blk, ps) + cn * dotM(ps, ps);
  }
  // unshifted point: adjoint n_j (LU - n_j/9 M T_j)
  rec_blk(Ri, kBlkLU, blk);
  applyM(cn, Tj, tmp);
#pragma unroll
  for (int c = 0; c < 9; c++) L[c] = nj * (blk[c] + tmp[c]);
  grad_psi(L, p0, ir0, u0, du0, w0, dw0, g3);
  g[0] += ni2 * g3[0]; g[1] += ni2 * g3[1]; g[2] += ni2 * g3[2];
  xt += dot9(blk, psi0) + cn * dotM(Tj, psi0);  // with the j-shift loop's: -2 n_j/9 <T_j, psi0>
  // pair-rule points: coincident angular term, pair potential, density
  const double Fp = Ri[kRecFP];
#pragma unroll 1
  for (int pt = 0; pt < PairRule<RULE>::NP; pt++) {
    double p[3], ir, o2, ps[9], ds_i[3], ds_j[3], uu, du, ww, dw;
    const double wt = PairRule<RULE>::point(geo, T, pt, si, sj, p, m, o2, ir, ds_i, ds_j);
    uw4(T, t, m, o2, uu, du, ww, dw);
    psi_from(uu, ww, p, ps);
    applyM(2.0 * wt * nj * nj, ps, L);
    grad_psi(L, p, ir, uu, du, ww, dw, g3);
    const double* cp = T.ph + ((size_t)t * T.nr + m) * 4;
    const double* cr = T.rh + ((size_t)spj * T.nr + m) * 4;
    const double cs = wt * (0.5 * ni * nj * cder(cp, o2) + ni * Fp * nj * cder(cr, o2)) * ir;
    double gt[3];
#pragma unroll
    for (int a = 0; a < 3; a++) gt[a] = ni2 * g3[a] + cs * p[a];
    g[0] += gt[0]; g[1] += gt[1]; g[2] += gt[2];
    dsi += gt[0] * ds_i[0] + gt[1] * ds_i[1] + gt[2] * ds_i[2];
    dsj += gt[0] * ds_j[0] + gt[1] * ds_j[1] + gt[2] * ds_j[2];
    xt += 2.0 * wt * nj * dotM(ps, ps);
    xs += wt * (0.5 * ni * cval(cp, o2) + embx * cval(cr, o2));
  }
  dxj = ni2 * xt + xs;
}

template <int RULE>
__global__ __launch_bounds__(64) void k_site_dv0(State s, Tables T, const double* __restrict__ rec,
                                                int nq, const int* __restrict__ qu,
                                                const int* __restrict__ qi,
                                                double* __restrict__ out) {
  const int q = (int)blockIdx.x;
  if (q >= nq) return;
  const int lane = (int)threadIdx.x;
  const int u = qu[q], i = qi[q];
  double* o = out + (size_t)q * 5;
  // inactive site i or u: zeros (MgHx-mf-V-bulk.cpp:1862-1863, 1891-1892, 2133, 2157)
  if (!(s.ne[i] > 0.0) || !(s.ne[u] > 0.0)) {
    if (lane < 5) o[lane] = 0.0;
    return;
  }
  const double* Ri = rec + (size_t)i * kRec;
  const Box box = s.boxes[s.inst[i]];
  const int spi = s.spc[i];
  const double ni = s.ne[i];
  // embedding factor of dV0_i/dx_j: x_i F'(rho_i) for H, F'(rho_i) for Mg
  const double embx = spi == 1 ? ni * Ri[kRecFP] : Ri[kRecFP];
  const int nn = s.nnb[i];
  const int* row = s.nbr + (size_t)i * kCap;
  double a0 = 0, a1 = 0, a2 = 0, as = 0, ax = 0, hit = 0;
  for (int kk = lane; kk < nn; kk += 64) {
    const int j = row[kk];
    if (u != i && j != u) continue;
    double g[3], dsi, dsj, dxj;
    pair_terms<RULE>(s, T, Ri, box, i, j, embx, g, dsi, dsj, dxj);
    if (u == i) {
      a0 += g[0]; a1 += g[1]; a2 += g[2]; as += dsi;
    } else {
      a0 = -g[0]; a1 = -g[1]; a2 = -g[2]; as = dsj; ax = dxj; hit = 1.0;
    }
  }
  a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2); as = wave_sum(as);
  ax = wave_sum(ax); hit = wave_sum(hit);
  if (lane == 0) {
    if (u == i) {
      // dV0_i/dx_i = F + 1/2 Ep + 2 n_i A (record), less F for an Mg centre
      double x = Ri[kRecDVDN];
      if (spi == 0) {
        double pe;
        const int me = seg_index(Ri[kRecRHO], T.dxe[0], T.inv_dxe[0], T.ne[0], pe);
        x -= cval(T.em + (size_t)me * 4, pe);
      }
      ax = x;
    } else if (hit == 0.0) {  // u is neither i nor a neighbour of i
      a0 = a1 = a2 = as = ax = 0.0;
    }
    o[0] = a0; o[1] = a1; o[2] = a2; o[3] = as; o[4] = ax;
  }
}

}  // namespace

hipError_t launch_site_dv0(hipStream_t st, int rule, const State& s, const Tables& T,
                           const double* rec, int nq, const int* qu, const int* qi,
                           double* out) {
  if (nq <= 0) return hipSuccess;
  if (rule == 6)
    k_site_dv0<6><<<nq, 64, 0, st>>>(s, T, rec, nq, qu, qi, out);
  else if (rule == 5)
    k_site_dv0<5><<<nq, 64, 0, st>>>(s, T, rec, nq, qu, qi, out);
  else if (rule == 44)
    k_site_dv0<44><<<nq, 64, 0, st>>>(s, T, rec, nq, qu, qi, out);
  else
    k_site_dv0<3><<<nq, 64, 0, st>>>(s, T, rec, nq, qu, qi, out);
  return hipGetLastError();
}

}  // namespace dmd
