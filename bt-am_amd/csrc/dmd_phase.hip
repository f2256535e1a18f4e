// dmd_phase.hip -- phase-1 / phase-2 kernels of the mean-field DMD residual.
//
// Work decomposition: in k_phase1, K consecutive lanes of a wavefront own one
// site; each lane walks every K-th neighbour of the site and the K partial sums
// are combined with a butterfly of cross-lane shuffles.  The pair kernels
// (k_phase1_pairs, k_phase2) evaluate each unordered pair once, at its owner,
// from per-chunk lists of owned pairs, and write both ends' contributions into
// their neighbour-row slots; the site sums read the slots in row order.  All
// reductions run in a fixed order, so a site's result does not depend on what
// else is in flight.
//
// The per-point quantities are the reference's Gaussian-averaged ADP terms
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
//
// Precision: P is the type of the quadrature-point evaluation (splines,
// feature vectors, adjoint combination, gradients).  Per-site sums, records,
// the embedding, entropy terms and totals are FP64 for both P.
#include "dmd_device.h"

namespace dmd {

// ---------------------------------------------------------------------------
// Phase 1: per-site sums.  For site i:
//   G0 = sum_j n_j psi(r_ij),  G_ls = sum_j n_j psi(r_ij + 3 s sigma_i e_l),
//   U = sum_j n_j T_j with T_j = sum_ls psi(r_ij + 3 s sigma_j e_l),
//   rho_bar, pair sum, coincident and self terms on the sqrt(6) sigma points.
// Writes V0_i, rho_bar_i and the adjoint record of site i.
//
// The neighbour row is walked 6 / kShiftsPerWalk times for the i-shift points
// (l, s), each walk accumulating only its kShiftsPerWalk G_ls (9 values each),
// and once more for the other 19 points of each pair.  Re-walking the row
// (L1/L2-resident neighbour data plus the pair geometry) costs far less than
// keeping all 54 G_ls accumulators live, which would leave room for a single
// wavefront per SIMD.
// ---------------------------------------------------------------------------
template <typename P> struct PrecCode;
template <> struct PrecCode<double> { static constexpr int v = kFP64; };
template <> struct PrecCode<float> { static constexpr int v = kSplit32; };

// Site and lane of this thread: the block maps onto the precision's list of
// chunks to evaluate (K blocks per 256-site chunk), in XCD-contiguous order.
template <typename P, int K>
DEV bool site_of(const State& s, int& i, int& g) {
  const int* act = s.act + PrecCode<P>::v * s.act_stride;
  const int nb = act[0] * K;
  if ((int)blockIdx.x >= nb) return false;  // uniform per block
  const int lb = xcd_block(blockIdx.x, nb);
  i = act[1 + lb / K] * kChunk + (lb % K) * (256 / K) + (int)threadIdx.x / K;
  g = (int)threadIdx.x % K;
  return i < s.n;
}

// i-shift points per walk of the neighbour row in phase 1 (1, 2, 3 or 6;
// measured on MI355X at the 54k-site workload: 3 is fastest in both precisions):
// fewer walks recompute the pair geometry less often, more accumulators live.
constexpr int kShiftsPerWalk = 3;

// Cooperative gathers (dmd_device.h) per phase and precision, as measured on
// MI355X at the 54k-site workload: phase 1 gains in both precisions; phase 2
// gains in FP64 only (its FP32 body is already VALU-heavy).
template <typename P> constexpr bool kCoop1 = true;
template <typename P> constexpr bool kCoop2 = sizeof(P) == 8;
// Fused pair-rule rows (phi | rho_Mg | rho_H) through the cooperative gather
// in phase 2, in both precisions (FP32 phase 2: 0.79 -> 0.75 ms).
template <typename P> constexpr bool kCoopF2 = true;
// The pair kernels step through their pair lists in groups of this many
// consecutive lanes (the largest cooperative gather group: 128-B fused rows in
// FP64), so every gather group runs each step together.
constexpr int kPairGroup = 8;


// Pair-rule points of phase 1, once per unordered pair: their sums are
// symmetric in the pair (C6 = sum w <psi,psi>, Phi = sum w phi) except the
// densities, one per end (rho of the other end's species), and both come from
// the same points.  The owner writes (C6, Phi, density at that end) into the
// slots of both ends (the slots are free until phase 2); k_phase1 adds them up
// in row order.  Work layout as in phase 2 (per-chunk owned-pair lists).
// VALUES = false (relaxation ticks with the residual-norm line search): the
// densities only -- one density row (rho_Mg | rho_H, the same coefficients as
// the fused row's) per point, no u/w or phi row, no C6 or Phi.
template <typename P, int B, int RULE, bool VALUES>
DEV void phase1_pairs(const State& s, const Tables& T, const Outputs& o, int inst, int ch,
                      int part, u32x4* stage) {
  constexpr bool CO = kCoop1<P>;
  const Box box = s.boxes[inst];
  const P* uwt = tab_uw<P>(T);
  const P* prt = tab_pr<P>(T);
  const P* rrt = tab_rr<P>(T);
  const int nr = T.nr;
  const int c0 = ch * kChunk;
  const int np = s.chunk_np[ch];
  const int* own = s.own + (size_t)ch * kChunk * kCap;
  const int tq = (int)threadIdx.x & (kPairGroup - 1);
  for (int p0 = part * 256 + ((int)threadIdx.x & ~(kPairGroup - 1)); p0 < np; p0 += B * 256) {
    const bool inb = p0 + tq < np;  // lanes past the end repeat the group's first pair
    const int ent = own[inb ? p0 + tq : p0];
    const int u = c0 + (ent & 255), kk = (ent >> 8) & 255, pos = ent >> 16;
    const int k = s.nbr[(size_t)u * kCap + kk];
    const double4 qu = s.qs[u], qk = s.qs[k];
    const int spu = s.spc[u], spk = s.spc[k];
    const int t = pair_slot(spu, spk);
    double rd[3];
    min_image(box, qu, qk, rd);
    PairGeo<P> geo;
    geo.init(rd);
    const P su = (P)qu.w, sk = (P)qk.w;
    double C6 = 0, Phi = 0, Ru = 0, Rk = 0;  // per-pair sums in FP64
#pragma unroll 1
    for (int pt = 0; pt < PairRule<RULE>::NP; pt++) {
      P p[3], off, ir, dsi[3], dsj[3], cw[8], f[16], cr[4];
      int m;
      const double wt = PairRule<RULE>::point(geo, T, pt, su, sk, p, m, off, ir, dsi, dsj);
      if constexpr (VALUES) {
        gather_cpt<P, CO, CO, true>(uwt, t * nr + m, prt, t * nr + m, stage, cw, f);
        const P r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        C6 += wt * (double)self_psi(cval(cw, off), cval(cw + 4, off), r2);
        Phi += wt * (double)cval(f, off);
        fused_rho(f, spk, cr);
        Ru += wt * (double)cval(cr, off);
        fused_rho(f, spu, cr);
        Rk += wt * (double)cval(cr, off);
      } else {
        gather_row8<P, CO>(rrt, m, stage, cw);
        P ck[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          ck[e] = spk ? cw[4 + e] : cw[e];
          cr[e] = spu ? cw[4 + e] : cw[e];
        }
        Ru += wt * (double)cval(ck, off);
        Rk += wt * (double)cval(cr, off);
      }
    }
    if (inb) {
      double* eu = o.slot + ((size_t)u * kCap + kk) * kSlotW;
      double* ek = o.slot + ((size_t)k * kCap + pos) * kSlotW;
      eu[0] = C6; eu[1] = Phi; eu[2] = Ru;
      ek[0] = C6; ek[1] = Phi; ek[2] = Rk;
    }
  }
}

template <typename P, int B, int RULE>
__global__ __launch_bounds__(256) void k_phase1_pairs(State s, Tables T, Outputs o) {
  constexpr bool CO = kCoop1<P>;
  __shared__ u32x4 stage_all[4][CptRounds<P, CO, CO>::v * kStageSlots];
  u32x4* stage = stage_all[threadIdx.x >> 6];
  const int* act = s.act + PrecCode<P>::v * s.act_stride;
  const int nb = act[0] * B;
  if ((int)blockIdx.x >= nb) return;  // uniform per block
  const int lb = xcd_block(blockIdx.x, nb);
  const int ch = act[1 + lb / B];
  const int inst = s.inst[ch * kChunk];
  if (inst < 0 || s.prec[inst] != PrecCode<P>::v) return;
  if (s.noxr && (s.noxr[inst] & kSkipValues))
    phase1_pairs<P, B, RULE, false>(s, T, o, inst, ch, lb % B, stage);
  else
    phase1_pairs<P, B, RULE, true>(s, T, o, inst, ch, lb % B, stage);
}

// VALUES = false (relaxation ticks with the residual-norm line search): the
// adjoint records, F'(rho_bar) and rho_bar only -- no V0 (its coincident,
// pair and self terms are left out).
template <typename P, int K, int RULE, bool VALUES>
DEV void phase1_site(const State& s, const Tables& T, const Outputs& o, int i, int g, int inst,
                     u32x4* stage) {
  constexpr bool CO = kCoop1<P>;
  const Box box = s.boxes[inst];
  double* rec = o.rec + (size_t)i * kRec;
  float* rec32 = o.rec32 + (size_t)i * kRec32;
  constexpr bool kF32 = PrecCode<P>::v == kSplit32;
  const double ni = s.ne[i];
  if (ni == 0.0) {  // all K lanes of the site leave together
    for (int c = g; c < kRec; c += K) rec[c] = 0.0;
    if (kF32)
      for (int c = g; c < kRec32; c += K) rec32[c] = 0.0f;
    if (g == 0) { o.V0[i] = 0.0; o.rho[i] = 0.0; }
    return;
  }
  const P* uwt = tab_uw<P>(T);
  const int nr = T.nr;
  const double4 qi = s.qs[i];
  const P si = (P)qi.w;
  const int spi = s.spc[i];
  const int nn = s.nnb[i];
  const int* row = s.nbr + (size_t)i * kCap;
  double tmp[9];

  // i-shift points, kShiftsPerWalk of the six (l, s) per walk of the row
  double GG = 0, selfA = 0;
#pragma unroll 1
  for (int ls0 = 0; ls0 < 6; ls0 += kShiftsPerWalk) {
    double G[kShiftsPerWalk][9];
#pragma unroll
    for (int b = 0; b < kShiftsPerWalk; b++)
#pragma unroll
      for (int c = 0; c < 9; c++) G[b][c] = 0;
    for (int kk0 = 0; kk0 < nn; kk0 += K) {  // uniform over the site's lanes
      const int kk = kk0 + g;
      const bool inb = kk < nn;
      const int j = inb ? row[kk] : i;
      const double nj = inb ? s.ne[j] : 0.0;
      const double4 qj = s.qs[j];
      const int rb = pair_slot(spi, s.spc[j]) * nr;
      double rd[3];
      min_image(box, qi, qj, rd);
      idle_disp(inb, rd);
      PairGeo<P> geo;
      geo.init(rd);
      if constexpr (CO) {  // the walk's rows in one gather
        P p[kShiftsPerWalk][3], off[kShiftsPerWalk], cw[kShiftsPerWalk][8];
        int rows[kShiftsPerWalk];
#pragma unroll
        for (int b = 0; b < kShiftsPerWalk; b++) {
          const int ls = ls0 + b;
          const P sh = (ls & 1 ? P(-3) : P(3)) * si;
          P ir;
          int m;
          geo.point(T, ls >> 1, sh, p[b], m, off[b], ir);
          rows[b] = rb + m;
        }
        gather_rows8<P, kShiftsPerWalk, CO>(uwt, rows, stage, cw);
#pragma unroll
        for (int b = 0; b < kShiftsPerWalk; b++) {
          P ps[9];
          const P uu = cval(cw[b], off[b]), ww = cval(cw[b] + 4, off[b]);
          psi_from(uu, ww, p[b], ps);
#pragma unroll
          for (int e = 0; e < 9; e++) G[b][e] += nj * (double)ps[e];
          if constexpr (VALUES) {
            const P r2 = p[b][0] * p[b][0] + p[b][1] * p[b][1] + p[b][2] * p[b][2];
            selfA += nj * nj * (double)self_psi(uu, ww, r2);
          }
        }
      } else {
#pragma unroll
        for (int b = 0; b < kShiftsPerWalk; b++) {
          const int ls = ls0 + b;
          const P sh = (ls & 1 ? P(-3) : P(3)) * si;
          P p[3], off, ir, ps[9];
          int m;
          geo.point(T, ls >> 1, sh, p, m, off, ir);
          const P* c = uwt + ((size_t)(rb + m) << 3);
          const P uu = cval(c, off), ww = cval(c + 4, off);
          psi_from(uu, ww, p, ps);
#pragma unroll
          for (int e = 0; e < 9; e++) G[b][e] += nj * (double)ps[e];
          if constexpr (VALUES)
            selfA += nj * nj * (double)self_psi(uu, ww, p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        }
      }
    }
#pragma unroll
    for (int b = 0; b < kShiftsPerWalk; b++) {
      const int ls = ls0 + b;
      if (K > 1) {
#pragma unroll
        for (int c = 0; c < 9; c++) G[b][c] = group_sum<K>(G[b][c]);
      }
      if constexpr (VALUES) GG += dotM(G[b], G[b]);
      applyM(1.0 / 9.0, G[b], tmp);
#pragma unroll
      for (int c = 0; c < 9; c++)
        if (((ls * 9 + c) % K) == g) {
          if (kF32) rec32[rec_slot(ls, c)] = (float)tmp[c];
          else rec[rec_slot(ls, c)] = tmp[c];
        }
    }
  }

  // unshifted, j-shift and pair-rule points
  double rho = 0, Ep = 0, Cc = 0, TPsi = 0;
  double G0[9], U[9];
#pragma unroll
  for (int c = 0; c < 9; c++) { G0[c] = 0; U[c] = 0; }
  for (int kk0 = 0; kk0 < nn; kk0 += K) {  // uniform over the site's lanes
      const int kk = kk0 + g;
    const bool inb = kk < nn;
    const int j = inb ? row[kk] : i;
    const double nj = inb ? s.ne[j] : 0.0;
    const double4 qj = s.qs[j];
    const P sj = (P)qj.w;
    const int spj = s.spc[j];
    const int t = pair_slot(spi, spj);
    const int rb = t * nr;
    double rd[3];
    min_image(box, qi, qj, rd);
    idle_disp(inb, rd);
    PairGeo<P> geo;
    geo.init(rd);
    P p[3], off, ir, psi0[9], ps[9], Tj[9], cw[8];
    int m;
    geo.point(T, 0, P(0), p, m, off, ir);
    gather_row8<P, CO>(uwt, rb + m, stage, cw);
    psi_from(cval(cw, off), cval(cw + 4, off), p, psi0);
#pragma unroll
    for (int c = 0; c < 9; c++) { G0[c] += nj * (double)psi0[c]; Tj[c] = P(0); }
    const double nj2 = nj * nj;
#pragma unroll 1
    for (int pt = 0; pt < 6; pt++) {
      geo.point(T, pt >> 1, (pt & 1 ? P(-3) : P(3)) * sj, p, m, off, ir);
      gather_row8<P, CO>(uwt, rb + m, stage, cw);
      psi_from(cval(cw, off), cval(cw + 4, off), p, ps);
#pragma unroll
      for (int e = 0; e < 9; e++) Tj[e] += ps[e];
    }
#pragma unroll
    for (int c = 0; c < 9; c++) U[c] += nj * (double)Tj[c];
    if constexpr (VALUES) TPsi += nj2 * (double)dotM(Tj, psi0);
    // pair and coincident points of the pair rule: per-pair sums from k_phase1_pairs
    if (inb) {
      const double* e = o.slot + ((size_t)i * kCap + kk) * kSlotW;
      if constexpr (VALUES) {
        Cc += nj2 * e[0];
        Ep += nj * e[1];
      }
      rho += nj * e[2];
    }
  }
  if (K > 1) {
    rho = group_sum<K>(rho);
    if constexpr (VALUES) {
      Ep = group_sum<K>(Ep); Cc = group_sum<K>(Cc);
      selfA = group_sum<K>(selfA); TPsi = group_sum<K>(TPsi);
    }
#pragma unroll
    for (int c = 0; c < 9; c++) {
      G0[c] = group_sum<K>(G0[c]);
      U[c] = group_sum<K>(U[c]);
    }
  }
  const double At = (GG - selfA + 2.0 * dotM(U, G0) - 2.0 * TPsi) * (1.0 / 18.0) + Cc;
  double pe;
  const int me = seg_index(rho, T.dxe[spi], T.inv_dxe[spi], T.ne[spi], pe);
  const double* ce = T.em + ((size_t)spi * T.ne_stride + me) * 4;
  const double F = cval(ce, pe), Fp = cder(ce, pe);
  applyM(1.0 / 9.0, U, tmp);
#pragma unroll
  for (int c = 0; c < 9; c++)
    if (((kBlkLU * 9 + c) % K) == g) {
      if (kF32) rec32[rec_slot(kBlkLU, c)] = (float)tmp[c];
      else rec[rec_slot(kBlkLU, c)] = tmp[c];
    }
  applyM(1.0 / 9.0, G0, tmp);
#pragma unroll
  for (int c = 0; c < 9; c++)
    if (((kBlkLG0 * 9 + c) % K) == g) {
      if (kF32) rec32[rec_slot(kBlkLG0, c)] = (float)tmp[c];
      else rec[rec_slot(kBlkLG0, c)] = tmp[c];
    }
  if (g == 0) {
    rec[kRecFP] = Fp;
    rec[kRecDVDN] = F + 0.5 * Ep + 2.0 * ni * At;
    rec[kRecRHO] = rho;
    o.V0[i] = ni * F + 0.5 * ni * Ep + ni * ni * At;
    o.rho[i] = rho;
  }
}

template <typename P, int K, int RULE>
__global__ __launch_bounds__(256) void k_phase1(State s, Tables T, Outputs o) {
  static_assert(K >= 4, "a site's lanes must cover whole gather groups");
  constexpr bool CO = kCoop1<P>;
  // largest gather: the rows of one i-shift walk
  constexpr int kRounds = StageRounds<P, kShiftsPerWalk, CO>::v;
  __shared__ u32x4 stage_all[4][kRounds * kStageSlots];
  u32x4* stage = stage_all[threadIdx.x >> 6];
  int i, g;
  if (!site_of<P, K>(s, i, g)) return;
  const int inst = s.inst[i];
  if (inst < 0 || s.prec[inst] != PrecCode<P>::v) return;  // uniform per block
  if (s.noxr && (s.noxr[inst] & kSkipValues))
    phase1_site<P, K, RULE, false>(s, T, o, i, g, inst, stage);
  else
    phase1_site<P, K, RULE, true>(s, T, o, i, g, inst, stage);
}

// Block b of a record as P, negating the dipole part when flip is set.
template <typename S, typename P>
DEV void load_blk(const S* r, int b, bool flip, P* o) {
  const S* m = r + 8 * b;
  const P sgn = flip ? P(-1) : P(1);
  o[0] = sgn * (P)m[0]; o[1] = sgn * (P)m[1]; o[2] = sgn * (P)m[2];
#pragma unroll
  for (int c = 3; c < 8; c++) o[c] = (P)m[c];
  o[8] = (P)r[kRecTail + b];
}

// The adjoint blocks of site i in the precision phase 2 reads: the FP64 record
// for FP64, the FP32 copy for the split path (adj_base + adj_rows: the same in
// 8-value rows).
DEV const double* adj_rec(const Outputs& o, int i, double) { return o.rec + (size_t)i * kRec; }
DEV const float* adj_rec(const Outputs& o, int i, float) { return o.rec32 + (size_t)i * kRec32; }
DEV const double* adj_base(const Outputs& o, double) { return o.rec; }
DEV const float* adj_base(const Outputs& o, float) { return o.rec32; }
template <typename P> constexpr int adj_rows() { return sizeof(P) == 8 ? kRec / 8 : kRec32 / 8; }

// The 9 components of a block from its 8-value row and its tail value.
template <typename P>
DEV void blk_from(const P* row, P tail, bool flip, P* o) {
  const P sgn = flip ? P(-1) : P(1);
  o[0] = sgn * row[0]; o[1] = sgn * row[1]; o[2] = sgn * row[2];
#pragma unroll
  for (int c = 3; c < 8; c++) o[c] = row[c];
  o[8] = tail;
}

// Block b of neighbour k's record (k differs between lanes): a cooperative
// gather of its row plus the tail value.  ib: the adjoint rows of k's instance
// (inst_adj), k0: that instance's first site.
template <typename P, bool CO>
DEV void nbr_blk(const Outputs& o, const P* ib, int k0, int k, int b, bool flip, u32x4* stage,
                 P* out) {
  const P tail = adj_rec(o, k, P(0))[kRecTail + b];
  const int r = (k - k0) * adj_rows<P>() + b;
  P row[1][8];
  gather8<P, 1, CO>(&ib, &r, stage, row);
  blk_from(row[0], tail, flip, out);
}

// The adjoint rows of the instance whose first site is k0: row offsets from
// here stay below 4 GiB (kMaxInstanceSites), as the row gathers need.
template <typename P>
DEV const P* inst_adj(const Outputs& o, int k0) {
  return adj_base(o, P(0)) + (size_t)k0 * adj_rows<P>() * 8;
}

// ---------------------------------------------------------------------------
// Phase 2: gradient of sum_c beta_c V0_c with respect to (q, sigma, x).
// Role (a): V0_u (u is the centre); role (b): V0_k of each neighbour k (u is a
// neighbour).  Both roles are evaluated on the same 25 points
// p = r_uk + shift; their adjoints are combined before one gradient per point.
// psi(-p) = flip(psi(p)) (the dipole part is odd, the quadrupole part even).
//
// The points of pair (u, k) seen from k are those seen from u mirrored
// (p -> -p), with the same combined adjoint, so each unordered pair is
// evaluated once, by its owner (pair_owner): k_phase2 writes the pair's
// contributions to both ends into their row slots -- position gradient +g at
// u and -g at k, each end's width derivative from the points that shift with
// its own sigma, and each end's occupancy derivative from the other end's
// adjoints -- and k_phase2_sum adds up each site's slots in row order.  No
// atomics: every slot has one writer and the sums run in a fixed order.
// ---------------------------------------------------------------------------
// Returns x unchanged, through an empty asm statement the compiler cannot see
// into, so a value recomputed from it is not merged with an earlier copy.
DEV double kept_out(double x) { asm volatile("" : "+v"(x)); return x; }
DEV float kept_out(float x) { asm volatile("" : "+v"(x)); return x; }

// XR: also form the occupancy terms (instances that relax positions and widths
// only skip them; their R_x output is zero).
//
// Work: the owned pairs of a chunk form one list; the chunk's B blocks stride
// through it, each group of kPairGroup lanes taking as many consecutive pairs
// (the cooperative gathers need the lanes of a group together).  A pair with
// an inactive end has exact-zero contributions at the active end (every
// weight carries both occupancies); the inactive end reads no slots.
template <typename P, int B, bool UNR, int RULE, bool XR>
DEV void phase2_pairs(const State& s, const Tables& T, const Outputs& o, int inst, int ch,
                      int part, u32x4* stage) {
  constexpr bool CO = kCoop2<P>;
  const Box box = s.boxes[inst];
  const P* uwt = tab_uw<P>(T);
  const P* prt = tab_pr<P>(T);
  const P* adjI = inst_adj<P>(o, box.base);
  const int nr = T.nr;
  const int c0 = ch * kChunk;
  // Mg positions and widths only: the list's H-H tail adds to nothing needed
  const int np = s.noxr && (s.noxr[inst] & kSkipHH) ? s.chunk_npx[ch] : s.chunk_np[ch];
  const int* own = s.own + (size_t)ch * kChunk * kCap;
  const int tq = (int)threadIdx.x & (kPairGroup - 1);
  for (int p0 = part * 256 + ((int)threadIdx.x & ~(kPairGroup - 1)); p0 < np; p0 += B * 256) {
    const bool inb = p0 + tq < np;  // lanes past the end repeat the group's first pair
    const int ent = own[inb ? p0 + tq : p0];
    const int u = c0 + (ent & 255), kk = (ent >> 8) & 255, pos = ent >> 16;
    const double nu = s.ne[u];
    const double4 qu = s.qs[u];
    const double su = qu.w;
    const int spu = s.spc[u];
    const double bu = s.beta[u];
    const double* Ru = o.rec + (size_t)u * kRec;
    const P* Au = adj_rec(o, u, P(0));
    const double FPu = Ru[kRecFP];
    const int k = s.nbr[(size_t)u * kCap + kk];
    const double nk = s.ne[k];
    const double4 qk = s.qs[k];
    const double sk = qk.w;
    const int spk = s.spc[k];
    const double bk = s.beta[k];
    const double* Rk = o.rec + (size_t)k * kRec;
    const P* Ak = adj_rec(o, k, P(0));
    const double wa = bu * nu * nu * nk, wb = bk * nk * nk * nu;
    const double xb = bk * nk * nk, xbk = bu * nu * nu;  // occupancy weights, ends u and k
    const P waP = (P)wa, wbP = (P)wb;
    const P cnk = P(-nk * (1.0 / 9.0)), cnu = P(-nu * (1.0 / 9.0));
    const int t = pair_slot(spu, spk);
    const int rb = t * nr;
    double q0 = 0, q1 = 0, q2 = 0, sgu = 0, sgk = 0, xu = 0, xk = 0;  // this pair
    double rd[3];
    min_image(box, qu, qk, rd);
    PairGeo<P> geo;
    geo.init(rd);
    P pz[3], ir0, off, tmp[9];
    int m;
    geo.point(T, 0, P(0), pz, m, off, ir0);
    P cz[8];
    gather_row8<P, CO>(uwt, rb + m, stage, cz);
    const P u0 = cval(cz, off), du0 = cder(cz, off), w0 = cval(cz + 4, off), dw0 = cder(cz + 4, off);
    // The two shifted-point walks each hold one constant adjoint and one
    // accumulator: the j-shift walk the role-(a) adjoint laB and the sum Tuk of
    // psi over it, the i-shift walk the role-(b) adjoint lbA and the sum SA.
    // The unshifted point's gradient is linear in its adjoint, so its role-(a)
    // half is taken after the first walk and its role-(b) half after the
    // second; psi at the unshifted point is formed again where it is needed
    // (kept_out hides the earlier copy from CSE so it is not held in registers
    // across a walk).
    P laB[9], Tuk[9];
    {
      P psi0[9];
      psi_from(u0, w0, pz, psi0);
      load_blk(Au, kBlkLG0, false, laB);
      applyM(cnk, psi0, tmp);
#pragma unroll
      for (int c = 0; c < 9; c++) { laB[c] += tmp[c]; Tuk[c] = P(0); }  // unscaled by waP
    }
    const P skP = (P)sk, suP = (P)su;

    // j-shift points of centre u (shift sigma_k) == i-shift points (l,-s) of k
    auto bpt = [&](int pt) {
      const int l = pt >> 1, sg = pt & 1;
      const P sh = sg ? P(-3) : P(3);
      P p[3], ir, o2, ps[9], lb[9], L[9], g3[3], c[2][8];
      int mm;
      geo.point(T, l, sh * skP, p, mm, o2, ir);
      const int blk = 2 * l + (1 - sg);
      const P tail = Ak[kRecTail + blk];
      const int r8[2] = {rb + mm, (k - box.base) * adj_rows<P>() + blk};
      const P* t8[2] = {uwt, adjI};
      gather8<P, 2, CO>(t8, r8, stage, c);
      blk_from(c[1], tail, true, lb);
      const P uu = cval(c[0], o2), du = cder(c[0], o2), ww = cval(c[0] + 4, o2),
              dw = cder(c[0] + 4, o2);
      psi_from(uu, ww, p, ps);
#pragma unroll
      for (int e = 0; e < 9; e++) Tuk[e] += ps[e];
      P mps[9];
      applyM(cnu, ps, mps);
#pragma unroll
      for (int e = 0; e < 9; e++) {
        lb[e] += mps[e];
        L[e] = waP * laB[e] + wbP * lb[e];
      }
      grad_psi(L, p, ir, uu, du, ww, dw, g3);
      q0 += (double)g3[0]; q1 += (double)g3[1]; q2 += (double)g3[2];
      const P gl = l == 0 ? g3[0] : (l == 1 ? g3[1] : g3[2]);
      sgk += (double)(sh * gl);
      if (XR) {
        xu += xb * (double)dot9(lb, ps);
        xk += xbk * (double)dot9(laB, ps);
      }
    };
    if (UNR) {
#pragma unroll 2
      for (int pt = 0; pt < 6; pt++) bpt(pt);
    } else {
#pragma unroll 1
      for (int pt = 0; pt < 6; pt++) bpt(pt);
    }
    // unshifted point, role (a): waP * (LU_u - nk/9 M Tuk)
    {
      P la[9], g3[3];
      load_blk(Au, kBlkLU, false, la);
      applyM(cnk, Tuk, tmp);
#pragma unroll
      for (int c = 0; c < 9; c++) la[c] += tmp[c];
      if (XR) {
        P psi0[9];
        psi_from(kept_out(u0), w0, pz, psi0);
        xk += xbk * (double)dot9(la, psi0);
      }
#pragma unroll
      for (int c = 0; c < 9; c++) la[c] *= waP;
      grad_psi(la, pz, ir0, u0, du0, w0, dw0, g3);
      q0 += (double)g3[0]; q1 += (double)g3[1]; q2 += (double)g3[2];
    }
    P lbA[9], SA[9];
    {
      P psi0[9];
      psi_from(kept_out(u0), w0, pz, psi0);
      nbr_blk<P, CO>(o, adjI, box.base, k, kBlkLG0, true, stage, lbA);
      applyM(cnu, psi0, tmp);
#pragma unroll
      for (int c = 0; c < 9; c++) { lbA[c] += tmp[c]; SA[c] = P(0); }
    }
    // i-shift points of centre u (shift sigma_u) == j-shift points of k
    auto apt = [&](int pt) {
      const int l = pt >> 1, sg = pt & 1;
      const P sh = sg ? P(-3) : P(3);
      P p[3], ir, o2, ps[9], la[9], L[9], g3[3], c[8];
      int mm;
      geo.point(T, l, sh * suP, p, mm, o2, ir);
      load_blk(Au, 2 * l + sg, false, la);
      gather_row8<P, CO>(uwt, rb + mm, stage, c);
      const P uu = cval(c, o2), du = cder(c, o2), ww = cval(c + 4, o2), dw = cder(c + 4, o2);
      psi_from(uu, ww, p, ps);
#pragma unroll
      for (int e = 0; e < 9; e++) SA[e] += ps[e];
      P mps[9];
      applyM(cnk, ps, mps);
#pragma unroll
      for (int e = 0; e < 9; e++) {
        la[e] += mps[e];
        L[e] = waP * la[e] + wbP * lbA[e];
      }
      grad_psi(L, p, ir, uu, du, ww, dw, g3);
      q0 += (double)g3[0]; q1 += (double)g3[1]; q2 += (double)g3[2];
      const P gl = l == 0 ? g3[0] : (l == 1 ? g3[1] : g3[2]);
      sgu += (double)(sh * gl);
      if (XR) {
        xu += xb * (double)dot9(lbA, ps);
        xk += xbk * (double)dot9(la, ps);
      }
    };
    if (UNR) {
#pragma unroll 2
      for (int pt = 0; pt < 6; pt++) apt(pt);
    } else {
#pragma unroll 1
      for (int pt = 0; pt < 6; pt++) apt(pt);
    }
    // unshifted point, role (b): wbP * (LU_k - nu/9 M SA)
    {
      P lb[9], L[9], g3[3];
      nbr_blk<P, CO>(o, adjI, box.base, k, kBlkLU, true, stage, lb);
      applyM(cnu, SA, tmp);
#pragma unroll
      for (int c = 0; c < 9; c++) {
        lb[c] += tmp[c];
        L[c] = wbP * lb[c];
      }
      grad_psi(L, pz, ir0, u0, du0, w0, dw0, g3);
      q0 += (double)g3[0]; q1 += (double)g3[1]; q2 += (double)g3[2];
      if (XR) {
        P psi0[9];
        psi_from(kept_out(u0), w0, pz, psi0);
        xu += xb * (double)dot9(lb, psi0);
      }
    }
    // pair-rule points: coincident angular term, pair and density
    {
      const double FPk = Rk[kRecFP];
      const P cphi = (P)(0.5 * nu * nk * (bu + bk));
      const P cra = (P)(bu * nu * FPu * nk);
      const P crb = (P)(bk * nk * FPk * nu);
      const P cang = (P)(2.0 * (wa * nk + wb * nu));
      // Every term at a pair-rule point is radial (the angular one through
      // self_hgrad), so its gradient is one scalar times p.
      auto cpt = [&](int pt) {
        // explicit FMAs only: the same operations with and without XR
#pragma clang fp contract(off)
        P p[3], ir, o2, dsu[3], dsk[3], c[8], f[16], ck[4], cu[4];
        int mm;
        const double wt =
            PairRule<RULE>::point(geo, T, pt, suP, skP, p, mm, o2, ir, dsu, dsk);
        const P wtP = (P)wt;
        gather_cpt<P, CO, kCoopF2<P>>(uwt, rb + mm, prt, rb + mm, stage, c, f);
        const P* cp = f;
        fused_rho(f, spk, ck);
        fused_rho(f, spu, cu);
        const P uu = cval(c, o2), du = cder(c, o2), ww = cval(c + 4, o2), dw = cder(c + 4, o2);
        const P r2 = fma(p[2], p[2], fma(p[1], p[1], p[0] * p[0]));
        const P rad = fma(cphi, cder(cp, o2), fma(cra, cder(ck, o2), crb * cder(cu, o2)));
        const P cs = wtP * fma(cang, self_hgrad(uu, du, ww, dw, r2 * ir, r2), rad * ir);
        const P g3[3] = {cs * p[0], cs * p[1], cs * p[2]};
        q0 += (double)g3[0]; q1 += (double)g3[1]; q2 += (double)g3[2];
        sgu += (double)fma(g3[2], dsu[2], fma(g3[1], dsu[1], g3[0] * dsu[0]));
        sgk += (double)fma(g3[2], dsk[2], fma(g3[1], dsk[1], g3[0] * dsk[0]));
        if (XR) {  // angular, density and pair terms of the other end's V0
          const double pp = (double)self_psi(uu, ww, r2), hp = 0.5 * (double)cval(cp, o2);
          xu += wt * (2.0 * xb * nu * pp + bk * nk * (FPk * (double)cval(cu, o2) + hp));
          xk += wt * (2.0 * xbk * nk * pp + bu * nu * (FPu * (double)cval(ck, o2) + hp));
        }
      };
      if (UNR && PairRule<RULE>::NP <= 27) {  // 44 points unrolled by three spill
#pragma unroll 3
        for (int pt = 0; pt < PairRule<RULE>::NP; pt++) cpt(pt);
      } else {
#pragma unroll 1
        for (int pt = 0; pt < PairRule<RULE>::NP; pt++) cpt(pt);
      }
    }
    if (inb) {  // with nk = 0 the values at u are exact zeros; k reads none
      double* eu = o.slot + ((size_t)u * kCap + kk) * kSlotW;
      double* ek = o.slot + ((size_t)k * kCap + pos) * kSlotW;
      eu[0] = q0; eu[1] = q1; eu[2] = q2; eu[3] = sgu;
      ek[0] = -q0; ek[1] = -q1; ek[2] = -q2; ek[3] = sgk;
      if (XR) { eu[4] = xu; ek[4] = xk; }
    }
  }
}

// Two wavefronts per SIMD: the FP64 pair body fits 256 VGPRs without spills.
// UNR: the point loops advance two (shifted) or three (pair-rule) points per
// iteration, so one point's arithmetic overlaps the next point's gathers.
template <typename P, int B, bool UNR, int RULE>
__global__ __launch_bounds__(256, 2) void k_phase2(State s, Tables T, Outputs o) {
  // largest gathers: two 8-value rows (j-shift points), an 8- and a 16-value
  // row (pair-rule points)
  constexpr bool CO = kCoop2<P>;
  constexpr int kRounds = StageRounds<P, 2, CO>::v > CptRounds<P, CO, kCoopF2<P>>::v
                              ? StageRounds<P, 2, CO>::v : CptRounds<P, CO, kCoopF2<P>>::v;
  __shared__ u32x4 stage_all[4][kRounds * kStageSlots];
  u32x4* stage = stage_all[threadIdx.x >> 6];
  const int* act = s.act + PrecCode<P>::v * s.act_stride;
  const int nb = act[0] * B;
  if ((int)blockIdx.x >= nb) return;  // uniform per block
  const int lb = xcd_block(blockIdx.x, nb);
  const int ch = act[1 + lb / B];
  const int inst = s.inst[ch * kChunk];
  if (inst < 0 || s.prec[inst] != PrecCode<P>::v) return;
  if (s.noxr && (s.noxr[inst] & kSkipRx))
    phase2_pairs<P, B, UNR, RULE, false>(s, T, o, inst, ch, lb % B, stage);
  else
    phase2_pairs<P, B, UNR, RULE, true>(s, T, o, inst, ch, lb % B, stage);
}

// Sums of a site's slots (row order, fixed lane order) and the site's own
// terms: entropy, width entropy, occupancy terms of its own V0 and S0.
constexpr int kSumLanes = 8;
template <typename P>
__global__ __launch_bounds__(256) void k_phase2_sum(State s, Tables T, Outputs o) {
  constexpr int K = kSumLanes;
  int u, g;
  if (!site_of<P, K>(s, u, g)) return;
  const int inst = s.inst[u];
  if (inst < 0 || s.prec[inst] != PrecCode<P>::v) return;  // uniform per block
  const bool xr = !(s.noxr && (s.noxr[inst] & kSkipRx));
  // Mg positions and widths only: an H site's slots are partly unwritten
  // (its H-H pairs were skipped) and its R_q, R_sigma are not needed
  const bool hq = !(s.noxr && (s.noxr[inst] & kSkipHH) && s.spc[u] == 1);
  const double nu = s.ne[u];
  if (nu == 0.0) {
    if (g == 0) {
      o.Rq[3 * u] = 0.0; o.Rq[3 * u + 1] = 0.0; o.Rq[3 * u + 2] = 0.0;
      o.Rs[u] = 0.0; o.Rx[u] = 0.0; o.S0[u] = 0.0;
    }
    return;
  }
  const int nn = hq ? s.nnb[u] : 0;
  const double* sl = o.slot + (size_t)u * kCap * kSlotW;
  double q0 = 0, q1 = 0, q2 = 0, sacc = 0, xacc = 0;
  for (int kk = g; kk < nn; kk += K) {  // every slot is written each evaluation
    const double* e = sl + (size_t)kk * kSlotW;
    q0 += e[0]; q1 += e[1]; q2 += e[2]; sacc += e[3];
    if (xr) xacc += e[4];
  }
  q0 = group_sum<K>(q0); q1 = group_sum<K>(q1); q2 = group_sum<K>(q2);
  sacc = group_sum<K>(sacc);
  xacc = group_sum<K>(xacc);
  if (g == 0) {
    const double4 qu = s.qs[u];
    const double su = qu.w;
    const int spu = s.spc[u];
    const double bu = s.beta[u];
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
    o.Rs[u] = hq ? sacc - 3.0 / su : 0.0;
    o.Rx[u] = xr ? bu * o.rec[(size_t)u * kRec + kRecDVDN] + xacc + dx : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename P, int K, int RULE>
static void p1(hipStream_t st, const State& s, const Tables& T, const Outputs& o) {
  const long long threads = (long long)s.n * K;
  const int nb = (int)((threads + 255) / 256);
  // blocks per chunk of the pair-rule kernel: 64 (about one owned pair per
  // lane at the 54k-site workload) is 2-4 % faster than 32, 16 or 8 there
  // (measured on MI355X)
  constexpr int kB = 64;
  // The pair-rule points in their own pass, once per unordered pair, in both
  // precisions and for every rule: faster than evaluating them from both ends
  // inside k_phase1 (split residual, 12-point rule: 1.48 -> 1.43 ms at 54k
  // sites, measured on MI355X).
  k_phase1_pairs<P, kB, RULE><<<(int)((long long)s.n * kB / 256), 256, 0, st>>>(s, T, o);
  k_phase1<P, K, RULE><<<nb, 256, 0, st>>>(s, T, o);
}
template <typename P, int K, bool UNR, int RULE>
static void p2(hipStream_t st, const State& s, const Tables& T, const Outputs& o) {
  const long long threads = (long long)s.n * K;
  const int nb = (int)((threads + 255) / 256);
  k_phase2<P, K, UNR, RULE><<<nb, 256, 0, st>>>(s, T, o);
  const int nbs = (int)(((long long)s.n * kSumLanes + 255) / 256);
  k_phase2_sum<P><<<nbs, 256, 0, st>>>(s, T, o);
}

template <typename P, int RULE>
static hipError_t phase1_p(hipStream_t st, int K, const State& s, const Tables& T,
                           const Outputs& o) {
  switch (K) {
    case 4: p1<P, 4, RULE>(st, s, T, o); break;
    case 8: p1<P, 8, RULE>(st, s, T, o); break;
    case 16: p1<P, 16, RULE>(st, s, T, o); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename P, bool UNR, int RULE>
static hipError_t phase2_pu(hipStream_t st, int K, const State& s, const Tables& T,
                            const Outputs& o) {
  switch (K) {
    case 4: p2<P, 4, UNR, RULE>(st, s, T, o); break;
    case 8: p2<P, 8, UNR, RULE>(st, s, T, o); break;
    case 16: p2<P, 16, UNR, RULE>(st, s, T, o); break;
    case 32: p2<P, 32, UNR, RULE>(st, s, T, o); break;
    case 64: p2<P, 64, UNR, RULE>(st, s, T, o); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename P>
static hipError_t phase1_r(hipStream_t st, const PhaseConfig& cfg, const State& s,
                           const Tables& T, const Outputs& o) {
  const int k1 = cfg.k1[PrecCode<P>::v];
  if (cfg.rule == 6) return phase1_p<P, 6>(st, k1, s, T, o);
  if (cfg.rule == 5) return phase1_p<P, 5>(st, k1, s, T, o);
  if (cfg.rule == 44) return phase1_p<P, 44>(st, k1, s, T, o);
  return phase1_p<P, 3>(st, k1, s, T, o);
}

template <typename P>
static hipError_t phase2_r(hipStream_t st, const PhaseConfig& cfg, const State& s,
                           const Tables& T, const Outputs& o) {
  const int k2 = cfg.k2[PrecCode<P>::v];
  const bool un = cfg.unroll2[PrecCode<P>::v] != 0;
  if (cfg.rule == 6)
    return un ? phase2_pu<P, true, 6>(st, k2, s, T, o) : phase2_pu<P, false, 6>(st, k2, s, T, o);
  if (cfg.rule == 5)
    return un ? phase2_pu<P, true, 5>(st, k2, s, T, o) : phase2_pu<P, false, 5>(st, k2, s, T, o);
  if (cfg.rule == 44)
    return un ? phase2_pu<P, true, 44>(st, k2, s, T, o) : phase2_pu<P, false, 44>(st, k2, s, T, o);
  return un ? phase2_pu<P, true, 3>(st, k2, s, T, o) : phase2_pu<P, false, 3>(st, k2, s, T, o);
}

hipError_t launch_phase1(hipStream_t st, const PhaseConfig& cfg, int precision, const State& s,
                         const Tables& T, const Outputs& o) {
  if (precision == kSplit32) return phase1_r<float>(st, cfg, s, T, o);
  return phase1_r<double>(st, cfg, s, T, o);
}

hipError_t launch_phase2(hipStream_t st, const PhaseConfig& cfg, int precision, const State& s,
                         const Tables& T, const Outputs& o) {
  if (precision == kSplit32) return phase2_r<float>(st, cfg, s, T, o);
  return phase2_r<double>(st, cfg, s, T, o);
}

}  // namespace dmd
