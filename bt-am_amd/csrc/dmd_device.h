// dmd_device.h -- device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "dmd_kernels.h"

namespace dmd {

#define DEV __device__ __forceinline__

constexpr double kSqrt6 = 2.449489742783178;  // src/Macros.hpp:42
constexpr double kSqrt2 = 1.4142135623730951;  // src/Macros.hpp:40
constexpr double kUma = 103.4993333E-6;        // src/Macros.hpp:851
constexpr double kHbar = 6.5821192815E-4;      // src/Macros.hpp:847

// ---------------------------------------------------------------------------
// Splines (reference clamp, src/Numerical/cubic-spline.cpp:84-128): segment
// m = min(trunc(x/dx), n-1), offset p = min(x - m dx, dx), value
// a + (b + (c + d p) p) p, derivative b + (2c + 3d p) p.
// ---------------------------------------------------------------------------
DEV int seg_index(double x, double dx, double inv_dx, int n, double& p) {
  int m = (int)(x * inv_dx);
  m = m > n - 1 ? n - 1 : m;
  m = m < 0 ? 0 : m;
  double pp = x - (double)m * dx;
  p = pp > dx ? dx : pp;
  return m;
}

template <typename P>
DEV P cval(const P* c, P p) {
  return c[0] + (c[1] + (c[2] + c[3] * p) * p) * p;
}
template <typename P>
DEV P cder(const P* c, P p) {
  return c[1] + (P(2) * c[2] + P(3) * c[3] * p) * p;
}

template <typename P> DEV const P* tab_uw(const Tables& T);
template <> DEV const double* tab_uw<double>(const Tables& T) { return T.uw; }
template <> DEV const float* tab_uw<float>(const Tables& T) { return T.uw32; }
template <typename P> DEV const P* tab_ph(const Tables& T);
template <> DEV const double* tab_ph<double>(const Tables& T) { return T.ph; }
template <> DEV const float* tab_ph<float>(const Tables& T) { return T.ph32; }
template <typename P> DEV const P* tab_rh(const Tables& T);
template <> DEV const double* tab_rh<double>(const Tables& T) { return T.rh; }
template <> DEV const float* tab_rh<float>(const Tables& T) { return T.rh32; }
template <typename P> DEV const P* tab_pr(const Tables& T);
template <> DEV const double* tab_pr<double>(const Tables& T) { return T.pr; }
template <> DEV const float* tab_pr<float>(const Tables& T) { return T.pr32; }
template <typename P> DEV const P* tab_rr(const Tables& T);
template <> DEV const double* tab_rr<double>(const Tables& T) { return T.rr; }
template <> DEV const float* tab_rr<float>(const Tables& T) { return T.rr32; }

// ---------------------------------------------------------------------------
// Quadrature-point geometry.  A point is p = r + d e_l (one axis shifted).
// PairGeo<P>::point returns p, the spline segment m and offset of |p|, and
// 1/|p|.  FP64: |p| directly.  FP32: the pair distance |r| is kept in FP64
// and |p| = |r| + D with the small increment D = (2 r_l d + d^2)/(|p| + |r|)
// in FP32, so the segment offset (|r| - m dx) + D carries FP64-quality
// position information while every other operation is FP32.
// ---------------------------------------------------------------------------
template <typename P> struct PairGeo;

// |p| and 1/|p| from |p|^2 in FP64: the hardware reciprocal square-root
// estimate refined by Newton's iteration y <- y (3 - a y^2) / 2 (each step
// doubles the correct bits of the ~2^-23-accurate estimate, so three steps
// reach the FP64 rounding level).  One transcendental and ~10 FMA-class
// operations instead of a square root and a division.
DEV void norm_and_inv(double a, double& x, double& ir) {
  double y = __builtin_amdgcn_rsq(a);
#pragma unroll
  for (int it = 0; it < 3; it++) {
    const double h = 0.5 * a * y;
    y = fma(-h, y, 1.5) * y;
  }
  ir = y;
  x = a * y;
}

template <> struct PairGeo<double> {
  double r[3];
  DEV void init(const double* rd) { r[0] = rd[0]; r[1] = rd[1]; r[2] = rd[2]; }
  DEV void point(const Tables& T, int l, double d, double* p, int& m, double& off,
                 double& ir) const {
    p[0] = r[0] + (l == 0 ? d : 0.0);
    p[1] = r[1] + (l == 1 ? d : 0.0);
    p[2] = r[2] + (l == 2 ? d : 0.0);
    double x;
    norm_and_inv(p[0] * p[0] + p[1] * p[1] + p[2] * p[2], x, ir);
    m = seg_index(x, T.dxr, T.inv_dxr, T.nr, off);
  }
  // p = r + d for a general displacement d.
  DEV void point3(const Tables& T, const double* d, double* p, int& m, double& off,
                  double& ir) const {
    p[0] = r[0] + d[0];
    p[1] = r[1] + d[1];
    p[2] = r[2] + d[2];
    double x;
    norm_and_inv(p[0] * p[0] + p[1] * p[1] + p[2] * p[2], x, ir);
    m = seg_index(x, T.dxr, T.inv_dxr, T.nr, off);
  }
};

template <> struct PairGeo<float> {
  float r[3];
  double R;
  float R32;
  DEV void init(const double* rd) {
    R = sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
    R32 = (float)R;
    r[0] = (float)rd[0]; r[1] = (float)rd[1]; r[2] = (float)rd[2];
  }
  DEV void point(const Tables& T, int l, float d, float* p, int& m, float& off, float& ir) const {
    p[0] = r[0] + (l == 0 ? d : 0.0f);
    p[1] = r[1] + (l == 1 ? d : 0.0f);
    p[2] = r[2] + (l == 2 ? d : 0.0f);
    const float rl = l == 0 ? r[0] : (l == 1 ? r[1] : r[2]);
    const float pn = __builtin_amdgcn_sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const float D = (2.0f * rl * d + d * d) * __builtin_amdgcn_rcpf(pn + R32);
    int mm = (int)((R32 + D) * T.inv_dxr32);
    mm = mm > T.nr - 1 ? T.nr - 1 : mm;
    mm = mm < 0 ? 0 : mm;
    float o = (float)(R - (double)mm * T.dxr) + D;
    off = o > T.dxr32 ? T.dxr32 : o;
    m = mm;
    ir = __builtin_amdgcn_rcpf(pn);
  }
  // p = r + d for a general displacement d; |p| = |r| + (2 r.d + d.d)/(|p| + |r|).
  DEV void point3(const Tables& T, const float* d, float* p, int& m, float& off, float& ir) const {
    p[0] = r[0] + d[0];
    p[1] = r[1] + d[1];
    p[2] = r[2] + d[2];
    const float pn = __builtin_amdgcn_sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const float rd = r[0] * d[0] + r[1] * d[1] + r[2] * d[2];
    const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float D = (2.0f * rd + dd) * __builtin_amdgcn_rcpf(pn + R32);
    int mm = (int)((R32 + D) * T.inv_dxr32);
    mm = mm > T.nr - 1 ? T.nr - 1 : mm;
    mm = mm < 0 ? 0 : mm;
    float o = (float)(R - (double)mm * T.dxr) + D;
    off = o > T.dxr32 ? T.dxr32 : o;
    m = mm;
    ir = __builtin_amdgcn_rcpf(pn);
  }
};

// ---------------------------------------------------------------------------
// Pair (and coincident-triplet) rules.  Point t of pair (i, j) lies at
// p = r_ij + delta_t with weight w_t (returned in FP64 for both precisions, so
// the weights sum to 1 to FP64 accuracy); dsi_t = d delta_t / d sigma_i and
// dsj_t = d delta_t / d sigma_j.
//   PairRule<3>: the reference's 6D Stroud rule -- sqrt(6) sigma shifts of
//                site i (t < 6) or site j along one axis, weight 1/12.
//   PairRule<5>: 3-point Gauss-Hermite per axis on the relative coordinate
//                r ~ N(r_ij, (sigma_i^2 + sigma_j^2) I): nodes 0 and
//                +-sqrt(3) sigma_rel, weights 2/3 and 1/6 (27 points, exact to
//                degree 5).
// ---------------------------------------------------------------------------
template <int RULE> struct PairRule;
template <> struct PairRule<3> {
  static constexpr int NP = 12;
  template <typename P>
  DEV static double point(const PairGeo<P>& geo, const Tables& T, int t, P si, P sj, P* p, int& m,
                          P& off, P& ir, P* dsi, P* dsj) {
    const int wsel = t / 6, l = (t >> 1) % 3;
    const P d = (t & 1) ? P(-kSqrt6) : P(kSqrt6);
    geo.point(T, l, d * (wsel ? sj : si), p, m, off, ir);
    dsi[0] = (!wsel && l == 0) ? d : P(0);
    dsi[1] = (!wsel && l == 1) ? d : P(0);
    dsi[2] = (!wsel && l == 2) ? d : P(0);
    dsj[0] = (wsel && l == 0) ? d : P(0);
    dsj[1] = (wsel && l == 1) ? d : P(0);
    dsj[2] = (wsel && l == 2) ? d : P(0);
    return 1.0 / 12.0;
  }
};
template <> struct PairRule<5> {
  static constexpr int NP = 27;
  template <typename P>
  DEV static double point(const PairGeo<P>& geo, const Tables& T, int t, P si, P sj, P* p, int& m,
                          P& off, P& ir, P* dsi, P* dsj) {
    const P s3 = P(1.7320508075688772);
    const int a0 = t % 3, a1 = (t / 3) % 3, a2 = t / 9;
    const P x0 = a0 == 0 ? P(0) : (a0 == 1 ? s3 : -s3);
    const P x1 = a1 == 0 ? P(0) : (a1 == 1 ? s3 : -s3);
    const P x2 = a2 == 0 ? P(0) : (a2 == 1 ? s3 : -s3);
    const P srel = sqrt(si * si + sj * sj);
    const P d[3] = {srel * x0, srel * x1, srel * x2};
    geo.point3(T, d, p, m, off, ir);
    const P f = si / srel, fj = sj / srel;
    dsi[0] = f * x0; dsi[1] = f * x1; dsi[2] = f * x2;
    dsj[0] = fj * x0; dsj[1] = fj * x1; dsj[2] = fj * x2;
    const double w0 = a0 == 0 ? 2.0 / 3.0 : 1.0 / 6.0;
    const double w1 = a1 == 0 ? 2.0 / 3.0 : 1.0 / 6.0;
    const double w2 = a2 == 0 ? 2.0 / 3.0 : 1.0 / 6.0;
    return w0 * w1 * w2;
  }
};

// PairRule<6>: the 6-point Stroud rule on the relative coordinate
// r ~ N(r_ij, (sigma_i^2 + sigma_j^2) I): r +- sqrt(3) sigma_rel e_l, weight
// 1/6 (degree 3).  With sigma_i = sigma_j it is PairRule<3> (both put
// r +- sqrt(6) sigma e_l); with unequal widths it matches the exact fourth
// moment 3 sigma_rel^4 along an axis, which PairRule<3> does not.  The point
// set is symmetric, so one evaluation serves both ends.
template <> struct PairRule<6> {
  static constexpr int NP = 6;
  template <typename P>
  DEV static double point(const PairGeo<P>& geo, const Tables& T, int t, P si, P sj, P* p, int& m,
                          P& off, P& ir, P* dsi, P* dsj) {
    const int l = t >> 1;
    const P x = (t & 1) ? P(-1.7320508075688772) : P(1.7320508075688772);
    const P srel = sqrt(si * si + sj * sj);
    geo.point(T, l, x * srel, p, m, off, ir);
    const P f = x * si / srel, fj = x * sj / srel;
    dsi[0] = l == 0 ? f : P(0); dsi[1] = l == 1 ? f : P(0); dsi[2] = l == 2 ? f : P(0);
    dsj[0] = l == 0 ? fj : P(0); dsj[1] = l == 1 ? fj : P(0); dsj[2] = l == 2 ? fj : P(0);
    return 1.0 / 6.0;
  }
};

// PairRule<44>: the reference's 6D degree-5 rule for a pair
// (src/Numerical/Quadrature-Hermitian-5th.cpp:31-157; defined there, not used
// by its evaluators): 22 generators zeta in R^6 and their negatives, site
// displacement sqrt(2) sigma_site zeta_site -- zeta = (1,...,1) with weight
// 1/128, sqrt(2) e_k with 1/16 (k = 0..5), and (1,...,1) with -1 at two
// positions a < b (lexicographic) with 1/128.  The relative point is
// p = r + sqrt(2) (sigma_i zeta_i - sigma_j zeta_j).  The set maps to itself
// when the two sites' halves are swapped, so one evaluation serves both ends.
DEV double rule44_gen(int t, double* z) {
  const int g = t % 22;
  const double sg = t < 22 ? 1.0 : -1.0;
  double w;
  if (g == 0) {
#pragma unroll
    for (int k = 0; k < 6; k++) z[k] = sg;
    w = 1.0 / 128.0;
  } else if (g <= 6) {
#pragma unroll
    for (int k = 0; k < 6; k++) z[k] = (k == g - 1) ? sg * kSqrt2 : 0.0;
    w = 1.0 / 16.0;
  } else {
    int idx = g - 7, a = 0;
    while (idx >= 5 - a) {
      idx -= 5 - a;
      a++;
    }
    const int b = a + 1 + idx;
#pragma unroll
    for (int k = 0; k < 6; k++) z[k] = (k == a || k == b) ? -sg : sg;
    w = 1.0 / 128.0;
  }
  return w;
}
template <> struct PairRule<44> {
  static constexpr int NP = 44;
  template <typename P>
  DEV static double point(const PairGeo<P>& geo, const Tables& T, int t, P si, P sj, P* p, int& m,
                          P& off, P& ir, P* dsi, P* dsj) {
    double z[6];
    const double w = rule44_gen(t, z);
    P d[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      dsi[k] = (P)(kSqrt2 * z[k]);
      dsj[k] = (P)(-kSqrt2 * z[3 + k]);
      d[k] = dsi[k] * si + dsj[k] * sj;
    }
    geo.point3(T, d, p, m, off, ir);
    return w;
  }
};

// ---------------------------------------------------------------------------
// Feature vectors psi = (u p, w p p^T) as 9 values
// (a_x a_y a_z B_xx B_yy B_zz B_xy B_xz B_yz) and the bilinear form
// <X,Y> = 1/2 a.a' + 1/2 B:B' - 1/6 trB trB' of the dipole + quadrupole terms
// (src/ADP/MgHx-ADP.cpp:1233-1281, 1781-1831).
// ---------------------------------------------------------------------------
template <typename P>
DEV void psi_from(P u, P w, const P* p, P* o) {
  o[0] = u * p[0]; o[1] = u * p[1]; o[2] = u * p[2];
  const P w0 = w * p[0], w1 = w * p[1];
  o[3] = w0 * p[0]; o[4] = w1 * p[1]; o[5] = w * p[2] * p[2];
  o[6] = w0 * p[1]; o[7] = w0 * p[2]; o[8] = w1 * p[2];
}

template <typename P>
DEV P dotM(const P* x, const P* y) {
  const P trx = x[3] + x[4] + x[5], tr_y = y[3] + y[4] + y[5];
  return P(0.5) * (x[0] * y[0] + x[1] * y[1] + x[2] * y[2] + x[3] * y[3] + x[4] * y[4] + x[5] * y[5]) +
         (x[6] * y[6] + x[7] * y[7] + x[8] * y[8]) - P(1.0 / 6.0) * trx * tr_y;
}

// o = c * M x
template <typename P>
DEV void applyM(P c, const P* x, P* o) {
  const P tr = (x[3] + x[4] + x[5]) * P(1.0 / 6.0);
  const P h = P(0.5) * c;
  o[0] = h * x[0]; o[1] = h * x[1]; o[2] = h * x[2];
  o[3] = c * (P(0.5) * x[3] - tr); o[4] = c * (P(0.5) * x[4] - tr); o[5] = c * (P(0.5) * x[5] - tr);
  o[6] = c * x[6]; o[7] = c * x[7]; o[8] = c * x[8];
}

template <typename P>
DEV P dot9(const P* x, const P* y) {
  P a = x[0] * y[0];
#pragma unroll
  for (int k = 1; k < 9; k++) a += x[k] * y[k];
  return a;
}

// Gradient in p of L . psi(p) given u, u', w, w' at |p| (ir = 1/|p|).
template <typename P>
DEV void grad_psi(const P* L, const P* p, P ir, P u, P du, P w, P dw, P* g) {
  const P la = L[0] * p[0] + L[1] * p[1] + L[2] * p[2];
  const P g0 = P(2) * L[3] * p[0] + L[6] * p[1] + L[7] * p[2];
  const P g1 = L[6] * p[0] + P(2) * L[4] * p[1] + L[8] * p[2];
  const P g2 = L[7] * p[0] + L[8] * p[1] + P(2) * L[5] * p[2];
  const P Q = P(0.5) * (p[0] * g0 + p[1] * g1 + p[2] * g2);
  const P c = (du * la + dw * Q) * ir;
  g[0] = c * p[0] + u * L[0] + w * g0;
  g[1] = c * p[1] + u * L[1] + w * g1;
  g[2] = c * p[2] + u * L[2] + w * g2;
}

// A point's product with itself is radial: with r = |p|,
//   <psi(p), psi(p)> = 1/2 u^2 r^2 + 1/3 w^2 r^4
// (the trace term removes a third of B:B = w^2 r^4), and the gradient of
// L . psi(p) at the fixed adjoint L = M psi(p) is half the gradient of that
// product: J(p)^T M psi(p) = h p with
//   h = 1/2 u (u' r + u) + w r^2 (1/3 w' r + 2/3 w).
// The coincident (pair-rule) angular terms and the j1 = j2 self terms of the
// triplet sums take these forms instead of the 9-component algebra.
// Written with explicit FMAs and no implicit contraction, so they compile to
// the same operations wherever they are inlined.
template <typename P>
DEV P self_psi(P u, P w, P r2) {
#pragma clang fp contract(off)
  return r2 * fma(P(0.5) * u, u, (P(1.0 / 3.0) * w) * (w * r2));
}
template <typename P>
DEV P self_hgrad(P u, P du, P w, P dw, P r, P r2) {
#pragma clang fp contract(off)
  const P a = fma(du, r, u);
  const P b = fma(P(1.0 / 3.0) * dw, r, P(2.0 / 3.0) * w);
  return fma(P(0.5) * u, a, (w * r2) * b);
}

DEV void min_image(const Box& b, const double4& qi, const double4& qj, double* r) {
  double v0 = qi.x - qj.x, v1 = qi.y - qj.y, v2 = qi.z - qj.z;
  if (b.per[0]) v0 -= b.L[0] * rint(v0 * b.invL[0]);
  if (b.per[1]) v1 -= b.L[1] * rint(v1 * b.invL[1]);
  if (b.per[2]) v2 -= b.L[2] * rint(v2 * b.invL[2]);
  r[0] = v0; r[1] = v1; r[2] = v2;
}

DEV int pair_slot(int si, int sj) { return si == sj ? si : 2; }

// Owner of the unordered pair {i, j}: the lower index when i + j is odd, the
// higher when even, so each site owns about half of its pairs.
DEV bool pair_owner(int i, int j) { return ((i ^ j) & 1) ? i < j : i > j; }

template <int K>
DEV double group_sum(double v) {
#pragma unroll
  for (int o = K / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, K);
  return v;
}

// Neighbour walks.  The K >= 4 lanes of a site walk its row together for
// ceil(nn / K) steps (the cooperative gathers below need every lane of a
// group of four): a lane past the row's end takes the site itself as the
// neighbour, with zero occupancy, and this fixed displacement (|r| = 2.4 A,
// inside every table), so all it forms is finite and enters the sums with
// weight zero.
DEV void idle_disp(bool inb, double* r) {
  if (!inb) { r[0] = 1.1; r[1] = 1.3; r[2] = 1.7; }
}

// ---------------------------------------------------------------------------
// Cooperative row gathers.  Each lane needs whole table rows (8 or 16
// coefficients) at lane-dependent indices.  A per-lane load touches one cache
// line per lane and per 16-byte load, and the vector memory pipeline's time
// grows with the number of distinct lines an instruction touches.  Here the
// NQ = row bytes / 16 lanes of an aligned group load the NQ chunks of one
// member's row with one instruction, member by member (NQ rounds: 64/NQ lines
// per instruction instead of 64), and the chunks pass through a per-wavefront
// LDS area to the lane that owns the row.  FP32 rows (32 and 16 B) gain at
// most a factor two in lines; the kernels choose per phase and precision
// (COOP) what measured faster on MI355X.
//
// Every lane of a group must execute the gather (the phase kernels walk the
// neighbour rows in wavefront-uniform loops, and whole groups leave early).
// Wavefront-scope fences keep the LDS writes before the reads and the reads
// before the next gather's writes; they emit no instructions.  All global
// loads of a gather are issued before its first fence.
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kStageSlots = 65;  // 16-byte slots per round: 64 lanes + 1 (LDS bank skew)
// 16-byte chunks per row of W values; stage rounds a gather of N 8-value rows
// uses.
template <typename P, int W> struct RowQ { static constexpr int v = (int)(W * sizeof(P) / 16); };
template <typename P, int N, bool COOP> struct StageRounds {
  static constexpr int v = COOP ? N * RowQ<P, 8>::v : 1;
};

// Value of v at lane R of this lane's group of NQ: a quad permutation (DPP)
// for NQ <= 4, a bit-mask swizzle (lane = (lane & 0x18) | R within each half
// wavefront) for NQ = 8.
template <int NQ, int R>
DEV int group_bcast(int v) {
  if constexpr (NQ == 8) {
    return __builtin_amdgcn_ds_swizzle(v, 0x18 | (R << 5));
  } else {
    constexpr int ctrl = NQ == 4 ? R * 0x55 : (R | (R << 2) | ((2 + R) << 4) | ((2 + R) << 6));
    return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xf, 0xf, false);
  }
}

// Row loads address base + a 32-bit byte offset: with a wavefront-uniform
// base the load takes the scalar-base + 32-bit vector-offset form, which saves
// the 64-bit address arithmetic (sign extension, shift, add) of every gathered
// row.  Callers keep offsets below 4 GiB: the tables are a few MB, and phase 2
// addresses adjoint rows from the first record of their instance
// (kMaxInstanceSites).
template <typename T>
DEV const T* at_off(const T* base, unsigned off) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
DEV u32x4 ld16(const u32x4* base, unsigned idx) { return *at_off(base, idx * 16u); }

template <int NQ>
DEV void coop_rounds(const u32x4* base, int row, int a, u32x4* v) {
  if constexpr (NQ == 1) {
    v[0] = ld16(base, (unsigned)row);
  } else if constexpr (NQ == 8) {
    v[0] = ld16(base, (unsigned)group_bcast<8, 0>(row) * 8u + (unsigned)a);
    v[1] = ld16(base, (unsigned)group_bcast<8, 1>(row) * 8u + (unsigned)a);
    v[2] = ld16(base, (unsigned)group_bcast<8, 2>(row) * 8u + (unsigned)a);
    v[3] = ld16(base, (unsigned)group_bcast<8, 3>(row) * 8u + (unsigned)a);
    v[4] = ld16(base, (unsigned)group_bcast<8, 4>(row) * 8u + (unsigned)a);
    v[5] = ld16(base, (unsigned)group_bcast<8, 5>(row) * 8u + (unsigned)a);
    v[6] = ld16(base, (unsigned)group_bcast<8, 6>(row) * 8u + (unsigned)a);
    v[7] = ld16(base, (unsigned)group_bcast<8, 7>(row) * 8u + (unsigned)a);
  } else if constexpr (NQ == 4) {
    v[0] = ld16(base, (unsigned)group_bcast<4, 0>(row) * 4u + (unsigned)a);
    v[1] = ld16(base, (unsigned)group_bcast<4, 1>(row) * 4u + (unsigned)a);
    v[2] = ld16(base, (unsigned)group_bcast<4, 2>(row) * 4u + (unsigned)a);
    v[3] = ld16(base, (unsigned)group_bcast<4, 3>(row) * 4u + (unsigned)a);
  } else {
    v[0] = ld16(base, (unsigned)group_bcast<2, 0>(row) * 2u + (unsigned)a);
    v[1] = ld16(base, (unsigned)group_bcast<2, 1>(row) * 2u + (unsigned)a);
  }
}

// o8[q] = row r8[q] of t8[q] (rows of 8 values), q < N, with one LDS round
// trip.  stage: this wavefront's LDS area, StageRounds<P, N, COOP> *
// kStageSlots slots.
template <typename P, int N, bool COOP>
DEV void gather8(const P* const* t8, const int* r8, u32x4* stage, P (*o8)[8]) {
  if constexpr (!COOP) {
#pragma unroll
    for (int q = 0; q < N; q++) {
      const P* c = at_off(t8[q], (unsigned)r8[q] * (unsigned)(8 * sizeof(P)));
#pragma unroll
      for (int e = 0; e < 8; e++) o8[q][e] = c[e];
    }
  } else {
    constexpr int Q8 = RowQ<P, 8>::v;
    const int lane = (int)__lane_id();
    const int a8 = lane & (Q8 - 1);
    u32x4 v8[N][Q8];
#pragma unroll
    for (int q = 0; q < N; q++)
      coop_rounds<Q8>(reinterpret_cast<const u32x4*>(t8[q]), r8[q], a8, v8[q]);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
    for (int q = 0; q < N; q++)
#pragma unroll
      for (int r = 0; r < Q8; r++) stage[(q * Q8 + r) * kStageSlots + lane] = v8[q][r];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
    for (int q = 0; q < N; q++) {
      u32x4 w[Q8];
#pragma unroll
      for (int c = 0; c < Q8; c++) w[c] = stage[(q * Q8 + a8) * kStageSlots + (lane - a8) + c];
      __builtin_memcpy(o8[q], w, sizeof(w));
    }
  }
}

// A pair-rule point's rows: o8 = u/w row r8 of t8 (8 values), o16 = fused
// row r16 of t16 (16 values: phi, rho_Mg, rho_H, 0), with one LDS round trip.
// C8 / C16: cooperative or direct for each; L8 = false loads the fused row
// only (o8 untouched).  stage: CptRounds<P, C8, C16> * kStageSlots slots.
template <typename P, bool C8, bool C16> struct CptRounds {
  static constexpr int v = (C8 ? RowQ<P, 8>::v : 0) + (C16 ? RowQ<P, 16>::v : 0) > 0
                               ? (C8 ? RowQ<P, 8>::v : 0) + (C16 ? RowQ<P, 16>::v : 0) : 1;
};
template <typename P, bool C8_, bool C16, bool L8 = true>
DEV void gather_cpt(const P* t8, int r8, const P* t16, int r16, u32x4* stage, P* o8, P* o16) {
  constexpr bool C8 = C8_ && L8;
  constexpr int Q8 = RowQ<P, 8>::v, Q16 = RowQ<P, 16>::v;
  const int lane = (int)__lane_id();
  const int a8 = lane & (Q8 - 1), a16 = lane & (Q16 - 1);
  const u32x4* b8 = reinterpret_cast<const u32x4*>(t8);
  const u32x4* b16 = reinterpret_cast<const u32x4*>(t16);
  u32x4 v8[Q8], v16[Q16];
  if constexpr (C8) coop_rounds<Q8>(b8, r8, a8, v8);
  else if constexpr (L8) {
#pragma unroll
    for (int c = 0; c < Q8; c++) v8[c] = ld16(b8, (unsigned)r8 * Q8 + c);
  }
  if constexpr (C16) coop_rounds<Q16>(b16, r16, a16, v16);
  else {
#pragma unroll
    for (int c = 0; c < Q16 - 1; c++) v16[c] = ld16(b16, (unsigned)r16 * Q16 + c);  // last chunk: padding
    v16[Q16 - 1] = u32x4{0, 0, 0, 0};
  }
  if constexpr (C8 || C16) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if constexpr (C8) {
#pragma unroll
      for (int r = 0; r < Q8; r++) stage[r * kStageSlots + lane] = v8[r];
    }
    if constexpr (C16) {
      constexpr int o = C8 ? Q8 : 0;
#pragma unroll
      for (int r = 0; r < Q16; r++) stage[(o + r) * kStageSlots + lane] = v16[r];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if constexpr (C8) {
#pragma unroll
      for (int c = 0; c < Q8; c++) v8[c] = stage[a8 * kStageSlots + (lane - a8) + c];
    }
    if constexpr (C16) {
      constexpr int o = C8 ? Q8 : 0;
#pragma unroll
      for (int c = 0; c < Q16; c++) v16[c] = stage[(o + a16) * kStageSlots + (lane - a16) + c];
    }
  }
  if constexpr (L8) __builtin_memcpy(o8, v8, sizeof(v8));
  __builtin_memcpy(o16, v16, sizeof(v16));
}

// rho coefficients of species slot sp (0 Mg, 1 H) in a fused row.
template <typename P>
DEV void fused_rho(const P* f, int sp, P* o) {
#pragma unroll
  for (int e = 0; e < 4; e++) o[e] = sp ? f[8 + e] : f[4 + e];
}

template <typename P, int NR, bool COOP>
DEV void gather_rows8(const P* tab, const int* rows, u32x4* stage, P (*out)[8]) {
  const P* t8[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) t8[q] = tab;
  gather8<P, NR, COOP>(t8, rows, stage, out);
}

template <typename P, bool COOP>
DEV void gather_row8(const P* tab, int row, u32x4* stage, P* out) {
  gather_rows8<P, 1, COOP>(tab, &row, stage, reinterpret_cast<P(*)[8]>(out));
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

}  // namespace dmd
