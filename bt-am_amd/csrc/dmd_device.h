// dmd_device.h -- device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "dmd_kernels.h"

namespace dmd {

#define DEV __device__ __forceinline__

constexpr double kSqrt6 = 2.449489742783178;  // src/Macros.hpp:42
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

// ---------------------------------------------------------------------------
// Quadrature-point geometry.  A point is p = r + d e_l (one axis shifted).
// PairGeo<P>::point returns p, the spline segment m and offset of |p|, and
// 1/|p|.  FP64: |p| directly.  FP32: the pair distance |r| is kept in FP64
// and |p| = |r| + D with the small increment D = (2 r_l d + d^2)/(|p| + |r|)
// in FP32, so the segment offset (|r| - m dx) + D carries FP64-quality
// position information while every other operation is FP32.
// ---------------------------------------------------------------------------
template <typename P> struct PairGeo;

template <> struct PairGeo<double> {
  double r[3];
  DEV void init(const double* rd) { r[0] = rd[0]; r[1] = rd[1]; r[2] = rd[2]; }
  DEV void point(const Tables& T, int l, double d, double* p, int& m, double& off,
                 double& ir) const {
    p[0] = r[0] + (l == 0 ? d : 0.0);
    p[1] = r[1] + (l == 1 ? d : 0.0);
    p[2] = r[2] + (l == 2 ? d : 0.0);
    const double x = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    ir = 1.0 / x;
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
    const float D = __fdividef(2.0f * rl * d + d * d, pn + R32);
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

}  // namespace dmd
