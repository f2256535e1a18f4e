// dmd_diffusion.hip -- implicit occupancy (diffusion) step of H sites.
//
// The reference builds the diffusive topology (src/Atoms/Neighbors.cpp:139-277:
// H-H pairs within r_cutoff_Eb = 2.6 A, at most max_chemical_neighs = 50,
// src/Macros.hpp:860,863) and declares the diffusion parameters
// (src/Variables.hpp:66-90) but has no rate law.  The rate law here is the
// lattice-gas gradient flow in Onsager form with the ideal (Fermi-Dirac) site
// entropy l(x) = ln(x / (1 - x)) and a logarithmic-mean mobility:
//
//   dx_i/dt = sum_j G_ij [ (x_j - x_i) + L(x_i, x_j) (mu_j - mu_i) ],
//   L(a, b) = (b - a) / (l(b) - l(a))      (a (1 - a) as b -> a),
//   G_ij = D0 / r_ij^2 * exp(-beta_ij Eb),  beta_ij = (beta_i + beta_j) / 2,
//
// i.e. G L (mu_tot_j - mu_tot_i) with mu_tot = mu + l(x).  mu_i is the
// interaction chemical potential in units of k_B T: the occupancy residual
// R_x(i) = d(sum_c S0_c)/dx_i without the site's own entropy terms
// (-3/(2x) - [x<1]/(1-x) - [x<0.9999] gamma).  Masked sites (x < 5e-4) do
// not interact, so L = 0 on any pair with a masked site (Fick exchange only);
// which sites are active is frozen at the start of a step.  Without
// interactions the law is exactly Fick's with D = D0 exp(-Eb / k_B T) (the
// Arrhenius rate); it conserves sum_i x_i (the pair flux is antisymmetric
// bit for bit) and its stationary states are mu_tot = const.
//
// The implicit (backward Euler) step x = x_n + dt xdot(x) is solved by Picard
// sweeps; each sweep needs one residual evaluation at the current iterate.
#include "dmd_device.h"
#include "dmd_solver.h"

namespace dmd {

constexpr double kMinOccD = 5e-4;  // src/Macros.hpp:871

// Periodic minimum-image distance without contraction (the membership test
// r <= 2.6 rounds as the CPU restatement does).
DEV double dist_nc(const Box& box, const double4& a, const double4& b) {
  double v[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
#pragma unroll
  for (int e = 0; e < 3; e++)
    if (box.per[e]) v[e] = __dsub_rn(v[e], __dmul_rn(box.L[e], rint(v[e] / box.L[e])));
  const double s = __dadd_rn(__dadd_rn(__dmul_rn(v[0], v[0]), __dmul_rn(v[1], v[1])),
                             __dmul_rn(v[2], v[2]));
  return sqrt(s);
}

// One thread per site of the instance [base, base + n): keeps the H neighbours
// of an H site within kRcEb, in neighbour-row order.
__global__ __launch_bounds__(256) void k_diff_neigh(int base, int n, Box box, const int* spc,
                                                   const double4* qs, const int* nnb,
                                                   const int* nbr, int* dnn, int* dnbr,
                                                   int* flags) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int i = base + t;
  if (spc[i] != 1) {
    dnn[i] = 0;
    return;
  }
  const double4 qi = qs[i];
  const int* row = nbr + (size_t)i * kCap;
  int* out = dnbr + (size_t)i * kDCap;
  int cnt = 0;
  bool over = false;
  for (int kk = 0; kk < nnb[i]; kk++) {
    const int j = row[kk];
    if (spc[j] != 1) continue;
    if (dist_nc(box, qi, qs[j]) <= kRcEb) {
      if (cnt < kDCap) out[cnt] = j;
      else over = true;
      cnt++;
    }
  }
  dnn[i] = cnt < kDCap ? cnt : kDCap;
  if (over) atomicOr(&flags[0], 1);
  if (cnt == 0) atomicOr(&flags[1], 1);
}

// Interaction chemical potential of site i (units of k_B T).
DEV double mu_int(const State& s, const Outputs& o, int i) {
  const double x = s.ne[i];
  if (x == 0.0) return 0.0;
  double dx = -1.5 / x;
  if (x < 1.0) dx -= 1.0 / (1.0 - x);
  if (x < 0.9999) dx -= s.gamma[i];
  return o.Rx[i] - dx;
}

// Logarithmic-mean mobility; symmetric in (a, b) bit for bit.
DEV double log_mean(double a, double b, bool act) {
  if (!act || !(a > 0.0 && a < 1.0 && b > 0.0 && b < 1.0)) return 0.0;
  const double d = log(b / (1.0 - b)) - log(a / (1.0 - a));
  if (fabs(d) < 1e-8) {
    const double m = 0.5 * (a + b);
    return m * (1.0 - m);
  }
  return (b - a) / d;
}

// Xn holds the occupancies at the start of the step (activity and, for
// kDiffStep, the backward-Euler base).
__global__ __launch_bounds__(256) void k_diff_flux(State s, const SolveOp* ops, SolveVecs v,
                                                  Outputs o, const int* dnn, const int* dnbr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= s.n) return;
  const int k = s.inst[i];
  if (k < 0) return;
  const SolveOp op = ops[k];
  if (op.diff == kDiffNone) return;
  if (s.spc[i] != 1) {
    v.Xt[i] = op.diff == kDiffStep ? s.xr[i] : 0.0;
    return;
  }
  const Box box = s.boxes[k];
  const double4 qi = s.qs[i];
  const double xi = s.xr[i], mi = mu_int(s, o, i), bi = s.beta[i];
  const bool ai = v.Xn[i] >= kMinOccD;
  const int* row = dnbr + (size_t)i * kDCap;
  double f = 0.0;
  for (int jj = 0; jj < dnn[i]; jj++) {
    const int j = row[jj];
    const double r = dist_nc(box, qi, s.qs[j]);
    const double xj = s.xr[j], mj = mu_int(s, o, j);
    const double g = op.d0 / (r * r) * exp(-(0.5 * (bi + s.beta[j])) * op.eb);
    const bool act = ai && v.Xn[j] >= kMinOccD;
    f += g * ((xj - xi) + log_mean(xi, xj, act) * (mj - mi));
  }
  v.Xt[i] = op.diff == kDiffStep ? v.Xn[i] + op.dt * f : f;
}

// Accepts the sweep: x = Xt; the effective occupancy follows x on the sites
// active at the start of the step.
__global__ __launch_bounds__(256) void k_diff_apply(int n, const int* inst, const int* spc,
                                                   const SolveOp* ops, SolveVecs v, double* xr,
                                                   double* ne) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int k = inst[i];
  if (k < 0 || ops[k].diff != kDiffStep || spc[i] != 1) return;
  const double x = v.Xt[i];
  xr[i] = x;
  ne[i] = v.Xn[i] >= kMinOccD ? x : 0.0;
}

hipError_t launch_diff_neighbors(hipStream_t st, int base, int n, const Box& box, const int* spc,
                                 const double4* qs, const int* nnb, const int* nbr, int* dnn,
                                 int* dnbr, int* flags) {
  if (n <= 0) return hipSuccess;
  k_diff_neigh<<<(n + 255) / 256, 256, 0, st>>>(base, n, box, spc, qs, nnb, nbr, dnn, dnbr, flags);
  return hipGetLastError();
}

hipError_t launch_diff_flux(hipStream_t st, const State& s, const SolveOp* ops, const SolveVecs& v,
                            const Outputs& o, const int* dnn, const int* dnbr) {
  if (s.n <= 0) return hipSuccess;
  k_diff_flux<<<(s.n + 255) / 256, 256, 0, st>>>(s, ops, v, o, dnn, dnbr);
  return hipGetLastError();
}

hipError_t launch_diff_apply(hipStream_t st, int n, const int* inst, const int* spc,
                             const SolveOp* ops, const SolveVecs& v, double* xr, double* ne) {
  if (n <= 0) return hipSuccess;
  k_diff_apply<<<(n + 255) / 256, 256, 0, st>>>(n, inst, spc, ops, v, xr, ne);
  return hipGetLastError();
}

}  // namespace dmd
