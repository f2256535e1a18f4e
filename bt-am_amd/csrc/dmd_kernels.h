// dmd_kernels.h -- device-side data layout and kernel launchers (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace dmd {

// Species slots: 0 = Mg, 1 = H.  Pair-table slots: 0 = MgMg, 1 = HH, 2 = MgH.
constexpr int kCap = 256;          // neighbour row stride (reference maxneigh = 250)
constexpr int kMaxNeigh = 250;     // src/Macros.hpp:862
// Phase-1 record of a site: eight 9-component adjoint blocks
//   blocks 0..5 : (1/9) M G_ls   (i-shift sums, block = 2 l + s)
//   block 6     : (1/9) M U      (j-shift sums)
//   block 7     : (1/9) M G0     (unshifted sums)
// stored as aligned 8-value rows (components 0..7 of block b at 8 b) and one
// tail row (component 8 of block b at kRecTail + b), then three scalars.  A
// block is one aligned row plus one value, so gathers of it touch few lines.
constexpr int kRec = 80;           // doubles per phase-1 record (10 rows of 8)
constexpr int kRecRows = kRec / 8;
constexpr int kBlkLU = 6;
constexpr int kBlkLG0 = 7;
constexpr int kRecTail = 64;
constexpr int kRecFP = 72;         // F'(rho_bar)
constexpr int kRecDVDN = 73;       // dV0_i / dx_i (own occupancy)
constexpr int kRecRHO = 74;        // rho_bar
constexpr int kRec32 = 72;         // floats per FP32 copy of the eight blocks (split path)
// Slot of component c of block b in a record (FP64 or FP32 copy).
__host__ __device__ constexpr int rec_slot(int b, int c) { return c < 8 ? 8 * b + c : kRecTail + b; }

// Spline tables resident in HBM (L2-resident in practice: ~0.8 MB for the
// reference tables).  All r-tables share one (nr, dx) grid.  Coefficient rows
// hold the value coefficients only (derivatives are b + (2c + 3d p) p):
//   uw[t][m] : u: a b c d | w: a b c d   (8 doubles, one 64-B line segment)
//   ph[t][m] : phi: a b c d               (4 doubles)
//   rh[s][m] : rho: a b c d
//   em[s][m] : embedding F: a b c d (own grid per species)
struct Tables {
  const double* uw;
  const double* ph;
  const double* rh;
  const double* em;
  const float* uw32;   // FP32 copies of uw, ph, rh (split-precision path)
  const float* ph32;
  const float* rh32;
  const double* pr;    // fused pair-rule rows [t][m]: phi_t | rho_Mg | rho_H | 0 (16 values)
  const float* pr32;
  const double* rr;    // density rows [m]: rho_Mg | rho_H (8 values; relaxation ticks)
  const float* rr32;
  int nr;
  double dxr, inv_dxr;
  float dxr32, inv_dxr32;
  int ne[2];
  int ne_stride;
  double dxe[2], inv_dxe[2];
  double mass[2];
};

struct Box {
  double lo[3], L[3], invL[3];
  int per[3];
  int base;  // first arena site of the instance
};

// Sites of all resident instances live in one arena.  Every instance starts
// at a multiple of kChunk sites, so a kChunk-aligned block of sites belongs to
// a single instance.
constexpr int kChunk = 256;
// Largest instance (sites, a multiple of kChunk): phase 2 addresses a
// neighbour's adjoint rows by a 32-bit byte offset from its instance's first
// record (kRec doubles per site), so an instance's records stay below 4 GiB.
constexpr int kMaxInstanceSites = (int)((4294967296ull / (kRec * sizeof(double))) / kChunk * kChunk);
constexpr int kSlotW = 5;

// What a solver tick's evaluation may leave out, per instance (State::noxr):
// the occupancy residual R_x (relaxations keep occupancies fixed); the pairs
// with two H ends and the H sites' sums (only Mg positions and widths are
// relaxed: an H-H pair adds to nothing else); the pair-rule values C6 and Phi
// (the residual-norm line search needs no energies: V0, S0 and the energy
// totals are then not formed, every gradient is).
constexpr int kSkipRx = 1;
constexpr int kSkipHH = 2;
constexpr int kSkipValues = 4;

struct State {
  int n;               // arena extent (sites)
  const int* inst;     // instance of each site, -1 for a free slot
  const Box* boxes;    // per instance
  const int* prec;     // per instance: Precision to evaluate with, -1 = skip
  // per instance (nullptr: none): kSkip* bits, what a solver tick's
  // evaluation leaves out
  const int* noxr;
  const double4* qs;   // (x, y, z, sigma)
  const double* ne;    // effective occupancy: x if active (x >= 5e-4) else 0
  const double* xr;    // raw occupancy
  const int* spc;      // species slot
  const double* beta;
  const double* gamma;
  const int* nnb;
  const int* nbr;      // n x kCap
  // Pairs owned by the sites of each chunk (phase 2 evaluates each unordered
  // pair once, at its owner): chunk c holds chunk_np[c] entries, first the
  // chunk_npx[c] pairs with an Mg end, then those with two H ends, each part in
  // site order:
  // own[c * kChunk * kCap + j] = li | kk << 8 | pos << 16 -- owner site
  // c * kChunk + li, kk the slot of the neighbour k in its row, pos the slot of
  // the owner in k's row.
  const int* chunk_np;
  const int* chunk_npx;
  const int* own;
  // Chunks to evaluate, per precision: act[p * act_stride] = count, then the
  // chunk indices in arena order.  The phase kernels map their blocks onto
  // these chunks only, so free slots and instances that skip this evaluation
  // cost no work and the XCD-contiguous block order spreads the real work
  // over all 8 XCDs.
  const int* act;
  int act_stride;
};

struct Outputs {
  double* rec;         // n x kRec
  float* rec32;        // n x kRec32: LA, LU, LG0 in FP32 for split-precision instances
                       // (phase 2 of the split path reads them as FP32 anyway)
  double* V0;
  double* rho;
  double* Rq;          // n x 3
  double* Rs;
  double* Rx;
  double* S0;
  double* slot;        // n x kCap x kSlotW: a pair's phase-2 contribution to the site
                       // of that row (dR_q x3, dR_sigma, dR_x), written by the owner
  double* partials;    // per-chunk [S0, V0, domain] partial sums
  double* totals;      // per instance [sum S0, sum V0, sites whose density left the
                       // embedding table's domain]
};

// Neighbour build (cell list).  Returns via flags[0] = overflow, flags[1] = empty row.
struct CellGrid {
  int nc[3];
  double lo[3], cs[3], inv_cs[3];
};

// Builds the lists of one instance whose sites are qs[base, base + n); the
// lists hold arena indices.
hipError_t launch_neighbors(hipStream_t st, const Box& box, const CellGrid& g, double rc, int base,
                            int n, const double4* qs, int* cell_of, int* cell_count,
                            int* cell_start, int* cell_cursor, int* cell_atoms, int* nnb, int* nbr,
                            int* flags);

// Owned-pair lists of one instance (after launch_neighbors on the same
// stream): flags[2] is set if a neighbour relation is not symmetric.
hipError_t launch_own_lists(hipStream_t st, int base, int n, const int* nnb, const int* nbr,
                            const int* spc, int* chunk_np, int* chunk_npx, int* own, int* flags);

// Precision of the quadrature-point evaluation: FP64 throughout, or the split
// form (FP32 points and splines, FP64 distances, sums, records and totals).
enum Precision { kFP64 = 0, kSplit32 = 1 };

struct PhaseConfig {
  int precision;     // Precision of one-instance evaluations
  // per Precision (index kFP64 / kSplit32):
  int k1[2], k2[2];  // lanes per site in phase 1 / phase 2 (1, 2, 4, 8, 16)
  int unroll2[2];    // 1: fully unrolled point loops in phase 2, 0: rolled
  int rule = 3;      // pair / coincident rule: 3 (reference, 12 points), 5 or 44
                     // (27-point tensor Gauss-Hermite on the relative coordinate)
  bool operator==(const PhaseConfig& o) const {
    return k1[0] == o.k1[0] && k1[1] == o.k1[1] && k2[0] == o.k2[0] && k2[1] == o.k2[1] &&
           unroll2[0] == o.unroll2[0] && unroll2[1] == o.unroll2[1] && rule == o.rule;
  }
};

// Evaluates every instance whose prec[] equals the given precision (both
// precisions are launched when instances of both kinds are present).
hipError_t launch_phase1(hipStream_t st, const PhaseConfig& cfg, int precision, const State& s,
                         const Tables& T, const Outputs& o);
hipError_t launch_phase2(hipStream_t st, const PhaseConfig& cfg, int precision, const State& s,
                         const Tables& T, const Outputs& o);
// Per-instance totals: chunk sums, then per instance over its chunk range
// [chunk_lo[k], chunk_hi[k]) in a fixed order.
// Diagonal 3x3 blocks D[(u - base) * 9] of dR_q/dq for sites [base, base + n)
// from the FP64 phase-1 records (rules 3, 5 and 44; exact: see dmd_hess.hip);
// sites of a species s with bit s of spc_mask clear get zero blocks.
hipError_t launch_diag(hipStream_t st, int rule, int exact, int spc_mask, const State& s,
                       const Tables& T, const double* rec, int base, int n, double* D);
hipError_t launch_totals(hipStream_t st, const State& s, const Tables& T, const Outputs& o,
                         int ninst, const int* chunk_lo, const int* chunk_hi);
// out_per_chunk[2 c] = active ordered pairs of chunk c, [2 c + 1] = those with an Mg end.
hipError_t launch_active_pairs(hipStream_t st, const State& s, long long* out_per_chunk);
// Packed vectors of one instance (sites [base, base + n) of the arena; perm =
// the caller's index of each arena site): X = [q (3n) | sigma (n)] into qs and
// the solver iterate Xs (bad[0] = 1 on a non-finite position or a width that
// is not positive); R = [R_q (3n) | R_sigma (n) | R_x (n)] out.
hipError_t launch_gather_unknowns(hipStream_t st, int base, int n, const int* perm, const double* X,
                                  double4* qs, double4* Xs, int* bad);
hipError_t launch_scatter_residual(hipStream_t st, int base, int n, const int* perm,
                                   const double* Rq, const double* Rs, const double* Rx,
                                   double* R);
// Packed-vector helpers (caller-order vectors of length m): out[0] = |v|^2
// (fixed-order sum), y = x + s v, d = (a - b) c.
hipError_t launch_sumsq(hipStream_t st, const double* v, size_t m, double* out);
hipError_t launch_axpy(hipStream_t st, const double* x, const double* v, double s, size_t m,
                       double* y);
hipError_t launch_diff_scale(hipStream_t st, const double* a, const double* b, double c, size_t m,
                             double* d);

}  // namespace dmd
