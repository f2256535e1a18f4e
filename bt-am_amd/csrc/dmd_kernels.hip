// dmd_kernels.hip -- neighbour build, totals and work counts (gfx950).
#include "dmd_device.h"

namespace dmd {

// ---------------------------------------------------------------------------
// Deterministic totals: per-block sums in a fixed tree, then one block.  The
// third sum counts the sites whose density left the embedding table's domain:
// the reference's spline takes segment trunc(x/dx), a negative index for
// x <= -dx (src/Numerical/cubic-spline.cpp:84-93), so such a density has no
// reference value; a non-finite density counts too.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_block_sums(int n, const int* inst, const int* prec,
                                                   const int* spc, const double* rho,
                                                   double dxe0, double dxe1, const double* S0,
                                                   const double* V0, double* partials) {
  __shared__ double sa[256], sb[256], sc[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int k = inst[blockIdx.x * 256];  // one instance per 256-site chunk
  const bool on = i < n && k >= 0 && prec[k] >= 0;
  sa[threadIdx.x] = on ? S0[i] : 0.0;
  sb[threadIdx.x] = on ? V0[i] : 0.0;
  sc[threadIdx.x] = on && !(rho[i] > -(spc[i] ? dxe1 : dxe0) && rho[i] < 1e300) ? 1.0 : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = sa[0];
    partials[3 * blockIdx.x + 1] = sb[0];
    partials[3 * blockIdx.x + 2] = sc[0];
  }
}

// One block per instance: sums its chunk partials in a fixed order.
__global__ __launch_bounds__(256) void k_inst_sums(const int* prec, const int* chunk_lo,
                                                  const int* chunk_hi, const double* partials,
                                                  double* totals) {
  __shared__ double sa[256], sb[256], sc[256];
  const int k = blockIdx.x;
  if (prec[k] < 0) return;
  double a = 0.0, b = 0.0, d = 0.0;
  for (int c = chunk_lo[k] + threadIdx.x; c < chunk_hi[k]; c += 256) {
    a += partials[3 * c];
    b += partials[3 * c + 1];
    d += partials[3 * c + 2];
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  sc[threadIdx.x] = d;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    totals[3 * k] = sa[0];
    totals[3 * k + 1] = sb[0];
    totals[3 * k + 2] = sc[0];
  }
}

// Active ordered pairs per 256-site chunk (for the work count): out[2 c] all
// of them, out[2 c + 1] those with an Mg end (what a tick that relaxes the Mg
// lattice only evaluates in phase 2).
__global__ __launch_bounds__(256) void k_active_pairs(State s, long long* out) {
  __shared__ long long sh[2][256];
  int i = blockIdx.x * 256 + threadIdx.x;
  long long c = 0, cm = 0;
  if (i < s.n && s.inst[i] >= 0 && s.ne[i] != 0.0) {
    const int* row = s.nbr + (size_t)i * kCap;
    const bool mi = s.spc[i] == 0;
    for (int k = 0; k < s.nnb[i]; k++) {
      const bool on = s.ne[row[k]] != 0.0;
      c += on;
      cm += on && (mi || s.spc[row[k]] == 0);
    }
  }
  sh[0][threadIdx.x] = c;
  sh[1][threadIdx.x] = cm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sh[0][0];
    out[2 * blockIdx.x + 1] = sh[1][0];
  }
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

__global__ __launch_bounds__(256) void k_build(int base, int n, Box box, CellGrid g, double rc,
                                              const double4* qs, const int* cell_of,
                                              const int* start, const int* atoms, int* nnb,
                                              int* nbr, int* flags) {
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
            if (cnt < kMaxNeigh) row[cnt] = base + j;
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

// Owned pairs of one chunk (block = chunk, thread = site): first the pairs with
// an Mg end, then those with two H ends (species 1), each part in site order
// and row order within a site: li | kk << 8 | pos << 16 (owner li of the chunk, kk
// the slot of the neighbour k in its row, pos the slot of the owner in k's
// row).  Rows are sorted, so the reverse slot is a binary search.
DEV int rev_slot(const int* nnb, const int* nbr, int k, int i) {
  const int* rk = nbr + (size_t)k * kCap;
  const int nk = nnb[k];
  int lo = 0, hi = nk;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rk[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  return lo < nk && rk[lo] == i ? lo : -1;
}

__global__ __launch_bounds__(256) void k_own_lists(int base, int n, const int* nnb, const int* nbr,
                                                   const int* spc, int* chunk_np, int* chunk_npx,
                                                   int* own, int* flags) {
  __shared__ int scan[2][256];
  const int li = blockIdx.x * 256 + threadIdx.x;
  const int i = base + li;
  const int cnt = li < n ? nnb[i] : 0;
  const bool hi = li < n && spc[i] == 1;
  const int* row = nbr + (size_t)i * kCap;
  int nx = 0, nh = 0;  // owned pairs with an Mg end, with two H ends
  for (int kk = 0; kk < cnt; kk++) {
    const int k = row[kk];
    if (!pair_owner(i, k)) continue;
    if (hi && spc[k] == 1) nh++;
    else nx++;
  }
  // exclusive scans over the chunk's sites (Hillis-Steele in LDS)
  scan[0][threadIdx.x] = nx;
  scan[1][threadIdx.x] = nh;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const int vx = threadIdx.x >= d ? scan[0][threadIdx.x - d] : 0;
    const int vh = threadIdx.x >= d ? scan[1][threadIdx.x - d] : 0;
    __syncthreads();
    scan[0][threadIdx.x] += vx;
    scan[1][threadIdx.x] += vh;
    __syncthreads();
  }
  const int ch = (base >> 8) + (int)blockIdx.x;
  const int totx = scan[0][255];
  int* out = own + (size_t)ch * kChunk * kCap;
  int jx = scan[0][threadIdx.x] - nx;         // pairs with an Mg end first
  int jh = totx + scan[1][threadIdx.x] - nh;  // then the H-H pairs
  for (int kk = 0; kk < cnt; kk++) {
    const int k = row[kk];
    if (!pair_owner(i, k)) continue;
    int& j = hi && spc[k] == 1 ? jh : jx;
    const int pos = rev_slot(nnb, nbr, k, i);
    if (pos < 0) {
      flags[2] = 1;
      out[j++] = (int)threadIdx.x | (kk << 8);  // counted: keep the list dense
      continue;
    }
    out[j++] = (int)threadIdx.x | (kk << 8) | (pos << 16);
  }
  if (threadIdx.x == 255) {
    chunk_np[ch] = totx + scan[1][255];
    chunk_npx[ch] = totx;
  }
}

// ---------------------------------------------------------------------------
// Packed SNES vectors in the caller's site order <-> the arena (internal
// order, site i of the instance holds the caller's site perm[base + i]).
// ---------------------------------------------------------------------------
// X = [q (3n) | sigma (n)]: positions and widths of the n sites into the
// evaluation state and the solver iterate; bad[0] is set for a non-finite
// position or a width that is not positive and finite.
__global__ __launch_bounds__(256) void k_gather_unknowns(int base, int n, const int* perm,
                                                         const double* X, double4* qs,
                                                         double4* Xs, int* bad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int o = perm[base + i];
  const double4 v = make_double4(X[3 * (size_t)o], X[3 * (size_t)o + 1], X[3 * (size_t)o + 2],
                                 X[3 * (size_t)n + o]);
  if (!(v.w > 0.0 && v.w < 1e300) || !(fabs(v.x) < 1e300 && fabs(v.y) < 1e300 &&
                                        fabs(v.z) < 1e300))
    bad[0] = 1;
  qs[base + i] = v;
  Xs[base + i] = v;
}

// R = [R_q (3n) | R_sigma (n) | R_x (n)] in the caller's order.
__global__ __launch_bounds__(256) void k_scatter_residual(int base, int n, const int* perm,
                                                          const double* Rq, const double* Rs,
                                                          const double* Rx, double* R) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int o = perm[base + i];
  const size_t a = base + (size_t)i;
  R[3 * (size_t)o] = Rq[3 * a];
  R[3 * (size_t)o + 1] = Rq[3 * a + 1];
  R[3 * (size_t)o + 2] = Rq[3 * a + 2];
  R[3 * (size_t)n + o] = Rs[a];
  R[4 * (size_t)n + o] = Rx[a];
}

// Packed-vector helpers of the matrix-free Jacobian-vector product.
// out[0] = sum of v[k]^2 over [0, m): one block, lanes stride through v, then
// a fixed tree in LDS (deterministic).
__global__ __launch_bounds__(1024) void k_sumsq(const double* v, size_t m, double* out) {
  __shared__ double sh[1024];
  double a = 0.0;
  for (size_t k = threadIdx.x; k < m; k += 1024) a += v[k] * v[k];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// y = x + s v over [0, m).
__global__ __launch_bounds__(256) void k_axpy(const double* x, const double* v, double s, size_t m,
                                              double* y) {
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (k < m) y[k] = x[k] + s * v[k];
}

// d[k] = (a[k] - b[k]) * c over [0, m).
__global__ __launch_bounds__(256) void k_diff_scale(const double* a, const double* b, double c,
                                                    size_t m, double* d) {
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (k < m) d[k] = (a[k] - b[k]) * c;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_neighbors(hipStream_t st, const Box& box, const CellGrid& g, double rc, int base,
                            int n, const double4* qs_all, int* cell_of, int* cell_count,
                            int* cell_start, int* cell_cursor, int* cell_atoms, int* nnb_all,
                            int* nbr_all, int* flags) {
  const double4* qs = qs_all + base;
  int* nnb = nnb_all + base;
  int* nbr = nbr_all + (size_t)base * kCap;
  int nc = g.nc[0] * g.nc[1] * g.nc[2];
  hipError_t e = hipMemsetAsync(cell_count, 0, sizeof(int) * nc, st);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(flags, 0, sizeof(int) * 3, st);
  if (e != hipSuccess) return e;
  int nb = (n + 255) / 256;
  k_cell_assign<<<nb, 256, 0, st>>>(n, box, g, qs, cell_of, cell_count);
  k_scan_cells<<<1, 1024, 0, st>>>(nc, cell_count, cell_start, cell_cursor);
  k_cell_fill<<<nb, 256, 0, st>>>(n, cell_of, cell_cursor, cell_atoms);
  k_build<<<nb, 256, 0, st>>>(base, n, box, g, rc, qs, cell_of, cell_start, cell_atoms, nnb, nbr,
                              flags);
  k_sort_rows<<<n, kCap, 0, st>>>(n, nnb, nbr);
  return hipGetLastError();
}

hipError_t launch_own_lists(hipStream_t st, int base, int n, const int* nnb, const int* nbr,
                            const int* spc, int* chunk_np, int* chunk_npx, int* own, int* flags) {
  k_own_lists<<<(n + 255) / 256, 256, 0, st>>>(base, n, nnb, nbr, spc, chunk_np, chunk_npx, own,
                                               flags);
  return hipGetLastError();
}

hipError_t launch_totals(hipStream_t st, const State& s, const Tables& T, const Outputs& o,
                         int ninst, const int* chunk_lo, const int* chunk_hi) {
  const int nb = (s.n + 255) / 256;
  k_block_sums<<<nb, 256, 0, st>>>(s.n, s.inst, s.prec, s.spc, o.rho, T.dxe[0], T.dxe[1], o.S0,
                                   o.V0, o.partials);
  k_inst_sums<<<ninst, 256, 0, st>>>(s.prec, chunk_lo, chunk_hi, o.partials, o.totals);
  return hipGetLastError();
}

hipError_t launch_gather_unknowns(hipStream_t st, int base, int n, const int* perm, const double* X,
                                  double4* qs, double4* Xs, int* bad) {
  if (n <= 0) return hipSuccess;
  k_gather_unknowns<<<(n + 255) / 256, 256, 0, st>>>(base, n, perm, X, qs, Xs, bad);
  return hipGetLastError();
}

hipError_t launch_scatter_residual(hipStream_t st, int base, int n, const int* perm,
                                   const double* Rq, const double* Rs, const double* Rx,
                                   double* R) {
  if (n <= 0) return hipSuccess;
  k_scatter_residual<<<(n + 255) / 256, 256, 0, st>>>(base, n, perm, Rq, Rs, Rx, R);
  return hipGetLastError();
}

hipError_t launch_sumsq(hipStream_t st, const double* v, size_t m, double* out) {
  k_sumsq<<<1, 1024, 0, st>>>(v, m, out);
  return hipGetLastError();
}

hipError_t launch_axpy(hipStream_t st, const double* x, const double* v, double s, size_t m,
                       double* y) {
  if (m == 0) return hipSuccess;
  k_axpy<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(x, v, s, m, y);
  return hipGetLastError();
}

hipError_t launch_diff_scale(hipStream_t st, const double* a, const double* b, double c, size_t m,
                             double* d) {
  if (m == 0) return hipSuccess;
  k_diff_scale<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(a, b, c, m, d);
  return hipGetLastError();
}

hipError_t launch_active_pairs(hipStream_t st, const State& s, long long* out) {
  k_active_pairs<<<(s.n + 255) / 256, 256, 0, st>>>(s, out);
  return hipGetLastError();
}

}  // namespace dmd
