// dmdgpu.cpp -- host side of the C ABI declared in include/dmdgpu.h.
//
// A context owns the spline tables and one device arena of sites in which any
// number of independent instances (systems) are resident at once.  Each
// instance occupies a 256-site-aligned slot range with its own box, site
// order and neighbour lists; one launch of the phase kernels evaluates every
// instance.  On top of the arena runs a batched inexact Newton-Krylov
// relaxation: every instance advances its own state machine, one residual
// evaluation per instance per tick, with the tick's device work optionally
// captured once as a hipGraph and replayed.
#include "../../include/dmdgpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dmd_hooks.h"
#include "dmd_kernels.h"
#include "dmd_noalloc.h"
#include "dmd_solver.h"

namespace {

thread_local std::string g_err;

dmdgpu_status fail(dmdgpu_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(DMDGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)
#define CHK(expr)                   \
  do {                              \
    dmdgpu_status s_ = (expr);      \
    if (s_ != DMDGPU_OK) return s_; \
  } while (0)

constexpr double kMinOcc = 5e-4;            // src/Macros.hpp:871
constexpr double kDefaultRc = 6.2934034034;  // src/Macros.hpp:857
constexpr int kChunk = dmd::kChunk;
struct Spline {
  int n = 0;
  double dx = 0.0;
  std::vector<double> c;  // n x 9 columns a b c d db dc dd ddc ddd
};

enum KernelId { KID_NEIGH = 0, KID_PHASE1, KID_PHASE2, KID_TOTALS, KID_DIAG, KID_COUNT };
const char* kKernelNames[KID_COUNT] = {"neighbour_build", "phase1_site_sums", "phase2_gradient",
                                       "totals", "diag_blocks"};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t bytes() const { return p ? sizeof(T) * std::max<size_t>(n, 1) : 0; }
  // Grows to at least cnt elements, keeping the first `keep` elements.  With
  // nothing to keep the old allocation is released first, so a growth holds
  // both copies only for buffers whose contents survive it.
  hipError_t ensure(size_t cnt, size_t keep = 0) {
    if (cnt <= n && p) return hipSuccess;
    if (!keep) release();
    T* q = nullptr;
    hipError_t e = hipMalloc((void**)&q, sizeof(T) * std::max<size_t>(cnt, 1));
    if (e != hipSuccess) return e;
    if (p && keep) {
      e = hipMemcpy(q, p, sizeof(T) * std::min(keep, n), hipMemcpyDeviceToDevice);
      if (e != hipSuccess) return e;
    }
    if (p) (void)hipFree(p);
    p = q;
    n = cnt;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

enum SolveStage {
  ST_IDLE = 0, ST_START, ST_BASE32, ST_PC, ST_CG, ST_LSPREP, ST_LS, ST_DIFF,
  ST_NG_LS, ST_NG_LSF, ST_NG_A,  // NGMRES: line-search trials, x^M, x^A
  ST_DONE
};

struct Solve {
  int stage = ST_IDLE;
  int result = 0;  // DMDGPU_RELAX_*
  dmdgpu_relax_opts o{};
  dmdgpu_diffuse_opts dop{};  // ST_DIFF
  bool diff64 = false;         // mixed diffusion: sweeps switched to FP64
  bool early = false;          // mixed relaxation, far from convergence: split residuals throughout
  bool have_r0f = false;       // R0f already holds the split residual at X
  double mass0 = 0, mass = 0;  // hydrogen content at x_n and at the iterate
  int newton = 0, cg = 0, ls = 0, evals = 0;
  int evals32 = 0;  // of those, split-precision evaluations
  double rnorm = 0, rnorm0 = 0, rr = 0, pp = 0, eps = 0, alpha_ls = 1;
  double rz = 0;          // r . M^-1 r of the preconditioned CG
  double energy = 0;      // free energy sum_i S0_i at the iterate X
  bool energy_lazy = false;  // ticks formed no energies: evaluate at the end
  double slope = 0;       // R . D, the energy's slope along the Newton step
  bool pc_setup = false;  // build the preconditioner at X before the next tick
  int ndof = 0;
  // NGMRES: history size, next slot, Gram matrix F_s . F_t of the history,
  // the smallest |F| since the last restart, failed-iteration count; line
  // search step, its previous trial and slope; step bound of the iterate;
  // x^M's |F|, F^M . F_s and step bound; the combination's coefficients.
  int nh = 0, ng_next = 0, ng_fail = 0;
  double ff[dmd::kNgHist][dmd::kNgHist] = {};
  double fmin = 0, lam = 0, lam_old = 0, dphi_old = 0, cap = 0;
  double fm = 0, fmf[dmd::kNgHist] = {}, cap_m = 0, em = 0;
  double beta[dmd::kNgHist] = {};
  dmd::SolveOp pending{};  // updates to apply at the start of the next tick
};

struct Instance {
  bool used = false;
  int base = 0, n = 0, cap = 0;
  dmd::Box box{};
  std::vector<int> perm, inv;  // internal site i holds the caller's site perm[i]
  std::vector<int> h_slot;     // species slot, internal order
  std::vector<char> h_act;     // site active (x >= 5e-4), internal order
  std::vector<double> h_q;     // positions, caller's order
  bool h_q_stale = false;      // positions changed on the device since h_q
  bool have_neigh = false;
  long long pairs = 0;     // active ordered pairs
  long long pairs_mg = 0;  // those with an Mg end
  int nactive = 0;
  Solve sv;
  // The residual buffers hold a relaxation tick's evaluation, which leaves
  // R_x out (and, with only Mg unknowns relaxed, R_q and R_sigma of the H
  // sites: hq_stale; with the residual-norm line search, the energies:
  // e_stale); rx_prec is the precision to re-evaluate in for a full residual.
  bool rx_stale = false;
  bool hq_stale = false;
  bool e_stale = false;
  int rx_prec = 0;
};

}  // namespace

// Page-locked host staging for the per-tick descriptor upload and dot
// download (asynchronous copies without a pageable bounce).
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, std::max<size_t>(b, 256), hipHostMallocDefault);
    if (e == hipSuccess) bytes = std::max<size_t>(b, 256);
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct dmdgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Second stream for the split-precision chain of a mixed evaluation; forks
  // from and joins back into `stream` through the two events.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  PinnedBuf pin_up, pin_down;
  PinnedBuf pin_pk;  // staging of the packed host-buffer residual
  double rc = kDefaultRc;
  dmd::PhaseConfig cfg{dmd::kFP64, {16, 16}, {8, 8}, {0, 1}, 3};
  // tables
  bool have[3] = {false, false, false};
  Spline rho[2], embed[2], pair[3], u[3], w[3];
  double mass[2] = {0.0, 0.0};
  bool tables_dirty = true;
  DevBuf<double> d_uw, d_ph, d_rh, d_em, d_pr, d_rr;
  DevBuf<float> d_uw32, d_ph32, d_rh32, d_pr32, d_rr32, d_rec32;
  dmd::Tables T{};
  // arena
  int extent = 0;     // sites in use, multiple of kChunk
  size_t arena_cap = 0;  // sites every per-site array holds
  size_t budget = 0;  // device bytes allowed for the arena (0 = unlimited)
  std::vector<Instance> inst;
  int primary = -1;         // instance of the one-instance API
  std::vector<int> h_inst;  // arena site -> instance
  DevBuf<int> d_inst, d_prec, d_chunk_lo, d_chunk_hi, d_noxr;
  DevBuf<dmd::Box> d_boxes;
  DevBuf<double4> d_qs;
  DevBuf<double> d_ne, d_xr, d_beta, d_gamma;
  DevBuf<int> d_spc, d_nnb, d_nbr, d_chunk_np, d_chunk_npx, d_own;
  DevBuf<int> d_perm;  // arena site -> the caller's index within its instance
  DevBuf<int> d_cell_of, d_cell_count, d_cell_start, d_cell_cursor, d_cell_atoms, d_flags;
  DevBuf<double> d_rec, d_V0, d_rho, d_Rq, d_Rs, d_Rx, d_S0, d_partials, d_totals, d_slot;
  DevBuf<double> d_diag;  // diagonal Hessian blocks (n x 9)
  DevBuf<double> d_pkin, d_pkout;  // packed X in / R out of the host-buffer residual
  // Jacobian-vector product: X and V staged from host memory (4 per site), the
  // second residual (5 per site), |V|^2
  DevBuf<double> d_jx, d_jv, d_jr, d_jscal;
  DevBuf<int> d_hookq;     // per-(u, i) queries (2 per query)
  DevBuf<double> d_hookv;  // their results (5 per query)
  DevBuf<double> d_minv;  // arena x 10: block-Jacobi preconditioner of the solver
  DevBuf<long long> d_pairs;
  std::vector<int> h_prec;
  // per instance, solver ticks: 1 = no occupancy residual needed, 2 = Mg
  // positions and widths only (State::noxr)
  std::vector<int> h_noxr;
  bool meta_dirty = true;
  // solver
  DevBuf<double4> d_X, d_R0, d_R0f, d_D, d_Rr, d_Pp;
  DevBuf<double4> d_hx, d_hf;  // NGMRES history (kNgHist per site)
  DevBuf<double> d_Xn, d_Xt;    // diffusion step: x_n and the next iterate
  DevBuf<int> d_dnn, d_dnbr;    // diffusive neighbour lists (n x kDCap)
  DevBuf<int> d_act;            // chunks to evaluate per precision (State::act)
  std::vector<int> h_chunks;
  DevBuf<dmd::SolveOp> d_ops;
  DevBuf<double> d_sdots, d_spart;
  std::vector<dmd::SolveOp> h_ops;
  std::vector<double> h_sdots;
  bool use_graph = true;
  // one captured tick per set of precisions evaluated in it (bit 0 FP64,
  // bit 1 split), so a tick launches no kernels of a precision it does not use
  hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
  int graph_extent = -1, graph_ninst = -1;
  dmd::PhaseConfig graph_cfg{};
  long long ticks = 0;
  // timing
  bool timing = false;
  struct Pending {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  long long launches[KID_COUNT] = {0};
  double ms[KID_COUNT] = {0};
};

namespace {

// Copies between host and device (or within the device) outside the kernel
// chain go through the context's stream and wait for it: the kernels run on a
// non-blocking stream, which a null-stream copy does not wait for, so a copy
// issued while an evaluation is still in flight (dmdgpu_residual_resident
// returns without waiting) would race with it.
hipError_t copy_sync(dmdgpu_ctx* c, void* dst, const void* src, size_t bytes,
                     hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

// ---- instrumentation --------------------------------------------------------
dmdgpu_status begin_timed(dmdgpu_ctx* c, hipEvent_t* a) {
  if (!c->timing) return DMDGPU_OK;
  HIPCHK(hipEventCreate(a));
  HIPCHK(hipEventRecord(*a, c->stream));
  return DMDGPU_OK;
}
dmdgpu_status end_timed(dmdgpu_ctx* c, int id, hipEvent_t a) {
  if (!c->timing) return DMDGPU_OK;
  hipEvent_t b;
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(b, c->stream));
  c->pending.push_back({id, a, b});
  return DMDGPU_OK;
}
dmdgpu_status drain_timing(dmdgpu_ctx* c) {
  if (c->pending.empty()) return DMDGPU_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, p.a, p.b));
    c->ms[p.id] += t;
    c->launches[p.id] += 1;
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  c->pending.clear();
  return DMDGPU_OK;
}

// ---- tables (src/ADP/MgHx-ADP.cpp:366-501 layout) ---------------------------
bool read_block(FILE* f, Spline& s) {
  int n;
  double dx;
  if (fscanf(f, "%d %lf", &n, &dx) != 2 || n <= 0 || !(dx > 0)) return false;
  s.n = n;
  s.dx = dx;
  s.c.assign((size_t)n * 9, 0.0);
  for (int i = 0; i < n; i++) {
    double* r = &s.c[(size_t)i * 9];
    if (fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", r, r + 1, r + 2, r + 3, r + 4, r + 5, r + 6,
               r + 7, r + 8) != 9)
      return false;
  }
  return true;
}

// Packs a spline into rows of 4 doubles: a b c d.
void pack4(const Spline& s, double* out) {
  for (int m = 0; m < s.n; m++)
    for (int k = 0; k < 4; k++) out[(size_t)m * 4 + k] = s.c[(size_t)m * 9 + k];
}

void drop_graph(dmdgpu_ctx* c);
void recompute_extent(dmdgpu_ctx* c);

dmdgpu_status upload_tables(dmdgpu_ctx* c) {
  if (!c->tables_dirty) return DMDGPU_OK;
  // Captured tick graphs hold the table pointers and grid in their kernel
  // arguments; new tables may move or resize them (after the work in flight).
  drop_graph(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!(c->have[0] && c->have[1] && c->have[2]))
    return fail(DMDGPU_ERR_STATE, "all three ADP table files (MgMg, HH, MgH) must be loaded");
  const int nr = c->rho[0].n;
  const double dx = c->rho[0].dx;
  const Spline* rt[] = {&c->rho[0], &c->rho[1], &c->pair[0], &c->pair[1], &c->pair[2],
                        &c->u[0],   &c->u[1],   &c->u[2],    &c->w[0],    &c->w[1], &c->w[2]};
  for (const Spline* s : rt)
    if (s->n != nr || s->dx != dx)
      return fail(DMDGPU_ERR_TABLES, "rho/phi/u/w tables must share one (n, dx) grid");
  std::vector<double> uw((size_t)3 * nr * 8, 0.0), ph((size_t)3 * nr * 4, 0.0),
      rh((size_t)2 * nr * 4, 0.0);
  for (int t = 0; t < 3; t++) {
    for (int m = 0; m < nr; m++) {
      double* row = &uw[((size_t)t * nr + m) * 8];
      for (int k = 0; k < 4; k++) {
        row[k] = c->u[t].c[(size_t)m * 9 + k];
        row[4 + k] = c->w[t].c[(size_t)m * 9 + k];
      }
    }
    pack4(c->pair[t], &ph[(size_t)t * nr * 4]);
  }
  for (int s = 0; s < 2; s++) pack4(c->rho[s], &rh[(size_t)s * nr * 4]);
  // fused rows for the pair-rule points: phi of the pair slot, rho of Mg, rho of H
  std::vector<double> pr((size_t)3 * nr * 16, 0.0);
  for (int t = 0; t < 3; t++)
    for (int m = 0; m < nr; m++) {
      double* row = &pr[((size_t)t * nr + m) * 16];
      for (int k = 0; k < 4; k++) {
        row[k] = ph[((size_t)t * nr + m) * 4 + k];
        row[4 + k] = rh[((size_t)0 * nr + m) * 4 + k];
        row[8 + k] = rh[((size_t)1 * nr + m) * 4 + k];
      }
    }
  // density rows for the pair-rule points of relaxation ticks: rho_Mg | rho_H
  std::vector<double> rr((size_t)nr * 8, 0.0);
  for (int m = 0; m < nr; m++)
    for (int k = 0; k < 4; k++) {
      rr[(size_t)m * 8 + k] = rh[((size_t)0 * nr + m) * 4 + k];
      rr[(size_t)m * 8 + 4 + k] = rh[((size_t)1 * nr + m) * 4 + k];
    }
  const int ne_stride = std::max(c->embed[0].n, c->embed[1].n);
  std::vector<double> em((size_t)2 * ne_stride * 4, 0.0);
  for (int s = 0; s < 2; s++) pack4(c->embed[s], &em[(size_t)s * ne_stride * 4]);
  HIPCHK(c->d_uw.ensure(uw.size()));
  HIPCHK(c->d_ph.ensure(ph.size()));
  HIPCHK(c->d_rh.ensure(rh.size()));
  HIPCHK(c->d_em.ensure(em.size()));
  HIPCHK(copy_sync(c, c->d_uw.p, uw.data(), uw.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_ph.p, ph.data(), ph.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_rh.p, rh.data(), rh.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_em.p, em.data(), em.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c->d_rr.ensure(rr.size()));
  HIPCHK(copy_sync(c, c->d_rr.p, rr.data(), rr.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c->d_pr.ensure(pr.size()));
  HIPCHK(copy_sync(c, c->d_pr.p, pr.data(), pr.size() * 8, hipMemcpyHostToDevice));
  // FP32 copies for the split-precision evaluation
  std::vector<float> uw32(uw.begin(), uw.end()), ph32(ph.begin(), ph.end()),
      rh32(rh.begin(), rh.end()), pr32(pr.begin(), pr.end()), rr32(rr.begin(), rr.end());
  HIPCHK(c->d_rr32.ensure(rr32.size()));
  HIPCHK(copy_sync(c, c->d_rr32.p, rr32.data(), rr32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c->d_pr32.ensure(pr32.size()));
  HIPCHK(copy_sync(c, c->d_pr32.p, pr32.data(), pr32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c->d_uw32.ensure(uw32.size()));
  HIPCHK(c->d_ph32.ensure(ph32.size()));
  HIPCHK(c->d_rh32.ensure(rh32.size()));
  HIPCHK(copy_sync(c, c->d_uw32.p, uw32.data(), uw32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_ph32.p, ph32.data(), ph32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_rh32.p, rh32.data(), rh32.size() * 4, hipMemcpyHostToDevice));
  dmd::Tables& T = c->T;
  T.uw = c->d_uw.p;
  T.ph = c->d_ph.p;
  T.rh = c->d_rh.p;
  T.em = c->d_em.p;
  T.uw32 = c->d_uw32.p;
  T.ph32 = c->d_ph32.p;
  T.rh32 = c->d_rh32.p;
  T.pr = c->d_pr.p;
  T.pr32 = c->d_pr32.p;
  T.rr = c->d_rr.p;
  T.rr32 = c->d_rr32.p;
  T.nr = nr;
  T.dxr = dx;
  T.inv_dxr = 1.0 / dx;
  T.dxr32 = (float)dx;
  T.inv_dxr32 = (float)(1.0 / dx);
  T.ne_stride = ne_stride;
  for (int s = 0; s < 2; s++) {
    T.ne[s] = c->embed[s].n;
    T.dxe[s] = c->embed[s].dx;
    T.inv_dxe[s] = 1.0 / c->embed[s].dx;
    T.mass[s] = c->mass[s];
  }
  c->tables_dirty = false;
  return DMDGPU_OK;
}

// ---- arena ------------------------------------------------------------------
dmd::State make_state(dmdgpu_ctx* c) {
  dmd::State s;
  s.n = c->extent;
  s.inst = c->d_inst.p;
  s.boxes = c->d_boxes.p;
  s.prec = c->d_prec.p;
  s.noxr = nullptr;  // every evaluation outside solver ticks forms R_x
  s.qs = c->d_qs.p;
  s.ne = c->d_ne.p;
  s.xr = c->d_xr.p;
  s.spc = c->d_spc.p;
  s.beta = c->d_beta.p;
  s.gamma = c->d_gamma.p;
  s.nnb = c->d_nnb.p;
  s.nbr = c->d_nbr.p;
  s.chunk_np = c->d_chunk_np.p;
  s.chunk_npx = c->d_chunk_npx.p;
  s.own = c->d_own.p;
  s.act = c->d_act.p;
  s.act_stride = (int)(c->d_act.n / 2);
  return s;
}

dmd::Outputs make_outputs(dmdgpu_ctx* c) {
  dmd::Outputs o;
  o.rec = c->d_rec.p;
  o.rec32 = c->d_rec32.p;
  o.V0 = c->d_V0.p;
  o.rho = c->d_rho.p;
  o.Rq = c->d_Rq.p;
  o.Rs = c->d_Rs.p;
  o.Rx = c->d_Rx.p;
  o.S0 = c->d_S0.p;
  o.slot = c->d_slot.p;
  o.partials = c->d_partials.p;
  o.totals = c->d_totals.p;
  return o;
}

dmd::SolveVecs make_vecs(dmdgpu_ctx* c) {
  return dmd::SolveVecs{c->d_X.p,  c->d_R0.p, c->d_R0f.p,  c->d_D.p,  c->d_Rr.p, c->d_Pp.p,
                        c->d_Xn.p, c->d_Xt.p, c->d_minv.p, c->d_hx.p, c->d_hf.p};
}

void drop_graph(dmdgpu_ctx* c) {
  for (int k = 0; k < 4; k++) {
    if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
    if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
    c->gexec[k] = nullptr;
    c->graph[k] = nullptr;
  }
  c->graph_extent = -1;
}

// The arena's per-site arrays at a capacity of `cap` sites: f(buffer,
// elements, keep) for each, keep = the contents of the sites in use survive a
// growth.  The same list sizes the arena, counts its bytes and bounds the
// memory held while it grows.
template <typename F>
void for_site_bufs(dmdgpu_ctx* c, size_t cap, F&& f) {
  const size_t ch = cap / kChunk;
  f(c->d_qs, cap, true);
  f(c->d_ne, cap, true);
  f(c->d_xr, cap, true);
  f(c->d_beta, cap, true);
  f(c->d_gamma, cap, true);
  f(c->d_spc, cap, true);
  f(c->d_inst, cap, true);
  f(c->d_perm, cap, true);
  f(c->d_nnb, cap, true);
  f(c->d_nbr, cap * dmd::kCap, true);
  f(c->d_chunk_np, ch, true);
  f(c->d_chunk_npx, ch, true);
  f(c->d_own, cap * dmd::kCap, true);
  f(c->d_slot, cap * dmd::kCap * dmd::kSlotW, false);
  f(c->d_cell_of, cap, false);
  f(c->d_cell_atoms, cap, false);
  f(c->d_rec, cap * dmd::kRec, true);
  f(c->d_rec32, cap * dmd::kRec32, true);
  f(c->d_V0, cap, true);
  f(c->d_rho, cap, true);
  f(c->d_Rq, cap * 3, true);
  f(c->d_Rs, cap, true);
  f(c->d_Rx, cap, true);
  f(c->d_S0, cap, true);
  f(c->d_partials, ch * 3, false);
  f(c->d_pairs, 2 * ch, false);
  f(c->d_spart, ch * dmd::kSolveDots, false);
  f(c->d_X, cap, true);
  f(c->d_R0, cap, true);
  f(c->d_R0f, cap, true);
  f(c->d_D, cap, true);
  f(c->d_Rr, cap, true);
  f(c->d_Pp, cap, true);
  f(c->d_hx, cap * dmd::kNgHist, true);
  f(c->d_hf, cap * dmd::kNgHist, true);
  f(c->d_Xn, cap, true);
  f(c->d_Xt, cap, true);
  f(c->d_minv, cap * 10, true);
  f(c->d_diag, cap * 9, false);
  f(c->d_dnn, cap, true);
  f(c->d_dnbr, cap * dmd::kDCap, true);
  f(c->d_act, 2 * (ch + 1), false);
  f(c->d_pkin, cap * 4, false);
  f(c->d_pkout, cap * 5, false);
  f(c->d_jx, cap * 4, false);
  f(c->d_jv, cap * 4, false);
  f(c->d_jr, cap * 5, false);
}

// Device bytes of the per-site arrays at a capacity of `cap` sites, and the
// largest of those whose contents a growth keeps (both copies of one such
// array are held at a time while the arena grows).
size_t site_bytes(dmdgpu_ctx* c, size_t cap, size_t* max_kept = nullptr) {
  size_t b = 0, mk = 0;
  for_site_bufs(c, cap, [&](auto& buf, size_t cnt, bool keep) {
    const size_t v = cnt * sizeof(*buf.p);
    b += v;
    if (keep) mk = std::max(mk, v);
  });
  if (max_kept) *max_kept = mk;
  return b;
}

// The context's device arrays outside the arena: tables, cell grid,
// per-instance metadata and solver descriptors.
template <typename F>
void for_other_bufs(dmdgpu_ctx* c, F&& f) {
  f(c->d_uw); f(c->d_ph); f(c->d_rh); f(c->d_em); f(c->d_pr); f(c->d_rr);
  f(c->d_uw32); f(c->d_ph32); f(c->d_rh32); f(c->d_pr32); f(c->d_rr32);
  f(c->d_cell_count); f(c->d_cell_start); f(c->d_cell_cursor); f(c->d_flags);
  f(c->d_boxes); f(c->d_prec); f(c->d_noxr); f(c->d_chunk_lo); f(c->d_chunk_hi);
  f(c->d_totals); f(c->d_ops); f(c->d_sdots); f(c->d_hookq); f(c->d_hookv); f(c->d_jscal);
}

// Every device allocation of the context.
size_t device_bytes(dmdgpu_ctx* c) {
  size_t b = 0;
  for_site_bufs(c, 0, [&](auto& buf, size_t, bool) { b += buf.bytes(); });
  for_other_bufs(c, [&](auto& buf) { b += buf.bytes(); });
  return b;
}

// Number of the per-site array (1, 2, ... in for_site_bufs order) whose next
// growth is refused as if the device were out of memory, from the environment
// variable DMDGPU_TEST_FAIL_GROWTH; 0 (unset): none.  Lets tests drive the
// recovery from a failed growth without exhausting the device.
int test_fail_growth() {
  const char* v = std::getenv("DMDGPU_TEST_FAIL_GROWTH");
  return v ? std::atoi(v) : 0;
}

// Makes every per-site array hold at least `sites` sites (keeping contents).
// The capacity grows by half again when it must grow at all, but never past
// what the budget allows: the context's other allocations, the new arrays and
// the one old array held during its copy must fit.  A growth that fails part
// way leaves every array holding at least the old capacity (arena_cap), so the
// resident instances stay intact; if even that cannot be restored, the
// instances are dropped.
dmdgpu_status ensure_sites(dmdgpu_ctx* c, size_t sites) {
  const size_t keep = c->extent;
  if (sites <= c->arena_cap) return DMDGPU_OK;
  auto round_up = [](size_t v) { return (v + kChunk - 1) / kChunk * kChunk; };
  size_t cap = round_up(std::max(sites, c->arena_cap * 3 / 2));
  if (c->budget) {
    size_t kept_old = 0;
    const size_t old = site_bytes(c, c->arena_cap, &kept_old);
    const size_t other = device_bytes(c) - std::min(device_bytes(c), old);
    auto fits = [&](size_t cp) { return other + site_bytes(c, cp) + kept_old <= c->budget; };
    const size_t need = round_up(sites);
    if (!fits(need))
      return fail(DMDGPU_ERR_BUDGET, "instance does not fit the device-memory budget");
    if (!fits(cap)) {  // largest chunk multiple in [need, cap) that fits
      size_t lo = need / kChunk, hi = cap / kChunk;
      while (hi - lo > 1) {
        const size_t mid = (lo + hi) / 2;
        (fits(mid * kChunk) ? lo : hi) = mid;
      }
      cap = lo * kChunk;
    }
  }
  drop_graph(c);
  HIPCHK(hipStreamSynchronize(c->stream));  // the arrays move: after the work in flight
  hipError_t err = hipSuccess;
  const int fail_at = test_fail_growth();
  int idx = 0;
  for_site_bufs(c, cap, [&](auto& buf, size_t cnt, bool kp) {
    if (err != hipSuccess) return;
    if (++idx == fail_at && cnt > buf.n) {
      if (!kp) buf.release();  // as a failed allocation of an array without contents to keep
      err = hipErrorOutOfMemory;
      return;
    }
    err = buf.ensure(cnt, kp ? cnt * keep / cap : 0);
  });
  if (err != hipSuccess) {
    // The arrays grown before the failure hold more than the old capacity and
    // those after it the old capacity; an array whose contents are not kept
    // was released before its allocation: back to the old capacity.
    const size_t oc = c->arena_cap;
    hipError_t e2 = hipSuccess;
    for_site_bufs(c, oc, [&](auto& buf, size_t cnt, bool kp) {
      if (e2 == hipSuccess && !kp && oc) e2 = buf.ensure(cnt, 0);
    });
    if (e2 != hipSuccess) {
      for (Instance& I : c->inst) I = Instance();
      std::fill(c->h_inst.begin(), c->h_inst.end(), -1);
      std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
      c->primary = -1;
      c->arena_cap = 0;
      recompute_extent(c);
      c->meta_dirty = true;
      return fail(DMDGPU_ERR_HIP, std::string("arena growth failed (") + hipGetErrorString(err) +
                                      "), and the arena could not be restored: every instance "
                                      "was removed");
    }
    return fail(DMDGPU_ERR_HIP, std::string("arena growth failed: ") + hipGetErrorString(err));
  }
  c->arena_cap = cap;
  HIPCHK(c->d_flags.ensure(3));
  return DMDGPU_OK;
}

// Per-instance metadata (boxes, chunk ranges) to the device, and tables
// loaded since the last evaluation.
dmdgpu_status upload_meta(dmdgpu_ctx* c) {
  CHK(upload_tables(c));
  if (!c->meta_dirty) return DMDGPU_OK;
  HIPCHK(hipStreamSynchronize(c->stream));  // the metadata arrays may move
  const int ni = (int)c->inst.size();
  const size_t cap = std::max(ni, 1);
  if (c->d_boxes.n < cap) drop_graph(c);
  HIPCHK(c->d_boxes.ensure(cap));
  HIPCHK(c->d_prec.ensure(cap));
  HIPCHK(c->d_noxr.ensure(cap));
  HIPCHK(c->d_chunk_lo.ensure(cap));
  HIPCHK(c->d_chunk_hi.ensure(cap));
  HIPCHK(c->d_totals.ensure(cap * 3));
  HIPCHK(c->d_ops.ensure(cap));
  HIPCHK(c->d_sdots.ensure(cap * dmd::kSolveDots));
  std::vector<dmd::Box> bx(cap);
  std::vector<int> lo(cap, 0), hi(cap, 0);
  if (c->h_prec.size() < cap) c->h_prec.resize(cap, -1);
  if (c->h_ops.size() < cap) c->h_ops.resize(cap);
  for (int k = 0; k < ni; k++) {
    const Instance& I = c->inst[k];
    if (!I.used) continue;
    bx[k] = I.box;
    lo[k] = I.base / kChunk;
    hi[k] = (I.base + I.cap) / kChunk;
  }
  HIPCHK(copy_sync(c, c->d_boxes.p, bx.data(), sizeof(dmd::Box) * cap, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_chunk_lo.p, lo.data(), sizeof(int) * cap, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_chunk_hi.p, hi.data(), sizeof(int) * cap, hipMemcpyHostToDevice));
  c->meta_dirty = false;
  return DMDGPU_OK;
}

// Active-chunk lists of both precisions from h_prec (State::act layout).
void build_act(dmdgpu_ctx* c) {
  const int stride = (int)(c->d_act.n / 2);
  c->h_chunks.assign(2 * (size_t)stride, 0);
  std::vector<std::pair<int, int>> order;  // (base, instance) in arena order
  for (int k = 0; k < (int)c->inst.size(); k++)
    if (c->inst[k].used && c->h_prec[k] >= 0) order.push_back({c->inst[k].base, k});
  std::sort(order.begin(), order.end());
  for (auto& bk : order) {
    const Instance& I = c->inst[bk.second];
    int* a = &c->h_chunks[(size_t)c->h_prec[bk.second] * stride];
    for (int ch = I.base / kChunk; ch < (I.base + I.cap) / kChunk; ch++) a[1 + a[0]++] = ch;
  }
}

dmdgpu_status upload_prec(dmdgpu_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->d_prec.p, c->h_prec.data(), sizeof(int) * c->inst.size(),
                        hipMemcpyHostToDevice, c->stream));
  build_act(c);
  HIPCHK(hipMemcpyAsync(c->d_act.p, c->h_chunks.data(), sizeof(int) * c->h_chunks.size(),
                        hipMemcpyHostToDevice, c->stream));
  return DMDGPU_OK;
}

// Cell grid with edges >= r_c over the box (periodic axes) or the bounding box
// of the mean positions (free axes).
dmd::CellGrid make_grid(const dmdgpu_ctx* c, const Instance& I) {
  dmd::CellGrid g{};
  const int n = I.n;
  for (int a = 0; a < 3; a++) {
    double lo, ext;
    if (I.box.per[a]) {
      lo = I.box.lo[a];
      ext = I.box.L[a];
    } else {
      double mn = 1e300, mx = -1e300;
      for (int i = 0; i < n; i++) {
        mn = std::min(mn, I.h_q[3 * i + a]);
        mx = std::max(mx, I.h_q[3 * i + a]);
      }
      lo = mn;
      ext = mx - mn;
    }
    int nc = (int)std::floor(ext / c->rc);
    nc = std::max(1, std::min(nc, 1024));
    g.nc[a] = nc;
    g.lo[a] = lo;
    g.cs[a] = ext > 0 ? ext / nc : 1.0;
    g.inv_cs[a] = 1.0 / g.cs[a];
  }
  return g;
}

// Internal site order: sites sorted by cell (x fastest), ties by the caller's
// index, so that a contiguous range of sites is a compact region of space and
// the neighbour records phase 2 gathers stay in one L2.
void make_order(const dmdgpu_ctx* c, Instance& I) {
  const int n = I.n;
  const dmd::CellGrid g = make_grid(c, I);
  std::vector<long long> key(n);
  for (int i = 0; i < n; i++) {
    int cc[3];
    for (int a = 0; a < 3; a++) {
      double v = I.h_q[3 * i + a] - g.lo[a];
      if (I.box.per[a]) v = v - I.box.L[a] * std::floor(v / I.box.L[a]);
      int k = (int)std::floor(v * g.inv_cs[a]);
      cc[a] = k < 0 ? 0 : (k > g.nc[a] - 1 ? g.nc[a] - 1 : k);
    }
    key[i] = ((long long)cc[2] * g.nc[1] + cc[1]) * g.nc[0] + cc[0];
  }
  I.perm.resize(n);
  const size_t ncell = (size_t)g.nc[0] * g.nc[1] * g.nc[2];
  if (ncell <= 4 * (size_t)n + 4096) {
    // counting sort by cell: stable, so the order of sorting by (cell, index)
    std::vector<int> start(ncell + 1, 0);
    for (int i = 0; i < n; i++) start[key[i] + 1]++;
    for (size_t k = 0; k < ncell; k++) start[k + 1] += start[k];
    for (int i = 0; i < n; i++) I.perm[start[key[i]]++] = i;
  } else {  // sparse grid (sites spread over a large free box)
    for (int i = 0; i < n; i++) I.perm[i] = i;
    std::stable_sort(I.perm.begin(), I.perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  I.inv.resize(n);
  for (int i = 0; i < n; i++) I.inv[I.perm[i]] = i;
}

// Copies the caller's unknowns of instance I into its arena range (internal
// order).  NULL leaves a field unchanged.
dmdgpu_status upload_unknowns(dmdgpu_ctx* c, Instance& I, const double* q, const double* sg,
                              const double* x, bool first) {
  const int n = I.n;
  for (size_t k = 0; q && k < 3 * (size_t)n; k++)
    if (!std::isfinite(q[k])) return fail(DMDGPU_ERR_ARG, "positions must be finite");
  for (int k = 0; x && k < n; k++)
    if (!std::isfinite(x[k])) return fail(DMDGPU_ERR_ARG, "occupancies must be finite");
  if (q) {
    I.h_q.assign(q, q + 3 * (size_t)n);
    I.h_q_stale = false;
  }
  if (q || sg) {
    std::vector<double4> qs(n);
    if (!first)
      HIPCHK(copy_sync(c, qs.data(), c->d_qs.p + I.base, sizeof(double4) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
      const int o = I.perm[i];
      if (q) {
        qs[i].x = q[3 * o];
        qs[i].y = q[3 * o + 1];
        qs[i].z = q[3 * o + 2];
      }
      if (sg) {
        if (!(sg[o] > 0) || !std::isfinite(sg[o]))
          return fail(DMDGPU_ERR_ARG, "widths must be positive and finite");
        qs[i].w = sg[o];
      }
    }
    HIPCHK(copy_sync(c, c->d_qs.p + I.base, qs.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    HIPCHK(copy_sync(c, c->d_X.p + I.base, c->d_qs.p + I.base, sizeof(double4) * n,
                     hipMemcpyDeviceToDevice));
  }
  if (x) {
    std::vector<double> ne(n), xr(n);
    int na = 0;
    for (int i = 0; i < n; i++) {
      const int o = I.perm[i];
      if (I.h_slot[i] == 0 && x[o] < kMinOcc)
        return fail(DMDGPU_ERR_SPECIES, "Mg site with occupancy below 5e-4 (site " +
                                            std::to_string(o) + ")");
      xr[i] = x[o];
      ne[i] = x[o] >= kMinOcc ? x[o] : 0.0;
      na += ne[i] != 0.0;
    }
    I.nactive = na;
    I.h_act.resize(n);
    for (int i = 0; i < n; i++) I.h_act[i] = ne[i] != 0.0;
    HIPCHK(copy_sync(c, c->d_ne.p + I.base, ne.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(copy_sync(c, c->d_xr.p + I.base, xr.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  return DMDGPU_OK;
}

// First-fit placement of `cap` sites in the arena.
int place(const dmdgpu_ctx* c, int cap) {
  std::vector<std::pair<int, int>> ranges;
  for (const Instance& I : c->inst)
    if (I.used) ranges.push_back({I.base, I.base + I.cap});
  std::sort(ranges.begin(), ranges.end());
  int at = 0;
  for (auto& r : ranges) {
    if (r.first - at >= cap) return at;
    at = std::max(at, r.second);
  }
  return at;
}

void recompute_extent(dmdgpu_ctx* c) {
  int e = 0;
  for (const Instance& I : c->inst)
    if (I.used) e = std::max(e, I.base + I.cap);
  if (e != c->extent) drop_graph(c);
  c->extent = e;
}

dmdgpu_status build_neighbors(dmdgpu_ctx* c, Instance& I) {
  // The lists are rebuilt in place: until the rebuild succeeds they are not
  // valid, and no evaluation may read them.
  I.have_neigh = false;
  I.pairs = I.pairs_mg = 0;
  if (I.h_q_stale) {  // positions were last set on the device
    std::vector<double4> qs(I.n);
    HIPCHK(copy_sync(c, qs.data(), c->d_qs.p + I.base, sizeof(double4) * I.n, hipMemcpyDeviceToHost));
    for (int i = 0; i < I.n; i++) {
      const int o = I.perm[i];
      I.h_q[3 * o] = qs[i].x;
      I.h_q[3 * o + 1] = qs[i].y;
      I.h_q[3 * o + 2] = qs[i].z;
    }
    I.h_q_stale = false;
  }
  const dmd::CellGrid g = make_grid(c, I);
  const size_t ncell = (size_t)g.nc[0] * g.nc[1] * g.nc[2];
  if (c->budget) {
    const size_t grow = sizeof(int) * ((ncell > c->d_cell_count.n ? ncell : 0) +
                                       (ncell + 1 > c->d_cell_start.n ? ncell + 1 : 0) +
                                       (ncell > c->d_cell_cursor.n ? ncell : 0));
    if (device_bytes(c) + grow > c->budget)
      return fail(DMDGPU_ERR_BUDGET, "the cell grid does not fit the device-memory budget");
  }
  HIPCHK(hipStreamSynchronize(c->stream));  // the cell arrays may move
  HIPCHK(c->d_cell_count.ensure(ncell));
  HIPCHK(c->d_cell_start.ensure(ncell + 1));
  HIPCHK(c->d_cell_cursor.ensure(ncell));
  hipEvent_t ev{};
  CHK(begin_timed(c, &ev));
  HIPCHK(dmd::launch_neighbors(c->stream, I.box, g, c->rc, I.base, I.n, c->d_qs.p, c->d_cell_of.p,
                               c->d_cell_count.p, c->d_cell_start.p, c->d_cell_cursor.p,
                               c->d_cell_atoms.p, c->d_nnb.p, c->d_nbr.p, c->d_flags.p));
  HIPCHK(dmd::launch_own_lists(c->stream, I.base, I.n, c->d_nnb.p, c->d_nbr.p, c->d_spc.p,
                               c->d_chunk_np.p, c->d_chunk_npx.p, c->d_own.p, c->d_flags.p));
  CHK(end_timed(c, KID_NEIGH, ev));
  int flags[3];
  HIPCHK(hipMemcpyAsync(flags, c->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (flags[0]) return fail(DMDGPU_ERR_MAXNEIGH, "a site has more than 250 neighbours");
  if (flags[1]) return fail(DMDGPU_ERR_NONEIGH, "a site has no neighbours");
  if (flags[2]) return fail(DMDGPU_ERR_HIP, "neighbour lists are not symmetric");
  // active ordered pairs of this instance
  CHK(upload_meta(c));
  HIPCHK(dmd::launch_active_pairs(c->stream, make_state(c), c->d_pairs.p));
  const int c0 = I.base / kChunk, nch = I.cap / kChunk;
  std::vector<long long> pp(2 * (size_t)nch);
  HIPCHK(hipMemcpyAsync(pp.data(), c->d_pairs.p + 2 * c0, sizeof(long long) * 2 * nch,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  I.pairs = I.pairs_mg = 0;
  for (int k = 0; k < nch; k++) {
    I.pairs += pp[2 * k];
    I.pairs_mg += pp[2 * k + 1];
  }
  I.have_neigh = true;
  return DMDGPU_OK;
}

dmdgpu_status remove_instance(dmdgpu_ctx* c, int id) {
  if (id < 0 || id >= (int)c->inst.size() || !c->inst[id].used)
    return fail(DMDGPU_ERR_ARG, "no such instance");
  Instance& I = c->inst[id];
  for (int i = 0; i < I.cap; i++) c->h_inst[I.base + i] = -1;
  HIPCHK(copy_sync(c, c->d_inst.p + I.base, c->h_inst.data() + I.base, sizeof(int) * I.cap,
                   hipMemcpyHostToDevice));
  I = Instance();
  c->h_prec[id] = -1;
  if (c->primary == id) c->primary = -1;
  recompute_extent(c);
  c->meta_dirty = true;
  return DMDGPU_OK;
}

dmdgpu_status add_instance(dmdgpu_ctx* c, int n, const double* q, const double* sg, const double* x,
                           const int32_t* spc, const double* beta, const double* gamma,
                           const double* box, const int32_t* periodic, int* out_id) {
  if (n <= 0 || !q || !sg || !x || !spc || !beta || !gamma || !box || !periodic)
    return fail(DMDGPU_ERR_ARG, "null argument or n <= 0");
  if (n > dmd::kMaxInstanceSites)
    return fail(DMDGPU_ERR_ARG, "an instance holds at most " +
                                    std::to_string(dmd::kMaxInstanceSites) + " sites");
  HIPCHK(hipSetDevice(c->device));
  for (int i = 0; i < n; i++)
    if (spc[i] != 1 && spc[i] != 12)
      return fail(DMDGPU_ERR_SPECIES,
                  "species other than H (1) or Mg (12) at site " + std::to_string(i));
  for (size_t k = 0; k < 3 * (size_t)n; k++)
    if (!std::isfinite(q[k])) return fail(DMDGPU_ERR_ARG, "positions must be finite");
  for (int k = 0; k < 6; k++)
    if (!std::isfinite(box[k])) return fail(DMDGPU_ERR_ARG, "box bounds must be finite");
  Instance I;
  for (int a = 0; a < 3; a++) {
    I.box.lo[a] = box[a];
    I.box.L[a] = box[3 + a] - box[a];
    I.box.per[a] = periodic[a] ? 1 : 0;
    if (I.box.per[a] && !(I.box.L[a] >= 2.0 * c->rc))
      return fail(DMDGPU_ERR_BOX, "periodic box edge shorter than 2 r_c");
    I.box.invL[a] = I.box.L[a] > 0 ? 1.0 / I.box.L[a] : 0.0;
  }
  CHK(upload_tables(c));
  I.n = n;
  I.cap = (n + kChunk - 1) / kChunk * kChunk;
  I.base = place(c, I.cap);
  I.box.base = I.base;
  const size_t need = (size_t)I.base + I.cap;
  // The arena in use after placement, with the context's other allocations,
  // must fit the budget (ensure_sites also bounds what a growth allocates).
  if (c->budget) {
    const size_t held = site_bytes(c, c->arena_cap);
    const size_t other = device_bytes(c) - std::min(device_bytes(c), held);
    if (other + site_bytes(c, need) > c->budget)
      return fail(DMDGPU_ERR_BUDGET, "instance does not fit the device-memory budget");
  }
  CHK(ensure_sites(c, need));
  I.used = true;
  I.h_q.assign(q, q + 3 * (size_t)n);
  make_order(c, I);
  I.h_slot.resize(n);
  std::vector<double> bp(I.cap, 0.0), gp(I.cap, 0.0);
  std::vector<int> sl(I.cap, 0), pm(I.cap, -1);
  for (int i = 0; i < n; i++) {
    const int o = I.perm[i];
    I.h_slot[i] = spc[o] == 1 ? 1 : 0;
    sl[i] = I.h_slot[i];
    pm[i] = o;
    bp[i] = beta[o];
    gp[i] = gamma[o];
  }
  int id = -1;
  for (int k = 0; k < (int)c->inst.size(); k++)
    if (!c->inst[k].used) {
      id = k;
      break;
    }
  if (id < 0) {
    id = (int)c->inst.size();
    c->inst.emplace_back();
  }
  const int b = I.base, cp = I.cap;
  HIPCHK(copy_sync(c, c->d_spc.p + b, sl.data(), sizeof(int) * cp, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_perm.p + b, pm.data(), sizeof(int) * cp, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_beta.p + b, bp.data(), sizeof(double) * cp, hipMemcpyHostToDevice));
  HIPCHK(copy_sync(c, c->d_gamma.p + b, gp.data(), sizeof(double) * cp, hipMemcpyHostToDevice));
  // Zero on the device (no host copies): occupancies (the padding sites of the
  // range stay inactive, x = 0), neighbour counts and the solver vectors; done
  // before the sites' own values are written below.
  {
    auto zero = [&](auto* p, size_t cnt) {
      return hipMemsetAsync(p, 0, sizeof(*p) * cnt, c->stream);
    };
    HIPCHK(zero(c->d_ne.p + b, cp));
    HIPCHK(zero(c->d_xr.p + b, cp));
    HIPCHK(zero(c->d_nnb.p + b, cp));
    HIPCHK(zero(c->d_chunk_np.p + b / kChunk, cp / kChunk));
    HIPCHK(zero(c->d_chunk_npx.p + b / kChunk, cp / kChunk));
    HIPCHK(zero(c->d_dnn.p + b, cp));
    HIPCHK(zero(c->d_Xn.p + b, cp));
    HIPCHK(zero(c->d_Xt.p + b, cp));
    double4* vz[] = {c->d_R0.p, c->d_R0f.p, c->d_D.p, c->d_Rr.p, c->d_Pp.p};
    for (double4* v : vz) HIPCHK(zero(v + b, cp));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (cp > n) {  // padding sites: unit width at the origin (the sites' own are uploaded below)
    std::vector<double4> q0(cp - n, make_double4(0.0, 0.0, 0.0, 1.0));
    HIPCHK(copy_sync(c, c->d_qs.p + b + n, q0.data(), sizeof(double4) * (cp - n),
                     hipMemcpyHostToDevice));
    HIPCHK(copy_sync(c, c->d_X.p + b + n, q0.data(), sizeof(double4) * (cp - n),
                     hipMemcpyHostToDevice));
  }
  c->inst[id] = std::move(I);
  Instance& J = c->inst[id];
  dmdgpu_status st = upload_unknowns(c, J, q, sg, x, true);
  if (st) {
    J = Instance();
    return st;
  }
  if ((int)c->h_inst.size() < J.base + J.cap) c->h_inst.resize(J.base + J.cap, -1);
  for (int i = 0; i < J.cap; i++) c->h_inst[J.base + i] = id;
  HIPCHK(copy_sync(c, c->d_inst.p + J.base, c->h_inst.data() + J.base, sizeof(int) * J.cap,
                   hipMemcpyHostToDevice));
  recompute_extent(c);
  c->meta_dirty = true;
  if (c->h_prec.size() < c->inst.size()) c->h_prec.resize(c->inst.size(), -1);
  if (c->h_ops.size() < c->inst.size()) c->h_ops.resize(c->inst.size());
  c->h_prec[id] = -1;
  *out_id = id;
  return DMDGPU_OK;
}

Instance* get_inst(dmdgpu_ctx* c, int id) {
  if (!c || id < 0 || id >= (int)c->inst.size() || !c->inst[id].used) return nullptr;
  return &c->inst[id];
}

// FP64 chain on the main stream, split-precision chain on the side stream,
// joined back into the main stream (also valid under stream capture, where
// the fork and join become graph edges).
dmdgpu_status eval_chains(dmdgpu_ctx* c, const dmd::State& s, const dmd::Outputs& o,
                          bool phase2) {
  HIPCHK(hipEventRecord(c->ev_fork, c->stream));
  HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  HIPCHK(dmd::launch_phase1(c->side, c->cfg, dmd::kSplit32, s, c->T, o));
  if (phase2) HIPCHK(dmd::launch_phase2(c->side, c->cfg, dmd::kSplit32, s, c->T, o));
  HIPCHK(dmd::launch_phase1(c->stream, c->cfg, dmd::kFP64, s, c->T, o));
  if (phase2) HIPCHK(dmd::launch_phase2(c->stream, c->cfg, dmd::kFP64, s, c->T, o));
  HIPCHK(hipEventRecord(c->ev_join, c->side));
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return DMDGPU_OK;
}

// One residual evaluation of every instance whose h_prec >= 0.
dmdgpu_status run_eval(dmdgpu_ctx* c, bool phase2 = true) {
  CHK(upload_meta(c));
  CHK(upload_prec(c));
  bool any[2] = {false, false};
  for (int k = 0; k < (int)c->inst.size(); k++)
    if (c->inst[k].used && c->h_prec[k] >= 0) {
      any[c->h_prec[k]] = true;
      if (phase2) c->inst[k].rx_stale = c->inst[k].hq_stale = c->inst[k].e_stale = false;
    }
  const dmd::State s = make_state(c);
  const dmd::Outputs o = make_outputs(c);
  if (any[0] && any[1] && !c->timing) {
    // Both precisions present: the two chains touch disjoint instances, so
    // they run concurrently and fill the device together.
    CHK(eval_chains(c, s, o, phase2));
  } else {
    hipEvent_t a{};
    CHK(begin_timed(c, &a));
    for (int p = 0; p < 2; p++)
      if (any[p]) HIPCHK(dmd::launch_phase1(c->stream, c->cfg, p, s, c->T, o));
    CHK(end_timed(c, KID_PHASE1, a));
    if (phase2) {
      CHK(begin_timed(c, &a));
      for (int p = 0; p < 2; p++)
        if (any[p]) HIPCHK(dmd::launch_phase2(c->stream, c->cfg, p, s, c->T, o));
      CHK(end_timed(c, KID_PHASE2, a));
    }
  }
  // Totals (after phase 1 alone: the table-domain count of the new densities;
  // the energy sums then still hold the last full evaluation's S0)
  hipEvent_t a{};
  CHK(begin_timed(c, &a));
  HIPCHK(dmd::launch_totals(c->stream, s, c->T, o, (int)c->inst.size(), c->d_chunk_lo.p,
                            c->d_chunk_hi.p));
  CHK(end_timed(c, KID_TOTALS, a));
  return DMDGPU_OK;
}

// Evaluates one instance in the context's default precision.
dmdgpu_status eval_one(dmdgpu_ctx* c, int id, bool phase2 = true) {
  std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
  c->h_prec[id] = c->cfg.precision;
  return run_eval(c, phase2);
}

dmdgpu_status eval_all(dmdgpu_ctx* c) {
  for (int k = 0; k < (int)c->inst.size(); k++)
    c->h_prec[k] = c->inst[k].used && c->inst[k].have_neigh ? c->cfg.precision : -1;
  return run_eval(c, true);
}

// Status of an evaluation from its count of sites whose density left the
// embedding table's domain (the third per-instance total).
dmdgpu_status domain_status(double count) {
  if (count > 0.0)
    return fail(DMDGPU_ERR_DOMAIN, std::to_string((long long)count) +
                                       " site(s) have a density at or below -drho of the "
                                       "embedding table (outside the spline's domain)");
  return DMDGPU_OK;
}

// The table-domain status of the last evaluation of instance id.
dmdgpu_status check_domain(dmdgpu_ctx* c, int id) {
  double v = 0.0;
  HIPCHK(hipMemcpyAsync(&v, c->d_totals.p + 3 * id + 2, sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return domain_status(v);
}

// Copies `width` doubles per site of instance I back to the caller's order.
dmdgpu_status d2h_sites(dmdgpu_ctx* c, const Instance& I, double* h, const double* d, int width) {
  if (!h) return DMDGPU_OK;
  const int n = I.n;
  std::vector<double> tmp((size_t)n * width);
  HIPCHK(hipMemcpyAsync(tmp.data(), d + (size_t)I.base * width, tmp.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++)
    for (int k = 0; k < width; k++) h[(size_t)I.perm[i] * width + k] = tmp[(size_t)i * width + k];
  return DMDGPU_OK;
}

dmdgpu_status copy_residual(dmdgpu_ctx* c, int id, double* Rq, double* Rs, double* Rx,
                            double* tot) {
  Instance& I = c->inst[id];
  if ((Rx && I.rx_stale) || ((Rq || Rs) && I.hq_stale) || (tot && I.e_stale)) {
    // Last evaluated by a relaxation tick: evaluate again, in the same
    // precision, at the same point, with R_x, the H sites' sums and the
    // energies.
    std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
    c->h_prec[id] = I.rx_prec;
    CHK(run_eval(c, true));
  }
  CHK(d2h_sites(c, I, Rq, c->d_Rq.p, 3));
  CHK(d2h_sites(c, I, Rs, c->d_Rs.p, 1));
  CHK(d2h_sites(c, I, Rx, c->d_Rx.p, 1));
  double t3[3];
  HIPCHK(hipMemcpyAsync(t3, c->d_totals.p + 3 * id, sizeof(t3), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (tot) {
    tot[0] = t3[0];
    tot[1] = t3[1];
  }
  return domain_status(t3[2]);
}

dmdgpu_status check_primary(dmdgpu_ctx* c) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  if (c->primary < 0) return fail(DMDGPU_ERR_STATE, "dmdgpu_set_state has not been called");
  if (!c->inst[c->primary].have_neigh)
    return fail(DMDGPU_ERR_STATE, "dmdgpu_neighbors has not been called");
  return DMDGPU_OK;
}

// ---- batched Newton-Krylov relaxation ----------------------------------------
//
// Per instance (unknowns: q and sigma of its active sites; occupancies fixed):
//   START  : FP64 residual at X; stop when |R|_2 <= atol.
//   BASE32 : (mixed) split-precision residual at X, the base of FP32 probes.
//   PC     : (precond) P = M^-1 r with the block-Jacobi M built at X between
//            ticks (pc_setup); no evaluation in this tick.
//   CG     : conjugate gradients (preconditioned with precond) on H d = -R with
//            H v = (R(X + eps v) - R(X))/eps; stop at |r| <= eta |R|, max_cg
//            iterations or negative curvature.
//   LSPREP : step cap so that max |dq| <= 0.2 A and sigma stays >= sigma/2.
//   LS     : backtracking on |R(X + a d)| (sufficient decrease), FP64.
// Every stage that needs a residual gets exactly one per tick.
// RMS displacement of a Hessian probe (A).  The one-sided difference has a
// truncation error ~ probe |d3 S| / |d2 S| and a rounding error ~ noise /
// probe: 1e-7 balances them for FP64 residuals, 5e-5 for split-precision
// ones (RMS noise ~1e-6 per component).
constexpr double kProbe64 = 1e-7;
constexpr double kProbe32 = 5e-5;
constexpr double kMaxStep = 0.2;   // max |dq| of a Newton step (A)
// NGMRES (Oosterlee and Washio, SIAM J. Sci. Comput. 21 (2000) 1670; the
// selection and restart tests with the constants of PETSc's SNESNGMRES
// defaults): x^A is kept when |F^A| < kNgGammaA min|F| and either
// kNgEpsB |x^A - x^M| < min_s |x^A - x_s| or |F^A|^2 < kNgDeltaB min|F|^2;
// the history restarts after kNgRestart iterations in a row with
// |F| > kNgGammaC min|F|.  The critical-point line search takes at most
// kNgLsIts secant trials from lambda = min(2 lambda_prev, bound), lambda_0 =
// kNgLam0 (A^2/eV).
constexpr double kNgGammaA = 2.0, kNgGammaC = 2.0, kNgEpsB = 0.1, kNgDeltaB = 0.9;
constexpr int kNgRestart = 2, kNgLsIts = 3;
constexpr double kNgLam0 = 1e-3;
// Mixed relaxation far from convergence: while |R|_2 > kSplitSwitch sqrt(ndof)
// -- a thousand times the split residual's RMS noise of ~1e-6 per component --
// the residuals of the Newton iteration and of its line search are split
// evaluations too (they also serve as the base of the split probes, so no
// separate base evaluation is needed).  Below it every accepted residual and
// the stopping test are FP64.
constexpr double kSplitSwitch = 1e-3;

// Relative rounding level of the free-energy sum below which the energy merit
// cannot tell a decrease (sum of ~1e5 site terms in FP64).
constexpr double kEnergyRound = 1e-11;

constexpr double kA2PerM2 = 1e20;  // diffusivity m^2/s -> A^2/s
// Mixed diffusion steps sweep with split-precision residuals until the
// iterate moves by less than this, then finish with FP64 sweeps (the split
// residual's occupancy error bounds what FP32 sweeps can resolve).
constexpr double kDiffSwitch = 1e-7;

void set_diff_op(dmd::SolveOp& op, const dmdgpu_diffuse_opts& d, int mode) {
  op.diff = mode;
  op.dt = d.dt;
  op.d0 = d.D0 * kA2PerM2;
  op.eb = d.Eb;
}

void plan_tick(dmdgpu_ctx* c) {
  const int ni = (int)c->inst.size();
  for (int k = 0; k < ni; k++) {
    Instance& I = c->inst[k];
    dmd::SolveOp op = I.used ? I.sv.pending : dmd::SolveOp{};
    op.make = dmd::kMakeNone;
    op.dot = dmd::kDotNone;
    op.mixed = I.sv.o.mixed;
    op.dof = I.sv.o.dof_mask;
    op.pc = I.sv.o.precond;
    c->h_prec[k] = -1;
    if (I.used) {
      Solve& S = I.sv;
      switch (S.stage) {
        case ST_START:
          op.make = dmd::kMakeX;
          op.dot = S.o.solver ? dmd::kDotNg : dmd::kDotNorm;
          c->h_prec[k] = S.early ? dmd::kSplit32 : dmd::kFP64;
          break;
        case ST_BASE32:
          op.make = dmd::kMakeX;
          c->h_prec[k] = dmd::kSplit32;
          break;
        case ST_PC:
          op.dot = dmd::kDotPc;  // CgInit with P = M^-1 Rr; no evaluation
          break;
        case ST_CG:
          S.eps = (S.o.probe > 0 ? S.o.probe : (S.o.mixed ? kProbe32 : kProbe64)) *
                  std::sqrt((double)S.ndof / std::max(S.pp, 1e-300));
          op.eps = S.eps;
          op.make = dmd::kMakeXP;
          op.dot = dmd::kDotCg;
          c->h_prec[k] = S.o.mixed ? dmd::kSplit32 : dmd::kFP64;
          break;
        case ST_LSPREP:
          op.dot = dmd::kDotLs;
          break;
        case ST_LS:
          op.make = dmd::kMakeXD;
          op.alpha_ls = S.alpha_ls;
          op.dot = dmd::kDotNorm;
          c->h_prec[k] = S.early ? dmd::kSplit32 : dmd::kFP64;
          break;
        case ST_NG_LS:
        case ST_NG_LSF:
          op.make = dmd::kMakeXR;  // X - lambda R0
          op.alpha_ls = S.lam;
          op.dot = dmd::kDotNg;
          op.nh = S.nh;
          c->h_prec[k] = dmd::kFP64;
          break;
        case ST_NG_A:
          op.make = dmd::kMakeNgA;
          for (int t = 0; t < dmd::kNgHist; t++) op.b[t] = S.beta[t];
          op.dot = dmd::kDotNg;
          op.nh = S.nh;
          c->h_prec[k] = dmd::kFP64;
          break;
        case ST_DIFF:
          set_diff_op(op, S.dop, dmd::kDiffStep);
          op.dot = dmd::kDotDiff;
          c->h_prec[k] = S.dop.mixed && !S.diff64 ? dmd::kSplit32 : dmd::kFP64;
          break;
        default:
          break;
      }
      if (c->h_prec[k] >= 0) S.evals++;
      if (c->h_prec[k] == dmd::kSplit32) S.evals32++;
      S.pending = dmd::SolveOp{};
    }
    c->h_ops[k] = op;
  }
}

void finish(Solve& S, int result) {
  S.stage = ST_DONE;
  S.result = result;
}

// Start conjugate gradients at X (|R|^2 = rr): D = 0, Rr = -R0, P = Rr.  With
// the preconditioner the host first builds M at X (between ticks), and the
// next tick forms P = M^-1 Rr and its dots (ST_PC) before the first probe.
void begin_cg(Solve& S, dmd::SolveOp& nx, double rr) {
  nx.upd |= dmd::kUpdCgInit;
  S.rr = S.pp = rr;
  S.cg = 0;
  if (S.o.precond) {
    // rebuilt every `precond` Newton steps (lagged in between: still SPD)
    S.stage = ST_PC;
    S.pc_setup = S.newton % S.o.precond == 0;
  } else {
    S.stage = ST_CG;
  }
}

// ---- NGMRES ------------------------------------------------------------------
// Step bound of an iterate whose residual has max |F_q| = dq (component) and
// max F_sigma / sigma = ds over relaxed widths: the step -lambda F moves no
// position by more than kMaxStep and no width below half its value.
double ng_bound(double dq, double ds) {
  double b = 1e300;
  if (dq > 0.0) b = std::min(b, kMaxStep / dq);
  if (ds > 0.0) b = std::min(b, 0.5 / ds);
  return b;
}

// Line search from the iterate along -F (|F|^2 = f2): the first trial at S.lam.
void ng_ls_begin(Solve& S, double f2) {
  S.lam_old = 0.0;
  S.dphi_old = -f2;  // d/dlambda of the energy at lambda = 0: -F . F
  S.ls = 0;
  S.stage = ST_NG_LS;
}

// Coefficients beta of x^A = x^M + sum_s beta_s (x_s - x^M) minimising the
// linearised residual |F^M + sum_s beta_s (F_s - F^M)|: normal equations from
// the history's Gram matrix and F^M . F_s, solved by Cholesky with a diagonal
// shift of 1e-13 of the largest diagonal entry (a Gram matrix is only
// semi-definite when residuals repeat).
void ng_combine(Solve& S) {
  const int n = S.nh;
  const double fm2 = S.fm * S.fm;
  double H[dmd::kNgHist][dmd::kNgHist], r[dmd::kNgHist];
  double dmax = 0.0;
  for (int i = 0; i < n; i++) {
    r[i] = fm2 - S.fmf[i];
    for (int j = 0; j < n; j++) H[i][j] = S.ff[i][j] - S.fmf[i] - S.fmf[j] + fm2;
    dmax = std::max(dmax, H[i][i]);
  }
  for (int t = 0; t < dmd::kNgHist; t++) S.beta[t] = 0.0;
  if (!(dmax > 0.0) || !std::isfinite(dmax)) return;
  for (int i = 0; i < n; i++) H[i][i] += 1e-13 * dmax;
  double L[dmd::kNgHist][dmd::kNgHist] = {};
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      double v = H[i][j];
      for (int k = 0; k < j; k++) v -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(v > 0.0)) return;  // not positive definite: x^A = x^M
        L[i][i] = std::sqrt(v);
      } else {
        L[i][j] = v / L[j][j];
      }
    }
  double y[dmd::kNgHist];
  for (int i = 0; i < n; i++) {
    double v = r[i];
    for (int k = 0; k < i; k++) v -= L[i][k] * y[k];
    y[i] = v / L[i][i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double v = y[i];
    for (int k = i + 1; k < n; k++) v -= L[k][i] * S.beta[k];
    S.beta[i] = v / L[i][i];
  }
  for (int i = 0; i < n; i++)
    if (!std::isfinite(S.beta[i])) {
      for (int t = 0; t < dmd::kNgHist; t++) S.beta[t] = 0.0;
      return;
    }
}

// The first residual (at X): the history holds (X, F) only.
void ng_start(Solve& S, dmd::SolveOp& nx, const double* d) {
  S.rnorm = std::sqrt(d[0]);
  S.energy = d[dmd::kDotEnergy];
  if (S.newton == 0) S.rnorm0 = S.rnorm;
  nx.upd |= dmd::kUpdR0FromRv;
  if (!std::isfinite(S.rnorm)) {
    finish(S, DMDGPU_RELAX_STALLED);
    return;
  }
  if (S.rnorm <= S.o.atol) {
    finish(S, DMDGPU_RELAX_CONVERGED);
    return;
  }
  nx.upd |= dmd::kUpdNgPush;
  nx.slot = 0;
  S.nh = 1;
  S.ng_next = 1 % S.o.ngmres_m;
  S.ng_fail = 0;
  S.ff[0][0] = d[0];
  S.fmin = S.rnorm;
  S.cap = ng_bound(d[3], d[4]);
  S.lam = std::min(kNgLam0, S.cap);
  ng_ls_begin(S, d[0]);
}

// One NGMRES stage's dots (ST_NG_LS, ST_NG_LSF, ST_NG_A).
void ng_digest(Solve& S, dmd::SolveOp& nx, const double* d) {
  const bool finite = std::isfinite(d[0]) && std::isfinite(d[1]) && d[2] == 0.0;
  if (S.stage == ST_NG_LS || S.stage == ST_NG_LSF) {
    if (!finite) {  // a trial left the residual's domain: shorter steps
      S.lam *= 0.5;
      if (++S.ls > 12) finish(S, DMDGPU_RELAX_STALLED);
      return;
    }
  }
  if (S.stage == ST_NG_LS) {
    // secant step on dphi(lambda) = -F(X - lambda F) . F towards its zero
    const double dphi = -d[1];
    double ln = S.lam - dphi * (S.lam - S.lam_old) / (dphi - S.dphi_old);
    if (!std::isfinite(ln) || !(ln > 0.0)) ln = dphi < 0.0 ? 2.0 * S.lam : 0.5 * S.lam;
    ln = std::min(ln, S.cap);
    const bool done = ++S.ls >= kNgLsIts || std::fabs(ln - S.lam) <= 1e-2 * ln;
    S.lam_old = S.lam;
    S.dphi_old = dphi;
    S.lam = ln;
    if (done) S.stage = ST_NG_LSF;
    return;
  }
  if (S.stage == ST_NG_LSF) {
    // x^M = X - lambda F and F^M: keep them (D, Rr), then the combination
    nx.upd |= dmd::kUpdNgKeepM;
    S.fm = std::sqrt(d[0]);
    S.em = d[dmd::kDotEnergy];
    for (int t = 0; t < S.nh; t++) S.fmf[t] = d[5 + t];
    S.cap_m = ng_bound(d[3], d[4]);
    ng_combine(S);
    S.stage = ST_NG_A;
    return;
  }
  // ST_NG_A: select x^A or x^M, push it into the history
  const double fa = std::sqrt(d[0]);
  double dmin = 1e300;
  for (int t = 0; t < S.nh; t++) dmin = std::min(dmin, d[9 + t]);
  dmin = std::sqrt(dmin);
  const double dn = std::sqrt(d[8]);
  const bool take_a = finite && fa < kNgGammaA * S.fmin &&
                      (kNgEpsB * dn < dmin || fa * fa < kNgDeltaB * S.fmin * S.fmin);
  double fn, cap, fd[dmd::kNgHist];
  if (take_a) {
    nx.upd |= dmd::kUpdAcceptY | dmd::kUpdR0FromRv;
    fn = fa;
    cap = ng_bound(d[3], d[4]);
    for (int t = 0; t < S.nh; t++) fd[t] = d[5 + t];
    S.energy = d[dmd::kDotEnergy];
  } else {
    nx.upd |= dmd::kUpdNgAcceptM;
    fn = S.fm;
    cap = S.cap_m;
    for (int t = 0; t < S.nh; t++) fd[t] = S.fmf[t];
    S.energy = S.em;
  }
  S.rnorm = fn;
  S.newton++;
  S.ng_fail = fn > kNgGammaC * S.fmin ? S.ng_fail + 1 : 0;
  S.fmin = std::min(S.fmin, fn);
  const int m = S.o.ngmres_m;
  if (S.ng_fail > kNgRestart) {  // restart: the history keeps the new iterate only
    S.nh = 0;
    S.ng_next = 0;
    S.ng_fail = 0;
  }
  const int slot = S.ng_next;
  S.ng_next = (S.ng_next + 1) % m;
  for (int t = 0; t < S.nh; t++)
    if (t != slot) S.ff[slot][t] = S.ff[t][slot] = fd[t];
  S.ff[slot][slot] = fn * fn;
  S.nh = std::min(S.nh + 1, m);
  nx.upd |= dmd::kUpdNgPush;
  nx.slot = slot;
  if (fn <= S.o.atol) {
    finish(S, DMDGPU_RELAX_CONVERGED);
  } else if (S.newton >= S.o.max_newton) {
    finish(S, DMDGPU_RELAX_MAXITER);
  } else {
    S.cap = cap;
    S.lam = std::min(2.0 * S.lam, cap);
    ng_ls_begin(S, fn * fn);
  }
}

// Consumes the tick's dots: advances each instance's state machine and queues
// the vector updates for the start of the next tick.
void digest_tick(dmdgpu_ctx* c) {
  const int ni = (int)c->inst.size();
  for (int k = 0; k < ni; k++) {
    Instance& I = c->inst[k];
    if (!I.used) continue;
    Solve& S = I.sv;
    const double* d = &c->h_sdots[(size_t)dmd::kSolveDots * k];
    dmd::SolveOp& nx = S.pending;
    if (d[dmd::kDotDomain] > 0.0 && S.stage != ST_DONE && S.stage != ST_IDLE) {
      // a density of this tick's evaluation left the embedding table's domain
      finish(S, DMDGPU_RELAX_DOMAIN);
      continue;
    }
    switch (S.stage) {
      case ST_NG_LS:
      case ST_NG_LSF:
      case ST_NG_A:
        ng_digest(S, nx, d);
        break;
      case ST_START: {
        if (S.o.solver) {
          ng_start(S, nx, d);
          break;
        }
        S.rnorm = std::sqrt(d[0]);
        S.energy = d[dmd::kDotEnergy];
        if (S.newton == 0) S.rnorm0 = S.rnorm;
        nx.upd |= dmd::kUpdR0FromRv;
        if (S.early) {
          // split residual at X: base of the probes as well
          nx.upd |= dmd::kUpdR0fFromRv;
          if (S.rnorm <= kSplitSwitch * std::sqrt((double)S.ndof)) {
            S.early = false;
            S.have_r0f = true;  // FP64 residual at X next, then CG
            break;
          }
          begin_cg(S, nx, S.rnorm * S.rnorm);
          break;
        }
        if (S.rnorm <= S.o.atol) {
          finish(S, DMDGPU_RELAX_CONVERGED);
          break;
        }
        if (S.o.mixed && S.have_r0f) {
          S.have_r0f = false;
          begin_cg(S, nx, S.rnorm * S.rnorm);
          break;
        }
        if (S.o.mixed) {
          S.stage = ST_BASE32;
          break;
        }
        begin_cg(S, nx, S.rnorm * S.rnorm);
        break;
      }
      case ST_BASE32: {
        nx.upd |= dmd::kUpdR0fFromRv;
        begin_cg(S, nx, S.rnorm * S.rnorm);
        break;
      }
      case ST_PC: {
        // d[7] = Rr . M^-1 Rr, d[13] = |P|^2 with P = M^-1 Rr
        S.rz = d[7];
        S.pp = d[13];
        if (!(S.rz > 0.0) || !std::isfinite(S.rz)) {
          finish(S, DMDGPU_RELAX_STALLED);
          break;
        }
        S.stage = ST_CG;
        break;
      }
      case ST_CG: {
        const double pAp = d[0], rAp = d[1], ApAp = d[2];
        if (!(pAp > 0.0)) {
          // negative curvature: keep the current step, or a short steepest step
          if (S.cg == 0) {
            nx.upd |= dmd::kUpdSdDir;
            nx.c_sd = kMaxStep / std::sqrt(std::max(S.pp, 1e-300) / std::max(S.ndof, 1));
          }
          S.stage = ST_LSPREP;
          break;
        }
        double a, b, rr1;
        bool stop = false;
        if (S.o.precond) {
          // r1 = r - a Ap, z1 = z - a y (y = M^-1 Ap), P1 = z1 + b P
          const double zAp = d[5], ApMAp = d[6], zz = d[8], zy = d[9], yy = d[10], zP = d[11],
                       yP = d[12], PP = d[13];
          a = S.rz / pAp;
          rr1 = std::max(S.rr - 2.0 * a * rAp + a * a * ApAp, 0.0);
          const double rz1 = std::max(S.rz - 2.0 * a * zAp + a * a * ApMAp, 0.0);
          b = rz1 / S.rz;
          const double z1z1 = std::max(zz - 2.0 * a * zy + a * a * yy, 0.0);
          const double z1P = zP - a * yP;
          S.pp = std::max(z1z1 + 2.0 * b * z1P + b * b * PP, 0.0);
          S.rz = rz1;
          stop = !(rz1 > 0.0);
        } else {
          a = S.rr / pAp;
          rr1 = std::max(S.rr - 2.0 * a * rAp + a * a * ApAp, 0.0);
          b = rr1 / S.rr;
          S.pp = rr1 + b * b * S.pp;
        }
        nx.upd |= dmd::kUpdCgStep;
        nx.eps_prev = S.eps;
        nx.alpha = a;
        nx.beta = b;
        S.rr = rr1;
        S.cg++;
        if (std::sqrt(rr1) <= S.o.eta * S.rnorm || S.cg >= S.o.max_cg || !(rr1 > 0.0) || stop)
          S.stage = ST_LSPREP;
        break;
      }
      case ST_LSPREP: {
        const double dq = d[3], ds = d[4];
        S.slope = d[0];
        double a = 1.0;
        if (dq > kMaxStep) a = std::min(a, kMaxStep / dq);
        if (ds > 0.5) a = std::min(a, 0.5 / ds);
        S.alpha_ls = a;
        S.ls = 0;
        S.stage = ST_LS;
        break;
      }
      case ST_LS: {
        const double rn = std::sqrt(d[0]);
        const double en = d[dmd::kDotEnergy];
        const bool rdec = std::isfinite(rn) && rn <= (1.0 - 1e-4 * S.alpha_ls) * S.rnorm;
        // energy merit: Armijo on sum S0; once the change is at the rounding
        // level of the sum, a step that reduces |R| is taken
        const bool edec =
            std::isfinite(en) && std::isfinite(rn) &&
            (en <= S.energy + 1e-4 * S.alpha_ls * std::min(S.slope, 0.0) ||
             (std::fabs(en - S.energy) <= kEnergyRound * std::fabs(S.energy) && rdec));
        if (S.o.merit ? edec : rdec) {
          nx.upd |= dmd::kUpdAcceptY | dmd::kUpdR0FromRv;
          S.rnorm = rn;
          S.energy = en;
          S.newton++;
          if (S.early) {
            nx.upd |= dmd::kUpdR0fFromRv;
            if (S.newton >= S.o.max_newton) {
              finish(S, DMDGPU_RELAX_MAXITER);
              break;
            }
            if (rn <= kSplitSwitch * std::sqrt((double)S.ndof)) {
              S.early = false;
              S.have_r0f = true;
              S.stage = ST_START;  // FP64 residual at the accepted iterate
              break;
            }
            begin_cg(S, nx, rn * rn);
            break;
          }
          if (rn <= S.o.atol) {
            finish(S, DMDGPU_RELAX_CONVERGED);
            break;
          }
          if (S.newton >= S.o.max_newton) {
            finish(S, DMDGPU_RELAX_MAXITER);
            break;
          }
          if (S.o.mixed) {
            S.stage = ST_BASE32;
            break;
          }
          begin_cg(S, nx, rn * rn);
        } else {
          S.alpha_ls *= 0.5;
          if (++S.ls > 12) finish(S, DMDGPU_RELAX_STALLED);
        }
        break;
      }
      case ST_DIFF: {
        // d[0] = H content of the new iterate, d[1] = of the current one,
        // d[3] = max |dx|, d[4] > 0 when an occupancy left [0, 1]
        if (S.newton == 0) S.mass0 = d[1];
        S.mass = d[0];
        S.rnorm = d[3];
        S.newton++;
        if (!(d[4] <= 0.0) || !std::isfinite(d[3])) finish(S, DMDGPU_RELAX_STALLED);
        else if (d[3] <= S.dop.tol) finish(S, DMDGPU_RELAX_CONVERGED);
        else if (S.newton >= S.dop.max_iter) finish(S, DMDGPU_RELAX_MAXITER);
        else if (S.dop.mixed && !S.diff64 && d[3] <= kDiffSwitch) S.diff64 = true;
        break;
      }
      default:
        break;
    }
    if (S.stage != ST_DONE && S.stage != ST_IDLE && S.evals >= S.o.max_evals)
      finish(S, DMDGPU_RELAX_MAXITER);
  }
}

// The device work of one tick: vector updates + evaluation points, residual
// of every instance that needs one (the precisions of mask: bit 0 FP64, bit 1
// split), totals, dots.
dmdgpu_status tick_device(dmdgpu_ctx* c, int mask) {
  dmd::State s = make_state(c);
  s.noxr = c->d_noxr.p;  // relaxations need no occupancy residual
  const dmd::Outputs o = make_outputs(c);
  const int ni = (int)c->inst.size();
  HIPCHK(dmd::launch_solver_step(c->stream, c->extent, c->d_inst.p, c->d_spc.p, c->d_ne.p,
                                 c->d_ops.p, make_vecs(c), c->d_qs.p, o));
  if (mask == 3) {
    CHK(eval_chains(c, s, o, true));
  } else {
    for (int p = 0; p < 2; p++)
      if (mask & (1 << p)) {
        HIPCHK(dmd::launch_phase1(c->stream, c->cfg, p, s, c->T, o));
        HIPCHK(dmd::launch_phase2(c->stream, c->cfg, p, s, c->T, o));
      }
  }
  HIPCHK(dmd::launch_totals(c->stream, s, c->T, o, ni, c->d_chunk_lo.p, c->d_chunk_hi.p));
  HIPCHK(dmd::launch_diff_flux(c->stream, s, c->d_ops.p, make_vecs(c), o, c->d_dnn.p,
                               c->d_dnbr.p));
  HIPCHK(dmd::launch_solver_dots(c->stream, c->extent, ni, c->d_inst.p, c->d_spc.p, c->d_ne.p,
                                 c->d_ops.p, c->d_prec.p, make_vecs(c), o, c->d_xr.p, c->d_qs.p,
                                 c->d_spart.p,
                                 c->d_chunk_lo.p, c->d_chunk_hi.p, c->d_sdots.p));
  HIPCHK(dmd::launch_diff_apply(c->stream, c->extent, c->d_inst.p, c->d_spc.p, c->d_ops.p,
                                make_vecs(c), c->d_xr.p, c->d_ne.p));
  return DMDGPU_OK;
}

// Re-reads the effective occupancies of an instance after its occupancies
// changed on the device: active-site flags and active pair count.
dmdgpu_status refresh_activity(dmdgpu_ctx* c, Instance& I) {
  std::vector<double> ne(I.n);
  HIPCHK(hipMemcpyAsync(ne.data(), c->d_ne.p + I.base, sizeof(double) * I.n,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int na = 0;
  for (int i = 0; i < I.n; i++) {
    I.h_act[i] = ne[i] != 0.0;
    na += I.h_act[i];
  }
  I.nactive = na;
  CHK(upload_meta(c));
  HIPCHK(dmd::launch_active_pairs(c->stream, make_state(c), c->d_pairs.p));
  const int c0 = I.base / kChunk, nch = I.cap / kChunk;
  std::vector<long long> pp(2 * (size_t)nch);
  HIPCHK(hipMemcpyAsync(pp.data(), c->d_pairs.p + 2 * c0, sizeof(long long) * 2 * nch,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  I.pairs = I.pairs_mg = 0;
  for (int k = 0; k < nch; k++) {
    I.pairs += pp[2 * k];
    I.pairs_mg += pp[2 * k + 1];
  }
  return DMDGPU_OK;
}

// After a diffusion step: the activity (x >= 5e-4), frozen during the step,
// follows the new occupancies again.
dmdgpu_status end_diffusion(dmdgpu_ctx* c, Instance& I) {
  std::vector<double> x(I.n), ne(I.n);
  HIPCHK(hipMemcpyAsync(x.data(), c->d_xr.p + I.base, sizeof(double) * I.n,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < I.n; i++) ne[i] = (I.h_slot[i] == 0 || x[i] >= kMinOcc) ? x[i] : 0.0;
  HIPCHK(copy_sync(c, c->d_ne.p + I.base, ne.data(), sizeof(double) * I.n, hipMemcpyHostToDevice));
  return refresh_activity(c, I);
}

// Block-Jacobi preconditioners of the instances about to start CG, outside
// the captured tick: FP64 phase 1 at their evaluation point (their iterate X:
// the last tick evaluated them there), the exact diagonal Hessian blocks, and
// their inverses into d_minv.
dmdgpu_status pc_setup(dmdgpu_ctx* c) {
  const int ni = (int)c->inst.size();
  bool any = false;
  for (int k = 0; k < ni; k++) any |= c->inst[k].used && c->inst[k].sv.pc_setup;
  if (!any) return DMDGPU_OK;
  std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
  for (int k = 0; k < ni; k++)
    if (c->inst[k].used && c->inst[k].sv.pc_setup) c->h_prec[k] = dmd::kFP64;
  CHK(run_eval(c, false));
  const dmd::State s = make_state(c);
  for (int k = 0; k < ni; k++) {
    Instance& I = c->inst[k];
    if (!I.used || !I.sv.pc_setup) continue;
    double* D = c->d_diag.p + 9 * (size_t)I.base;
    // blocks of the species with relaxed unknowns only: the others are not read
    const int dof = I.sv.o.dof_mask;
    const int mask = ((dof & (DMDGPU_DOF_MG_Q | DMDGPU_DOF_MG_SIGMA)) ? 1 : 0) |
                     ((dof & (DMDGPU_DOF_H_Q | DMDGPU_DOF_H_SIGMA)) ? 2 : 0);
    HIPCHK(dmd::launch_diag(c->stream, c->cfg.rule, 1, mask, s, c->T, c->d_rec.p, I.base, I.n,
                            D));
    HIPCHK(dmd::launch_pc_invert(c->stream, I.base, I.n, I.cap, I.sv.o.dof_mask, c->d_spc.p,
                                 c->d_ne.p, c->d_qs.p, D, c->d_minv.p));
    I.sv.pc_setup = false;
  }
  return DMDGPU_OK;
}

dmdgpu_status run_tick(dmdgpu_ctx* c) {
  CHK(upload_meta(c));
  plan_tick(c);
  const int ni = (int)c->inst.size();
  build_act(c);
  c->h_noxr.assign(ni, 0);
  for (int k = 0; k < ni; k++) {
    Instance& I = c->inst[k];
    if (!I.used || c->h_prec[k] < 0) continue;
    const bool h_dof = (I.sv.o.dof_mask & (DMDGPU_DOF_H_Q | DMDGPU_DOF_H_SIGMA)) != 0;
    int skip = 0;  // a diffusion sweep needs the whole evaluation (R_x)
    if (I.sv.stage != ST_DIFF)
      skip = dmd::kSkipRx | (h_dof ? 0 : dmd::kSkipHH) | (I.sv.o.merit ? 0 : dmd::kSkipValues);
    c->h_noxr[k] = skip;
    I.rx_stale = skip != 0;
    I.hq_stale = (skip & dmd::kSkipHH) != 0;
    I.e_stale = (skip & dmd::kSkipValues) != 0;
    if (I.e_stale) I.sv.energy_lazy = true;
    I.rx_prec = c->h_prec[k];
  }
  const size_t ob = sizeof(dmd::SolveOp) * ni, pb = sizeof(int) * ni;
  const size_t ab = sizeof(int) * c->h_chunks.size();
  HIPCHK(c->pin_up.ensure(ob + 2 * pb + ab));
  HIPCHK(c->pin_down.ensure(sizeof(double) * dmd::kSolveDots * ni));
  std::memcpy(c->pin_up.p, c->h_ops.data(), ob);
  std::memcpy((char*)c->pin_up.p + ob, c->h_prec.data(), pb);
  std::memcpy((char*)c->pin_up.p + ob + pb, c->h_noxr.data(), pb);
  std::memcpy((char*)c->pin_up.p + ob + 2 * pb, c->h_chunks.data(), ab);
  HIPCHK(hipMemcpyAsync(c->d_ops.p, c->pin_up.p, ob, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_prec.p, (char*)c->pin_up.p + ob, pb, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(c->d_noxr.p, (char*)c->pin_up.p + ob + pb, pb, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(c->d_act.p, (char*)c->pin_up.p + ob + 2 * pb, ab,
                        hipMemcpyHostToDevice, c->stream));
  int mask = 0;
  for (int k = 0; k < ni; k++)
    if (c->h_prec[k] >= 0) mask |= 1 << c->h_prec[k];
  const bool same_cfg = c->graph_cfg == c->cfg;
  if (c->use_graph) {
    if (c->graph_extent != c->extent || c->graph_ninst != ni || !same_cfg) drop_graph(c);
    if (!c->gexec[mask]) {
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      dmdgpu_status st = tick_device(c, mask);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (st || e != hipSuccess) {  // a failed capture leaves no graph behind
        if (g) (void)hipGraphDestroy(g);
        if (st) return st;
        return fail(DMDGPU_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      }
      c->graph[mask] = g;
      HIPCHK(hipGraphInstantiate(&c->gexec[mask], c->graph[mask], nullptr, nullptr, 0));
      c->graph_extent = c->extent;
      c->graph_ninst = ni;
      c->graph_cfg = c->cfg;
    }
    HIPCHK(hipGraphLaunch(c->gexec[mask], c->stream));
  } else {
    CHK(tick_device(c, mask));
  }
  const size_t nd = (size_t)dmd::kSolveDots * ni;
  c->h_sdots.resize(nd);
  HIPCHK(hipMemcpyAsync(c->pin_down.p, c->d_sdots.p, sizeof(double) * nd, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::memcpy(c->h_sdots.data(), c->pin_down.p, sizeof(double) * nd);
  std::vector<int> before(ni);
  for (int k = 0; k < ni; k++) before[k] = c->inst[k].sv.stage;
  digest_tick(c);
  c->ticks++;
  // Instances that finished this tick: apply their last updates (accepting the
  // final step) now, and leave their evaluation point at the final iterate.
  bool flush = false;
  for (int k = 0; k < ni; k++) {
    Instance& I = c->inst[k];
    dmd::SolveOp op{};
    if (I.used && before[k] == ST_DIFF && I.sv.stage == ST_DONE) {
      CHK(end_diffusion(c, I));
    } else if (I.used && before[k] != ST_DONE && I.sv.stage == ST_DONE) {
      op = I.sv.pending;
      op.make = dmd::kMakeX;
      op.mixed = I.sv.o.mixed;
      op.dof = I.sv.o.dof_mask;
      I.sv.pending = dmd::SolveOp{};
      flush = true;
    }
    c->h_ops[k] = op;
  }
  if (flush) {
    std::memcpy(c->pin_up.p, c->h_ops.data(), ob);
    HIPCHK(hipMemcpyAsync(c->d_ops.p, c->pin_up.p, ob, hipMemcpyHostToDevice, c->stream));
    HIPCHK(dmd::launch_solver_step(c->stream, c->extent, c->d_inst.p, c->d_spc.p, c->d_ne.p,
                                   c->d_ops.p, make_vecs(c), c->d_qs.p, make_outputs(c)));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return pc_setup(c);
}

int running_count(const dmdgpu_ctx* c) {
  int run = 0;
  for (const Instance& I : c->inst)
    if (I.used && I.sv.stage != ST_IDLE && I.sv.stage != ST_DONE) run++;
  return run;
}

}  // namespace

extern "C" {

const char* dmdgpu_last_error(void) { return g_err.c_str(); }

dmdgpu_status dmdgpu_create(int device, dmdgpu_ctx** out) {
  if (!out) return fail(DMDGPU_ERR_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(DMDGPU_ERR_ARG, "no such HIP device");
  HIPCHK(hipSetDevice(device));
  // device address of this device's scratch array (allocation-free entry point)
  HIPCHK(dmd::noalloc_init(device));
  dmdgpu_ctx* c = new dmdgpu_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) {
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(DMDGPU_ERR_HIP, std::string("stream/event creation: ") + hipGetErrorString(e));
  }
  *out = c;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_destroy(dmdgpu_ctx* c) {
  if (!c) return DMDGPU_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  drop_graph(c);
  for_site_bufs(c, 0, [](auto& buf, size_t, bool) { buf.release(); });
  for_other_bufs(c, [](auto& buf) { buf.release(); });
  (void)hipStreamSynchronize(c->side);
  c->pin_up.release();
  c->pin_down.release();
  c->pin_pk.release();
  (void)hipEventDestroy(c->ev_fork);
  (void)hipEventDestroy(c->ev_join);
  (void)hipStreamDestroy(c->side);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return DMDGPU_OK;
}

void* dmdgpu_stream(dmdgpu_ctx* c) { return c ? (void*)c->stream : nullptr; }

dmdgpu_status dmdgpu_sync(dmdgpu_ctx* c) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_init_adp_MgHx(dmdgpu_ctx* c, int comb, const char* folder) {
  if (!c || !folder) return fail(DMDGPU_ERR_ARG, "null argument");
  const char* names[3] = {"adpmg.dat", "adph.dat", "adpmgh.dat"};
  if (comb < 0 || comb > 2) return fail(DMDGPU_ERR_ARG, "species_comb must be MgMg, HH or MgH");
  std::string path = std::string(folder) + "/" + names[comb];
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return fail(DMDGPU_ERR_IO, "cannot open " + path);
  bool ok = true;
  if (comb < 2) {
    double factor, rcut;
    ok = fscanf(f, "%lf %lf %lf", &c->mass[comb], &factor, &rcut) == 3;
    ok = ok && read_block(f, c->rho[comb]) && read_block(f, c->embed[comb]);
    ok = ok && read_block(f, c->pair[comb]) && read_block(f, c->u[comb]) &&
         read_block(f, c->w[comb]);
  } else {
    double rcut;
    ok = fscanf(f, "%lf", &rcut) == 1;
    ok = ok && read_block(f, c->pair[2]) && read_block(f, c->u[2]) && read_block(f, c->w[2]);
  }
  fclose(f);
  if (!ok) return fail(DMDGPU_ERR_IO, "malformed table file " + path);
  c->have[comb] = true;
  c->tables_dirty = true;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_cutoff(dmdgpu_ctx* c, double rc) {
  if (!c || !(rc > 0)) return fail(DMDGPU_ERR_ARG, "cut-off must be positive");
  c->rc = rc;
  for (Instance& I : c->inst) I.have_neigh = false;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_tuning(dmdgpu_ctx* c, int k1, int k2, int unroll2) {
  auto ok = [](int k) { return k == 4 || k == 8 || k == 16; };
  if (!c || !ok(k1) || !(ok(k2) || k2 == 32 || k2 == 64))
    return fail(DMDGPU_ERR_ARG, "k_phase1 must be 4, 8 or 16 and k_phase2 4, 8, 16, 32 or 64");
  for (int p = 0; p < 2; p++) {
    c->cfg.k1[p] = k1;
    c->cfg.k2[p] = k2;
    c->cfg.unroll2[p] = unroll2 ? 1 : 0;
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_tuning_precision(dmdgpu_ctx* c, int precision, int k1, int k2,
                                          int unroll2) {
  auto ok = [](int k) { return k == 4 || k == 8 || k == 16; };
  if (!c || !ok(k1) || !(ok(k2) || k2 == 32 || k2 == 64) ||
      (precision != DMDGPU_FP64 && precision != DMDGPU_SPLIT32))
    return fail(DMDGPU_ERR_ARG, "bad tuning");
  const int p = precision == DMDGPU_SPLIT32 ? dmd::kSplit32 : dmd::kFP64;
  c->cfg.k1[p] = k1;
  c->cfg.k2[p] = k2;
  c->cfg.unroll2[p] = unroll2 ? 1 : 0;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_precision(dmdgpu_ctx* c, int precision) {
  if (!c || (precision != DMDGPU_FP64 && precision != DMDGPU_SPLIT32))
    return fail(DMDGPU_ERR_ARG, "precision must be DMDGPU_FP64 or DMDGPU_SPLIT32");
  c->cfg.precision = precision == DMDGPU_SPLIT32 ? dmd::kSplit32 : dmd::kFP64;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_rule(dmdgpu_ctx* c, int order) {
  if (!c || (order != 3 && order != 5 && order != 44 && order != 6))
    return fail(DMDGPU_ERR_ARG, "rule must be 3, 5, 44 or 6");
  c->cfg.rule = order;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_state(dmdgpu_ctx* c, int n, const double* q, const double* sg,
                               const double* x, const int32_t* spc, const double* beta,
                               const double* gamma, const double* box, const int32_t* periodic) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  for (int k = 0; k < (int)c->inst.size(); k++)
    if (c->inst[k].used) CHK(remove_instance(c, k));
  int id = -1;
  CHK(add_instance(c, n, q, sg, x, spc, beta, gamma, box, periodic, &id));
  c->primary = id;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_update_state(dmdgpu_ctx* c, const double* q, const double* sg,
                                  const double* x) {
  if (!c || c->primary < 0) return fail(DMDGPU_ERR_STATE, "dmdgpu_set_state has not been called");
  HIPCHK(hipSetDevice(c->device));
  return upload_unknowns(c, c->inst[c->primary], q, sg, x, false);
}

dmdgpu_status dmdgpu_neighbors(dmdgpu_ctx* c) {
  if (!c || c->primary < 0) return fail(DMDGPU_ERR_STATE, "dmdgpu_set_state has not been called");
  HIPCHK(hipSetDevice(c->device));
  return build_neighbors(c, c->inst[c->primary]);
}

dmdgpu_status dmdgpu_get_neighbors(dmdgpu_ctx* c, int32_t* h_nn, int32_t* h_list, int cap) {
  CHK(check_primary(c));
  if (cap < dmd::kMaxNeigh) return fail(DMDGPU_ERR_ARG, "cap must be >= 250");
  const Instance& I = c->inst[c->primary];
  const int n = I.n;
  std::vector<int> nn(n), rows((size_t)n * dmd::kCap);
  HIPCHK(hipMemcpyAsync(nn.data(), c->d_nnb.p + I.base, sizeof(int) * n, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipMemcpyAsync(rows.data(), c->d_nbr.p + (size_t)I.base * dmd::kCap,
                        sizeof(int) * rows.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<int> row;
  for (int o = 0; o < n; o++) {
    const int i = I.inv[o];
    if (h_nn) h_nn[o] = nn[i];
    if (!h_list) continue;
    row.assign(rows.begin() + (size_t)i * dmd::kCap, rows.begin() + (size_t)i * dmd::kCap + nn[i]);
    for (int& v : row) v = I.perm[v - I.base];
    std::sort(row.begin(), row.end());
    for (int k = 0; k < cap; k++) h_list[(size_t)o * cap + k] = k < nn[i] ? row[k] : -1;
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_mf_rho(dmdgpu_ctx* c, double* h_rho) {
  CHK(check_primary(c));
  HIPCHK(hipSetDevice(c->device));
  CHK(eval_one(c, c->primary, false));
  CHK(d2h_sites(c, c->inst[c->primary], h_rho, c->d_rho.p, 1));
  return check_domain(c, c->primary);
}

dmdgpu_status dmdgpu_site_energies(dmdgpu_ctx* c, double* h_V0, double* h_S0) {
  CHK(check_primary(c));
  HIPCHK(hipSetDevice(c->device));
  CHK(eval_one(c, c->primary));
  CHK(d2h_sites(c, c->inst[c->primary], h_V0, c->d_V0.p, 1));
  CHK(d2h_sites(c, c->inst[c->primary], h_S0, c->d_S0.p, 1));
  return check_domain(c, c->primary);
}

dmdgpu_status dmdgpu_residual(dmdgpu_ctx* c, double* h_Rq, double* h_Rs, double* h_Rx,
                              double* h_tot) {
  CHK(check_primary(c));
  HIPCHK(hipSetDevice(c->device));
  CHK(eval_one(c, c->primary));
  return copy_residual(c, c->primary, h_Rq, h_Rs, h_Rx, h_tot);
}

// The packed residual of the primary instance on device buffers: dX (4n,
// caller order) sets q and sigma first (nullptr: the current state); dR (5n,
// caller order) receives R (nullptr: none), and pin (nullptr: none) a copy of
// it.  Returns after the stream has finished.
dmdgpu_status packed_eval(dmdgpu_ctx* c, const double* dX, double* dR, double* pin,
                          double* h_totals) {
  const int id = c->primary;
  Instance& I = c->inst[id];
  const size_t n = (size_t)I.n;
  if (dX) {
    HIPCHK(hipMemsetAsync(c->d_flags.p, 0, sizeof(int), c->stream));
    HIPCHK(dmd::launch_gather_unknowns(c->stream, I.base, I.n, c->d_perm.p, dX, c->d_qs.p, c->d_X.p,
                                       c->d_flags.p));
    I.h_q_stale = true;
  }
  CHK(eval_one(c, id));
  if (dR) {
    HIPCHK(dmd::launch_scatter_residual(c->stream, I.base, I.n, c->d_perm.p, c->d_Rq.p, c->d_Rs.p,
                                        c->d_Rx.p, dR));
    if (pin)
      HIPCHK(hipMemcpyAsync(pin, dR, sizeof(double) * 5 * n, hipMemcpyDeviceToHost, c->stream));
  }
  double t3[3];
  int bad = 0;
  HIPCHK(hipMemcpyAsync(t3, c->d_totals.p + 3 * id, sizeof(t3), hipMemcpyDeviceToHost, c->stream));
  if (dX) HIPCHK(hipMemcpyAsync(&bad, c->d_flags.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (h_totals) {
    h_totals[0] = t3[0];
    h_totals[1] = t3[1];
  }
  if (bad) return fail(DMDGPU_ERR_ARG, "positions must be finite and widths positive and finite");
  return domain_status(t3[2]);
}

dmdgpu_status dmdgpu_residual_packed(dmdgpu_ctx* c, const double* X, double* R, double* h_totals,
                                     int on_device) {
  CHK(check_primary(c));
  if (on_device != 0 && on_device != 1) return fail(DMDGPU_ERR_ARG, "on_device must be 0 or 1");
  HIPCHK(hipSetDevice(c->device));
  const size_t n = (size_t)c->inst[c->primary].n;
  if (on_device) return packed_eval(c, X, R, nullptr, h_totals);
  HIPCHK(c->pin_pk.ensure(sizeof(double) * 5 * n));
  double* pin = static_cast<double*>(c->pin_pk.p);
  if (X) {
    std::memcpy(pin, X, sizeof(double) * 4 * n);
    HIPCHK(hipMemcpyAsync(c->d_pkin.p, pin, sizeof(double) * 4 * n, hipMemcpyHostToDevice,
                          c->stream));
  }
  const dmdgpu_status st =
      packed_eval(c, X ? c->d_pkin.p : nullptr, R ? c->d_pkout.p : nullptr, R ? pin : nullptr,
                  h_totals);
  if (R && (st == DMDGPU_OK || st == DMDGPU_ERR_DOMAIN))
    std::memcpy(R, pin, sizeof(double) * 5 * n);
  return st;
}

// J V of the (q, sigma) residual at X by central differences, on device
// buffers (caller order, 4n each): |V|^2 on the device, R(X + eps V) and
// R(X - eps V) through the packed residual, JV = their difference / (2 eps);
// the state is left at X.
dmdgpu_status jvp_device(dmdgpu_ctx* c, const double* dX, const double* dV, double* dJ) {
  Instance& I = c->inst[c->primary];
  const size_t n = (size_t)I.n, m = 4 * n;
  HIPCHK(c->d_jscal.ensure(1));
  HIPCHK(dmd::launch_sumsq(c->stream, dV, m, c->d_jscal.p));
  double vv = 0.0;
  HIPCHK(hipMemcpyAsync(&vv, c->d_jscal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!std::isfinite(vv)) return fail(DMDGPU_ERR_ARG, "direction must be finite");
  if (!(vv > 0.0)) {
    HIPCHK(hipMemsetAsync(dJ, 0, sizeof(double) * m, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DMDGPU_OK;
  }
  // RMS displacement of 1e-5 A: truncation ~1e-10 |d3 S| and rounding ~1e-8
  // relative
  const double eps = 1e-5 * std::sqrt((double)m / vv);
  dmdgpu_status st = DMDGPU_OK;
  for (int side = 0; side < 2 && st == DMDGPU_OK; side++) {
    HIPCHK(dmd::launch_axpy(c->stream, dX, dV, side == 0 ? eps : -eps, m, c->d_pkin.p));
    st = packed_eval(c, c->d_pkin.p, side == 0 ? c->d_pkout.p : c->d_jr.p, nullptr, nullptr);
  }
  if (st != DMDGPU_OK) return st;
  HIPCHK(dmd::launch_diff_scale(c->stream, c->d_pkout.p, c->d_jr.p, 0.5 / eps, m, dJ));
  // back to X (no evaluation)
  HIPCHK(hipMemsetAsync(c->d_flags.p, 0, sizeof(int), c->stream));
  HIPCHK(dmd::launch_gather_unknowns(c->stream, I.base, I.n, c->d_perm.p, dX, c->d_qs.p, c->d_X.p,
                                     c->d_flags.p));
  I.h_q_stale = true;
  HIPCHK(hipStreamSynchronize(c->stream));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_jvp_packed(dmdgpu_ctx* c, const double* X, const double* V, double* JV,
                                int on_device) {
  CHK(check_primary(c));
  if (!X || !V || !JV) return fail(DMDGPU_ERR_ARG, "null argument");
  if (on_device != 0 && on_device != 1) return fail(DMDGPU_ERR_ARG, "on_device must be 0 or 1");
  HIPCHK(hipSetDevice(c->device));
  if (on_device) return jvp_device(c, X, V, JV);
  const size_t m = 4 * (size_t)c->inst[c->primary].n;
  HIPCHK(c->pin_pk.ensure(sizeof(double) * 2 * m));
  double* pin = static_cast<double*>(c->pin_pk.p);
  std::memcpy(pin, X, sizeof(double) * m);
  std::memcpy(pin + m, V, sizeof(double) * m);
  HIPCHK(hipMemcpyAsync(c->d_jx.p, pin, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_jv.p, pin + m, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
  CHK(jvp_device(c, c->d_jx.p, c->d_jv.p, c->d_jv.p));
  HIPCHK(hipMemcpyAsync(pin, c->d_jv.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::memcpy(JV, pin, sizeof(double) * m);
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_jvp(dmdgpu_ctx* c, const double* X, const double* V, double* JV) {
  return dmdgpu_jvp_packed(c, X, V, JV, 0);
}

dmdgpu_status diag_blocks_of(dmdgpu_ctx* c, int id, int exact, double* h_D) {
  if (!h_D) return fail(DMDGPU_ERR_ARG, "null output");
  if (exact != 0 && exact != 1) return fail(DMDGPU_ERR_ARG, "exact must be 0 or 1");
  HIPCHK(hipSetDevice(c->device));
  // phase 1 in FP64 at the instance's state (the blocks read the FP64 adjoint records)
  std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
  c->h_prec[id] = dmd::kFP64;
  CHK(run_eval(c, false));
  const Instance& I = c->inst[id];
  hipEvent_t a{};
  CHK(begin_timed(c, &a));
  HIPCHK(dmd::launch_diag(c->stream, c->cfg.rule, exact, 3, make_state(c), c->T, c->d_rec.p, I.base,
                          I.n, c->d_diag.p));
  CHK(end_timed(c, KID_DIAG, a));
  std::vector<double> tmp((size_t)I.n * 9);
  HIPCHK(hipMemcpyAsync(tmp.data(), c->d_diag.p, tmp.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < I.n; i++)
    std::copy(tmp.begin() + 9 * (size_t)i, tmp.begin() + 9 * (size_t)(i + 1),
              h_D + 9 * (size_t)I.perm[i]);
  return check_domain(c, id);
}

// Per-(u, i) derivatives of V0_i for m queries in the caller's site numbering
// of the one-instance state: FP64 phase 1, then one wavefront per query.
dmdgpu_status site_hooks(dmdgpu_ctx* c, int m, const int32_t* hu, const int32_t* hi,
                         double* out4, double* outx) {
  CHK(check_primary(c));
  if (m < 0 || (m > 0 && (!hu || !hi))) return fail(DMDGPU_ERR_ARG, "bad query list");
  if (m == 0) return DMDGPU_OK;
  const int id = c->primary;
  const Instance& I = c->inst[id];
  std::vector<int> q(2 * (size_t)m);
  for (int k = 0; k < m; k++) {
    if (hu[k] < 0 || hu[k] >= I.n || hi[k] < 0 || hi[k] >= I.n)
      return fail(DMDGPU_ERR_ARG, "site index out of range in query " + std::to_string(k));
    q[k] = I.base + I.inv[hu[k]];
    q[m + k] = I.base + I.inv[hi[k]];
  }
  HIPCHK(hipSetDevice(c->device));
  std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
  c->h_prec[id] = dmd::kFP64;
  CHK(run_eval(c, false));
  HIPCHK(c->d_hookq.ensure(2 * (size_t)m));
  HIPCHK(c->d_hookv.ensure(5 * (size_t)m));
  HIPCHK(hipMemcpyAsync(c->d_hookq.p, q.data(), sizeof(int) * q.size(), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(dmd::launch_site_dv0(c->stream, c->cfg.rule, make_state(c), c->T, c->d_rec.p, m,
                              c->d_hookq.p, c->d_hookq.p + m, c->d_hookv.p));
  std::vector<double> v(5 * (size_t)m);
  HIPCHK(hipMemcpyAsync(v.data(), c->d_hookv.p, sizeof(double) * v.size(), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < m; k++) {
    if (out4)
      for (int a = 0; a < 4; a++) out4[4 * (size_t)k + a] = v[5 * (size_t)k + a];
    if (outx) outx[k] = v[5 * (size_t)k + 4];
  }
  return check_domain(c, id);
}

dmdgpu_status dmdgpu_site_DV0_Dq(dmdgpu_ctx* c, int m, const int32_t* h_u, const int32_t* h_i,
                                 double* h_out) {
  if (!h_out && m > 0) return fail(DMDGPU_ERR_ARG, "null output");
  return site_hooks(c, m, h_u, h_i, h_out, nullptr);
}

dmdgpu_status dmdgpu_site_DV0_Dxi(dmdgpu_ctx* c, int m, const int32_t* h_u, const int32_t* h_i,
                                  double* h_out) {
  if (!h_out && m > 0) return fail(DMDGPU_ERR_ARG, "null output");
  return site_hooks(c, m, h_u, h_i, nullptr, h_out);
}

dmdgpu_status dmdgpu_diag_blocks(dmdgpu_ctx* c, int exact, double* h_D) {
  CHK(check_primary(c));
  return diag_blocks_of(c, c->primary, exact, h_D);
}

dmdgpu_status dmdgpu_diag_blocks_noalloc(dmdgpu_ctx* c, int exact, double* d_D) {
  CHK(check_primary(c));
  if (!d_D) return fail(DMDGPU_ERR_ARG, "null output");
  if (exact != 0 && exact != 1) return fail(DMDGPU_ERR_ARG, "exact must be 0 or 1");
  const int id = c->primary;
  const Instance& I = c->inst[id];
  if (I.n > dmd::kScratchSites)
    return fail(DMDGPU_ERR_ARG, "dmdgpu_diag_blocks_noalloc takes at most " +
                                    std::to_string(dmd::kScratchSites) + " sites (the instance has " +
                                    std::to_string(I.n) + ")");
  HIPCHK(hipSetDevice(c->device));
  double* scr = dmd::noalloc_scratch(c->device);
  if (!scr) return fail(DMDGPU_ERR_STATE, "device scratch not set up");
  // Everything below runs on memory the context already holds: the metadata
  // and per-site arrays exist since dmdgpu_neighbors, and no timing events
  // are created.
  if (c->meta_dirty) return fail(DMDGPU_ERR_STATE, "dmdgpu_neighbors has not been called");
  if (c->tables_dirty)  // their upload may allocate
    return fail(DMDGPU_ERR_STATE, "tables loaded since the last evaluation: evaluate once with "
                                  "an allocating entry point first");
  const bool timing = c->timing;
  c->timing = false;
  std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
  c->h_prec[id] = dmd::kFP64;
  dmdgpu_status st = run_eval(c, false);
  c->timing = timing;
  CHK(st);
  HIPCHK(dmd::launch_diag(c->stream, c->cfg.rule, exact, 3, make_state(c), c->T, c->d_rec.p, I.base,
                          I.n, scr));
  HIPCHK(dmd::launch_scatter_blocks(c->stream, I.base, I.n, c->d_perm.p, scr, d_D));
  return check_domain(c, id);
}

dmdgpu_status dmdgpu_instance_diag_blocks(dmdgpu_ctx* c, int id, int exact, double* h_D) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (!I->have_neigh) return fail(DMDGPU_ERR_STATE, "instance has no neighbour lists");
  if (I->sv.stage != ST_IDLE && I->sv.stage != ST_DONE)
    return fail(DMDGPU_ERR_STATE, "instance has a solve running");
  return diag_blocks_of(c, id, exact, h_D);
}

dmdgpu_status dmdgpu_residual_resident(dmdgpu_ctx* c) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  return eval_all(c);
}

// ---- instances ----------------------------------------------------------------

dmdgpu_status dmdgpu_add_instance(dmdgpu_ctx* c, int n, const double* q, const double* sg,
                                  const double* x, const int32_t* spc, const double* beta,
                                  const double* gamma, const double* box,
                                  const int32_t* periodic, int* id) {
  if (!c || !id) return fail(DMDGPU_ERR_ARG, "null argument");
  CHK(add_instance(c, n, q, sg, x, spc, beta, gamma, box, periodic, id));
  dmdgpu_status st = build_neighbors(c, c->inst[*id]);
  if (st) {
    std::string msg = g_err;
    (void)remove_instance(c, *id);
    *id = -1;
    return fail(st, msg);
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_remove_instance(dmdgpu_ctx* c, int id) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  return remove_instance(c, id);
}

dmdgpu_status dmdgpu_instance_info(dmdgpu_ctx* c, int id, int* n, int64_t* pairs,
                                   int64_t* arena_sites) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (n) *n = I->n;
  if (pairs) *pairs = I->pairs;
  if (arena_sites) *arena_sites = c->extent;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_budget(dmdgpu_ctx* c, int64_t bytes) {
  if (!c || bytes < 0) return fail(DMDGPU_ERR_ARG, "bad budget");
  c->budget = (size_t)bytes;
  return DMDGPU_OK;
}

int64_t dmdgpu_bytes_per_site(void) {
  static dmdgpu_ctx* proto = new dmdgpu_ctx();  // types only: nothing is allocated
  return (int64_t)((site_bytes(proto, (size_t)1 << 20) + ((size_t)1 << 20) - 1) >> 20);
}

dmdgpu_status dmdgpu_device_bytes(dmdgpu_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return fail(DMDGPU_ERR_ARG, "null argument");
  *bytes = (int64_t)device_bytes(c);
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_evaluate(dmdgpu_ctx* c) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  return eval_all(c);
}

dmdgpu_status dmdgpu_instance_residual(dmdgpu_ctx* c, int id, double* Rq, double* Rs, double* Rx,
                                       double* tot) {
  if (!get_inst(c, id)) return fail(DMDGPU_ERR_ARG, "no such instance");
  return copy_residual(c, id, Rq, Rs, Rx, tot);
}

dmdgpu_status dmdgpu_instance_state(dmdgpu_ctx* c, int id, double* q, double* sg, double* x) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  const int n = I->n;
  std::vector<double4> X(n);
  HIPCHK(hipMemcpyAsync(X.data(), c->d_X.p + I->base, sizeof(double4) * n, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) {
    const int o = I->perm[i];
    if (q) {
      q[3 * o] = X[i].x;
      q[3 * o + 1] = X[i].y;
      q[3 * o + 2] = X[i].z;
    }
    if (sg) sg[o] = X[i].w;
  }
  if (x) CHK(d2h_sites(c, *I, x, c->d_xr.p, 1));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_instance_set_unknowns(dmdgpu_ctx* c, int id, const double* q,
                                           const double* sg, const double* x) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  return upload_unknowns(c, *I, q, sg, x, false);
}

// ---- relaxation ----------------------------------------------------------------

dmdgpu_status dmdgpu_relax_start(dmdgpu_ctx* c, int id, const dmdgpu_relax_opts* o) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (!I->have_neigh) return fail(DMDGPU_ERR_STATE, "instance has no neighbour lists");
  Solve S;
  S.o = o ? *o : dmdgpu_relax_opts{1e-8, 50, 60, 0.05, 0, 100000, 0, 0.0, 0, 0};
  if (!(S.o.atol > 0) || S.o.max_newton <= 0 || S.o.max_cg <= 0 || !(S.o.eta > 0) ||
      S.o.dof_mask < 0 || S.o.dof_mask > 15 || !(S.o.probe >= 0) || S.o.precond < 0 ||
      (S.o.merit != 0 && S.o.merit != 1) || (S.o.solver != 0 && S.o.solver != 1) ||
      S.o.ngmres_m < 0 || S.o.ngmres_m > dmd::kNgHist ||
      (S.o.solver == 1 && (S.o.mixed || S.o.precond)))
    return fail(DMDGPU_ERR_ARG, "bad relaxation options");
  if (S.o.max_evals <= 0) S.o.max_evals = 100000;
  if (S.o.ngmres_m == 0) S.o.ngmres_m = dmd::kNgHist;
  if (S.o.dof_mask == 0) S.o.dof_mask = DMDGPU_DOF_MG_Q | DMDGPU_DOF_MG_SIGMA;
  S.stage = ST_START;
  int nd = 0;
  for (int i = 0; i < I->n; i++) {
    if (!I->h_act[i]) continue;
    const bool mg = I->h_slot[i] == 0;
    nd += (S.o.dof_mask & (mg ? DMDGPU_DOF_MG_Q : DMDGPU_DOF_H_Q)) ? 3 : 0;
    nd += (S.o.dof_mask & (mg ? DMDGPU_DOF_MG_SIGMA : DMDGPU_DOF_H_SIGMA)) ? 1 : 0;
  }
  if (nd == 0) return fail(DMDGPU_ERR_ARG, "no degree of freedom selected");
  S.ndof = nd;
  S.early = S.o.mixed != 0;
  I->sv = S;
  I->h_q_stale = true;  // the solve moves the positions on the device
  // the iterate starts from the current evaluation state of the instance
  HIPCHK(hipMemcpyAsync(c->d_X.p + I->base, c->d_qs.p + I->base, sizeof(double4) * I->cap,
                        hipMemcpyDeviceToDevice, c->stream));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_relax_tick(dmdgpu_ctx* c, int* running) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  if (running_count(c) > 0) CHK(run_tick(c));
  if (running) *running = running_count(c);
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_relax_status(dmdgpu_ctx* c, int id, int* state, double* rnorm, int* newton,
                                  int* evals) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  const Solve& S = I->sv;
  if (state)
    *state = S.stage == ST_IDLE ? DMDGPU_RELAX_IDLE
             : S.stage == ST_DONE ? S.result
                                  : DMDGPU_RELAX_RUNNING;
  if (rnorm) *rnorm = S.rnorm;
  if (newton) *newton = S.newton;
  if (evals) *evals = S.evals;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_relax_evals(dmdgpu_ctx* c, int id, int* evals_fp64, int* evals_split) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (evals_fp64) *evals_fp64 = I->sv.evals - I->sv.evals32;
  if (evals_split) *evals_split = I->sv.evals32;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_instance_pairs(dmdgpu_ctx* c, int id, int64_t* pairs, int64_t* pairs_mg) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (pairs) *pairs = I->pairs;
  if (pairs_mg) *pairs_mg = I->pairs_mg;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_relax_energy(dmdgpu_ctx* c, int id, double* energy) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (!energy) return fail(DMDGPU_ERR_ARG, "null output");
  if (I->sv.energy_lazy) {
    // The solve's ticks formed no energies (residual-norm line search): the
    // energy of its final iterate, evaluated once in FP64.
    if (I->sv.stage != ST_DONE)
      return fail(DMDGPU_ERR_STATE, "energy of a running residual-merit solve");
    HIPCHK(hipSetDevice(c->device));
    std::fill(c->h_prec.begin(), c->h_prec.end(), -1);
    c->h_prec[id] = dmd::kFP64;
    CHK(run_eval(c, true));
    double t3[3];
    HIPCHK(hipMemcpyAsync(t3, c->d_totals.p + 3 * id, sizeof(t3), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    I->sv.energy = t3[0];
    I->sv.energy_lazy = false;
  }
  *energy = I->sv.energy;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_relax(dmdgpu_ctx* c, const dmdgpu_relax_opts* o, int max_ticks) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  for (int k = 0; k < (int)c->inst.size(); k++)
    if (c->inst[k].used && c->inst[k].have_neigh) CHK(dmdgpu_relax_start(c, k, o));
  for (int t = 0; t < max_ticks; t++) {
    int run = 0;
    CHK(dmdgpu_relax_tick(c, &run));
    if (run == 0) break;
  }
  return DMDGPU_OK;
}

// ---- diffusion step -------------------------------------------------------------

namespace {

dmdgpu_status check_diffuse_opts(const dmdgpu_diffuse_opts* o) {
  if (!o) return fail(DMDGPU_ERR_ARG, "null diffusion options");
  if (!(o->dt > 0) || !(o->D0 >= 0) || !std::isfinite(o->Eb) || !(o->tol > 0) ||
      o->max_iter <= 0)
    return fail(DMDGPU_ERR_ARG, "bad diffusion options");
  return DMDGPU_OK;
}

dmdgpu_status build_diff_neighbors(dmdgpu_ctx* c, Instance& I) {
  HIPCHK(hipMemsetAsync(c->d_flags.p, 0, 2 * sizeof(int), c->stream));
  HIPCHK(dmd::launch_diff_neighbors(c->stream, I.base, I.n, I.box, c->d_spc.p, c->d_qs.p,
                                    c->d_nnb.p, c->d_nbr.p, c->d_dnn.p, c->d_dnbr.p,
                                    c->d_flags.p));
  int flags[2];
  HIPCHK(hipMemcpyAsync(flags, c->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (flags[0]) return fail(DMDGPU_ERR_MAXNEIGH, "an H site has more than 50 diffusive neighbours");
  if (flags[1]) return fail(DMDGPU_ERR_NONEIGH, "an H site has no diffusive neighbour");
  return DMDGPU_OK;
}

}  // namespace

dmdgpu_status dmdgpu_diffuse_start(dmdgpu_ctx* c, int id, const dmdgpu_diffuse_opts* o) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (!I->have_neigh) return fail(DMDGPU_ERR_STATE, "instance has no neighbour lists");
  CHK(check_diffuse_opts(o));
  HIPCHK(hipSetDevice(c->device));
  CHK(build_diff_neighbors(c, *I));
  Solve S;
  S.dop = *o;
  S.o.max_evals = o->max_iter;  // one residual evaluation per sweep
  S.stage = ST_DIFF;
  I->sv = S;
  HIPCHK(hipMemcpyAsync(c->d_Xn.p + I->base, c->d_xr.p + I->base, sizeof(double) * I->cap,
                        hipMemcpyDeviceToDevice, c->stream));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_diffusion_rates(dmdgpu_ctx* c, int id, const dmdgpu_diffuse_opts* o,
                                     double* xdot) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (!I->have_neigh) return fail(DMDGPU_ERR_STATE, "instance has no neighbour lists");
  if (running_count(c) > 0) return fail(DMDGPU_ERR_STATE, "solves are running");
  CHK(check_diffuse_opts(o));
  HIPCHK(hipSetDevice(c->device));
  CHK(build_diff_neighbors(c, *I));
  CHK(eval_one(c, id));
  HIPCHK(hipMemcpyAsync(c->d_Xn.p + I->base, c->d_xr.p + I->base, sizeof(double) * I->cap,
                        hipMemcpyDeviceToDevice, c->stream));
  const int ni = (int)c->inst.size();
  std::vector<dmd::SolveOp> ops(ni);
  set_diff_op(ops[id], *o, dmd::kDiffRates);
  HIPCHK(copy_sync(c, c->d_ops.p, ops.data(), sizeof(dmd::SolveOp) * ni, hipMemcpyHostToDevice));
  HIPCHK(dmd::launch_diff_flux(c->stream, make_state(c), c->d_ops.p, make_vecs(c),
                               make_outputs(c), c->d_dnn.p, c->d_dnbr.p));
  return d2h_sites(c, *I, xdot, c->d_Xt.p, 1);
}

dmdgpu_status dmdgpu_diffusion_neighbors(dmdgpu_ctx* c, int id, int* counts, int* lists) {
  Instance* I = get_inst(c, id);
  if (!I) return fail(DMDGPU_ERR_ARG, "no such instance");
  if (!I->have_neigh) return fail(DMDGPU_ERR_STATE, "instance has no neighbour lists");
  if (!counts || !lists) return fail(DMDGPU_ERR_ARG, "null output");
  HIPCHK(hipSetDevice(c->device));
  CHK(build_diff_neighbors(c, *I));
  const int n = I->n;
  std::vector<int> cnt(n), lst((size_t)n * dmd::kDCap);
  HIPCHK(copy_sync(c, cnt.data(), c->d_dnn.p + I->base, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHK(copy_sync(c, lst.data(), c->d_dnbr.p + (size_t)I->base * dmd::kDCap,
                   sizeof(int) * n * dmd::kDCap, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) {
    const int o = I->perm[i];
    counts[o] = cnt[i];
    for (int k = 0; k < dmd::kDCap; k++)
      lists[(size_t)o * dmd::kDCap + k] =
          k < cnt[i] ? I->perm[lst[(size_t)i * dmd::kDCap + k] - I->base] : -1;
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_graph(dmdgpu_ctx* c, int on) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  c->use_graph = on != 0;
  if (!c->use_graph) drop_graph(c);
  return DMDGPU_OK;
}

// ---- instrumentation -------------------------------------------------------------

dmdgpu_status dmdgpu_set_timing(dmdgpu_ctx* c, int on) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  c->timing = on != 0;
  return DMDGPU_OK;
}

int dmdgpu_kernel_count(void) { return KID_COUNT; }

const char* dmdgpu_kernel_name(int id) {
  return (id >= 0 && id < KID_COUNT) ? kKernelNames[id] : "";
}

dmdgpu_status dmdgpu_kernel_times(dmdgpu_ctx* c, int64_t* h_launches, double* h_ms) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  CHK(drain_timing(c));
  for (int k = 0; k < KID_COUNT; k++) {
    if (h_launches) h_launches[k] = c->launches[k];
    if (h_ms) h_ms[k] = c->ms[k];
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_reset_kernel_times(dmdgpu_ctx* c) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  CHK(drain_timing(c));
  for (int k = 0; k < KID_COUNT; k++) {
    c->launches[k] = 0;
    c->ms[k] = 0.0;
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_work_counts(dmdgpu_ctx* c, int64_t* pairs, int64_t* pqe) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  long long p = 0;
  for (const Instance& I : c->inst)
    if (I.used) p += I.pairs;
  if (pairs) *pairs = p;
  if (pqe) *pqe = 50 * p;
  return DMDGPU_OK;
}

}  // extern "C"
