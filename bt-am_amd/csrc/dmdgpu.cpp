// dmdgpu.cpp -- host side of the C ABI declared in include/dmdgpu.h.
//
// Owns the device-resident swarm (SoA), the spline tables, the neighbour lists
// and the output buffers of one context, and sequences the kernels of
// dmd_kernels.hip on the context's stream.
#include "../../include/dmdgpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dmd_kernels.h"

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

constexpr double kMinOcc = 5e-4;            // src/Macros.hpp:871
constexpr double kDefaultRc = 6.2934034034;  // src/Macros.hpp:857

struct Spline {
  int n = 0;
  double dx = 0.0;
  std::vector<double> c;  // n x 9 columns a b c d db dc dd ddc ddd
};

enum KernelId { KID_NEIGH = 0, KID_PHASE1, KID_PHASE2, KID_TOTALS, KID_COUNT };
const char* kKernelNames[KID_COUNT] = {"neighbour_build", "phase1_site_sums", "phase2_gradient",
                                       "totals"};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t cnt) {
    if (cnt <= n && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(cnt, 1));
    if (e == hipSuccess) n = cnt;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct dmdgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double rc = kDefaultRc;
  dmd::PhaseConfig cfg{dmd::kFP64, 4, 4, 1};
  // tables
  bool have[3] = {false, false, false};
  Spline rho[2], embed[2], pair[3], u[3], w[3];
  double mass[2] = {0.0, 0.0};
  bool tables_dirty = true;
  DevBuf<double> d_uw, d_ph, d_rh, d_em;
  DevBuf<float> d_uw32, d_ph32, d_rh32;
  dmd::Tables T{};
  // state
  int n = 0;
  dmd::Box box{};
  std::vector<double> h_q;      // positions, caller's order
  std::vector<int> perm, inv;   // internal site i holds caller's site perm[i]
  std::vector<int> h_slot;      // species slot, internal order
  DevBuf<double4> d_qs;
  DevBuf<double> d_ne, d_xr, d_beta, d_gamma;
  DevBuf<int> d_spc;
  bool have_state = false;
  // neighbours
  DevBuf<int> d_nnb, d_nbr, d_cell_of, d_cell_count, d_cell_start, d_cell_cursor, d_cell_atoms,
      d_flags;
  bool have_neigh = false;
  long long pairs = 0;
  // outputs
  DevBuf<double> d_rec, d_V0, d_rho, d_Rq, d_Rs, d_Rx, d_S0, d_partials, d_totals;
  DevBuf<long long> d_pairs;
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

dmdgpu_status begin_timed(dmdgpu_ctx* c, int id, hipEvent_t* a) {
  if (!c->timing) return DMDGPU_OK;
  HIPCHK(hipEventCreate(a));
  HIPCHK(hipEventRecord(*a, c->stream));
  (void)id;
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

// ---- table files (src/ADP/MgHx-ADP.cpp:366-501 layout) ----------------------
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

dmdgpu_status upload_tables(dmdgpu_ctx* c) {
  if (!c->tables_dirty) return DMDGPU_OK;
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
  const int ne_stride = std::max(c->embed[0].n, c->embed[1].n);
  std::vector<double> em((size_t)2 * ne_stride * 4, 0.0);
  for (int s = 0; s < 2; s++) pack4(c->embed[s], &em[(size_t)s * ne_stride * 4]);
  HIPCHK(c->d_uw.ensure(uw.size()));
  HIPCHK(c->d_ph.ensure(ph.size()));
  HIPCHK(c->d_rh.ensure(rh.size()));
  HIPCHK(c->d_em.ensure(em.size()));
  HIPCHK(hipMemcpy(c->d_uw.p, uw.data(), uw.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_ph.p, ph.data(), ph.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_rh.p, rh.data(), rh.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_em.p, em.data(), em.size() * 8, hipMemcpyHostToDevice));
  // FP32 copies for the split-precision evaluation
  std::vector<float> uw32(uw.begin(), uw.end()), ph32(ph.begin(), ph.end()), rh32(rh.begin(), rh.end());
  HIPCHK(c->d_uw32.ensure(uw32.size()));
  HIPCHK(c->d_ph32.ensure(ph32.size()));
  HIPCHK(c->d_rh32.ensure(rh32.size()));
  HIPCHK(hipMemcpy(c->d_uw32.p, uw32.data(), uw32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_ph32.p, ph32.data(), ph32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_rh32.p, rh32.data(), rh32.size() * 4, hipMemcpyHostToDevice));
  dmd::Tables& T = c->T;
  T.uw = c->d_uw.p;
  T.ph = c->d_ph.p;
  T.rh = c->d_rh.p;
  T.em = c->d_em.p;
  T.uw32 = c->d_uw32.p;
  T.ph32 = c->d_ph32.p;
  T.rh32 = c->d_rh32.p;
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

dmd::State make_state(dmdgpu_ctx* c) {
  dmd::State s;
  s.n = c->n;
  s.qs = c->d_qs.p;
  s.ne = c->d_ne.p;
  s.xr = c->d_xr.p;
  s.spc = c->d_spc.p;
  s.beta = c->d_beta.p;
  s.gamma = c->d_gamma.p;
  s.nnb = c->d_nnb.p;
  s.nbr = c->d_nbr.p;
  return s;
}

dmd::Outputs make_outputs(dmdgpu_ctx* c) {
  dmd::Outputs o;
  o.rec = c->d_rec.p;
  o.V0 = c->d_V0.p;
  o.rho = c->d_rho.p;
  o.Rq = c->d_Rq.p;
  o.Rs = c->d_Rs.p;
  o.Rx = c->d_Rx.p;
  o.S0 = c->d_S0.p;
  o.partials = c->d_partials.p;
  o.totals = c->d_totals.p;
  return o;
}

dmdgpu_status check_ready(dmdgpu_ctx* c) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  if (!c->have_state) return fail(DMDGPU_ERR_STATE, "dmdgpu_set_state has not been called");
  if (!c->have_neigh) return fail(DMDGPU_ERR_STATE, "dmdgpu_neighbors has not been called");
  return DMDGPU_OK;
}

dmdgpu_status run_phase1(dmdgpu_ctx* c) {
  hipEvent_t a{};
  dmdgpu_status st = begin_timed(c, KID_PHASE1, &a);
  if (st) return st;
  HIPCHK(dmd::launch_phase1(c->stream, c->cfg, make_state(c), c->box, c->T, make_outputs(c)));
  return end_timed(c, KID_PHASE1, a);
}

dmdgpu_status run_residual(dmdgpu_ctx* c) {
  dmdgpu_status st = run_phase1(c);
  if (st) return st;
  hipEvent_t a{};
  if ((st = begin_timed(c, KID_PHASE2, &a))) return st;
  HIPCHK(dmd::launch_phase2(c->stream, c->cfg, make_state(c), c->box, c->T, make_outputs(c)));
  if ((st = end_timed(c, KID_PHASE2, a))) return st;
  if ((st = begin_timed(c, KID_TOTALS, &a))) return st;
  HIPCHK(dmd::launch_totals(c->stream, c->n, make_outputs(c)));
  return end_timed(c, KID_TOTALS, a);
}

dmdgpu_status d2h(dmdgpu_ctx* c, double* h, const double* d, size_t cnt) {
  if (!h) return DMDGPU_OK;
  HIPCHK(hipMemcpyAsync(h, d, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return DMDGPU_OK;
}

// Copies a per-site array of `width` doubles per site back to the caller's
// site order.
dmdgpu_status d2h_sites(dmdgpu_ctx* c, double* h, const double* d, int width) {
  if (!h) return DMDGPU_OK;
  const int n = c->n;
  std::vector<double> tmp((size_t)n * width);
  HIPCHK(hipMemcpyAsync(tmp.data(), d, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++)
    for (int k = 0; k < width; k++) h[(size_t)c->perm[i] * width + k] = tmp[(size_t)i * width + k];
  return DMDGPU_OK;
}

// Cell grid with edges >= r_c over the box (periodic axes) or the bounding box
// of the mean positions (free axes).
dmd::CellGrid make_grid(const dmdgpu_ctx* c) {
  dmd::CellGrid g{};
  const int n = c->n;
  for (int a = 0; a < 3; a++) {
    double lo, ext;
    if (c->box.per[a]) {
      lo = c->box.lo[a];
      ext = c->box.L[a];
    } else {
      double mn = 1e300, mx = -1e300;
      for (int i = 0; i < n; i++) {
        mn = std::min(mn, c->h_q[3 * i + a]);
        mx = std::max(mx, c->h_q[3 * i + a]);
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
void make_order(dmdgpu_ctx* c) {
  const int n = c->n;
  const dmd::CellGrid g = make_grid(c);
  std::vector<long long> key(n);
  for (int i = 0; i < n; i++) {
    int cc[3];
    for (int a = 0; a < 3; a++) {
      double v = c->h_q[3 * i + a] - g.lo[a];
      if (c->box.per[a]) v = v - c->box.L[a] * std::floor(v / c->box.L[a]);
      int k = (int)std::floor(v * g.inv_cs[a]);
      cc[a] = k < 0 ? 0 : (k > g.nc[a] - 1 ? g.nc[a] - 1 : k);
    }
    key[i] = ((long long)cc[2] * g.nc[1] + cc[1]) * g.nc[0] + cc[0];
  }
  c->perm.resize(n);
  for (int i = 0; i < n; i++) c->perm[i] = i;
  std::stable_sort(c->perm.begin(), c->perm.end(),
                   [&](int a, int b) { return key[a] < key[b]; });
  c->inv.resize(n);
  for (int i = 0; i < n; i++) c->inv[c->perm[i]] = i;
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
  dmdgpu_ctx* c = new dmdgpu_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(DMDGPU_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
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
  c->d_uw32.release();
  c->d_ph32.release();
  c->d_rh32.release();
  DevBuf<double>* dbs[] = {&c->d_uw,  &c->d_ph, &c->d_rh, &c->d_em, &c->d_ne, &c->d_xr,
                           &c->d_beta, &c->d_gamma, &c->d_rec, &c->d_V0, &c->d_rho, &c->d_Rq,
                           &c->d_Rs, &c->d_Rx, &c->d_S0, &c->d_partials, &c->d_totals};
  for (auto* b : dbs) b->release();
  DevBuf<int>* ibs[] = {&c->d_spc, &c->d_nnb, &c->d_nbr, &c->d_cell_of, &c->d_cell_count,
                        &c->d_cell_start, &c->d_cell_cursor, &c->d_cell_atoms, &c->d_flags};
  for (auto* b : ibs) b->release();
  c->d_qs.release();
  c->d_pairs.release();
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
  c->have_neigh = false;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_tuning(dmdgpu_ctx* c, int k1, int k2, int unroll2) {
  auto ok = [](int k) { return k == 1 || k == 2 || k == 4 || k == 8; };
  if (!c || !ok(k1) || !ok(k2)) return fail(DMDGPU_ERR_ARG, "lanes per site must be 1, 2, 4 or 8");
  c->cfg.k1 = k1;
  c->cfg.k2 = k2;
  c->cfg.unroll2 = unroll2 ? 1 : 0;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_precision(dmdgpu_ctx* c, int precision) {
  if (!c || (precision != DMDGPU_FP64 && precision != DMDGPU_SPLIT32))
    return fail(DMDGPU_ERR_ARG, "precision must be DMDGPU_FP64 or DMDGPU_SPLIT32");
  c->cfg.precision = precision == DMDGPU_SPLIT32 ? dmd::kSplit32 : dmd::kFP64;
  return DMDGPU_OK;
}

static dmdgpu_status upload_unknowns(dmdgpu_ctx* c, const double* q, const double* sg,
                                     const double* x, const int32_t* spc, bool first) {
  const int n = c->n;
  (void)spc;
  if (q) c->h_q.assign(q, q + 3 * (size_t)n);
  std::vector<double4> qs;
  if (q || sg) {
    qs.resize(n);
    if (!first) HIPCHK(hipMemcpy(qs.data(), c->d_qs.p, sizeof(double4) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
      const int o = c->perm[i];
      if (q) {
        qs[i].x = q[3 * o];
        qs[i].y = q[3 * o + 1];
        qs[i].z = q[3 * o + 2];
      }
      if (sg) {
        if (!(sg[o] > 0)) return fail(DMDGPU_ERR_ARG, "widths must be positive");
        qs[i].w = sg[o];
      }
    }
    HIPCHK(hipMemcpy(c->d_qs.p, qs.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
  }
  if (x) {
    std::vector<double> ne(n), xr(n);
    for (int i = 0; i < n; i++) {
      const int o = c->perm[i];
      if (c->h_slot[i] == 0 && x[o] < kMinOcc)
        return fail(DMDGPU_ERR_SPECIES, "Mg site with occupancy below 5e-4 (site " +
                                            std::to_string(o) + ")");
      xr[i] = x[o];
      ne[i] = x[o] >= kMinOcc ? x[o] : 0.0;
    }
    HIPCHK(hipMemcpy(c->d_ne.p, ne.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_xr.p, xr.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_set_state(dmdgpu_ctx* c, int n, const double* q, const double* sg,
                               const double* x, const int32_t* spc, const double* beta,
                               const double* gamma, const double* box, const int32_t* periodic) {
  if (!c || n <= 0 || !q || !sg || !x || !spc || !beta || !gamma || !box || !periodic)
    return fail(DMDGPU_ERR_ARG, "set_state: null argument or n <= 0");
  HIPCHK(hipSetDevice(c->device));
  for (int i = 0; i < n; i++)
    if (spc[i] != 1 && spc[i] != 12)
      return fail(DMDGPU_ERR_SPECIES, "species other than H (1) or Mg (12) at site " +
                                          std::to_string(i));
  dmd::Box b{};
  for (int a = 0; a < 3; a++) {
    b.lo[a] = box[a];
    b.L[a] = box[3 + a] - box[a];
    b.per[a] = periodic[a] ? 1 : 0;
    if (b.per[a] && !(b.L[a] >= 2.0 * c->rc))
      return fail(DMDGPU_ERR_BOX, "periodic box edge shorter than 2 r_c");
    b.invL[a] = b.L[a] > 0 ? 1.0 / b.L[a] : 0.0;
  }
  dmdgpu_status st = upload_tables(c);
  if (st) return st;
  c->n = n;
  c->box = b;
  HIPCHK(c->d_qs.ensure(n));
  HIPCHK(c->d_ne.ensure(n));
  HIPCHK(c->d_xr.ensure(n));
  HIPCHK(c->d_beta.ensure(n));
  HIPCHK(c->d_gamma.ensure(n));
  HIPCHK(c->d_spc.ensure(n));
  HIPCHK(c->d_nnb.ensure(n));
  HIPCHK(c->d_nbr.ensure((size_t)n * dmd::kCap));
  HIPCHK(c->d_cell_of.ensure(n));
  HIPCHK(c->d_cell_atoms.ensure(n));
  HIPCHK(c->d_flags.ensure(2));
  HIPCHK(c->d_rec.ensure((size_t)n * dmd::kRec));
  HIPCHK(c->d_V0.ensure(n));
  HIPCHK(c->d_rho.ensure(n));
  HIPCHK(c->d_Rq.ensure((size_t)3 * n));
  HIPCHK(c->d_Rs.ensure(n));
  HIPCHK(c->d_Rx.ensure(n));
  HIPCHK(c->d_S0.ensure(n));
  HIPCHK(c->d_partials.ensure((size_t)2 * ((n + 255) / 256)));
  HIPCHK(c->d_totals.ensure(2));
  HIPCHK(c->d_pairs.ensure((n + 255) / 256));
  c->h_q.assign(q, q + 3 * (size_t)n);
  make_order(c);
  c->h_slot.resize(n);
  std::vector<double> bp(n), gp(n);
  for (int i = 0; i < n; i++) {
    const int o = c->perm[i];
    c->h_slot[i] = spc[o] == 1 ? 1 : 0;
    bp[i] = beta[o];
    gp[i] = gamma[o];
  }
  HIPCHK(hipMemcpy(c->d_spc.p, c->h_slot.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_beta.p, bp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_gamma.p, gp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  st = upload_unknowns(c, q, sg, x, spc, true);
  if (st) return st;
  c->have_state = true;
  c->have_neigh = false;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_update_state(dmdgpu_ctx* c, const double* q, const double* sg,
                                  const double* x) {
  if (!c || !c->have_state) return fail(DMDGPU_ERR_STATE, "dmdgpu_set_state has not been called");
  HIPCHK(hipSetDevice(c->device));
  return upload_unknowns(c, q, sg, x, nullptr, false);
}

dmdgpu_status dmdgpu_neighbors(dmdgpu_ctx* c) {
  if (!c || !c->have_state) return fail(DMDGPU_ERR_STATE, "dmdgpu_set_state has not been called");
  HIPCHK(hipSetDevice(c->device));
  const dmd::CellGrid g = make_grid(c);
  const int n = c->n;
  const size_t ncell = (size_t)g.nc[0] * g.nc[1] * g.nc[2];
  HIPCHK(c->d_cell_count.ensure(ncell));
  HIPCHK(c->d_cell_start.ensure(ncell + 1));
  HIPCHK(c->d_cell_cursor.ensure(ncell));
  hipEvent_t ev{};
  dmdgpu_status st = begin_timed(c, KID_NEIGH, &ev);
  if (st) return st;
  HIPCHK(dmd::launch_neighbors(c->stream, c->box, g, c->rc, n, c->d_qs.p, c->d_cell_of.p,
                               c->d_cell_count.p, c->d_cell_start.p, c->d_cell_cursor.p,
                               c->d_cell_atoms.p, c->d_nnb.p, c->d_nbr.p, c->d_flags.p));
  if ((st = end_timed(c, KID_NEIGH, ev))) return st;
  int flags[2];
  HIPCHK(hipMemcpyAsync(flags, c->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (flags[0]) return fail(DMDGPU_ERR_MAXNEIGH, "a site has more than 250 neighbours");
  if (flags[1]) return fail(DMDGPU_ERR_NONEIGH, "a site has no neighbours");
  // active ordered pairs
  const int nb = (n + 255) / 256;
  HIPCHK(dmd::launch_active_pairs(c->stream, make_state(c), c->d_pairs.p, nb));
  std::vector<long long> pp(nb);
  HIPCHK(hipMemcpyAsync(pp.data(), c->d_pairs.p, sizeof(long long) * nb, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->pairs = 0;
  for (long long v : pp) c->pairs += v;
  c->have_neigh = true;
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_get_neighbors(dmdgpu_ctx* c, int32_t* h_nn, int32_t* h_list, int cap) {
  dmdgpu_status st = check_ready(c);
  if (st) return st;
  if (cap < dmd::kMaxNeigh) return fail(DMDGPU_ERR_ARG, "cap must be >= 250");
  const int n = c->n;
  std::vector<int> nn(n), rows((size_t)n * dmd::kCap);
  HIPCHK(hipMemcpyAsync(nn.data(), c->d_nnb.p, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(rows.data(), c->d_nbr.p, sizeof(int) * rows.size(), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<int> row;
  for (int o = 0; o < n; o++) {
    const int i = c->inv[o];
    if (h_nn) h_nn[o] = nn[i];
    if (!h_list) continue;
    row.assign(rows.begin() + (size_t)i * dmd::kCap, rows.begin() + (size_t)i * dmd::kCap + nn[i]);
    for (int& v : row) v = c->perm[v];
    std::sort(row.begin(), row.end());
    for (int k = 0; k < cap; k++) h_list[(size_t)o * cap + k] = k < nn[i] ? row[k] : -1;
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_mf_rho(dmdgpu_ctx* c, double* h_rho) {
  dmdgpu_status st = check_ready(c);
  if (st) return st;
  HIPCHK(hipSetDevice(c->device));
  if ((st = run_phase1(c))) return st;
  if ((st = d2h_sites(c, h_rho, c->d_rho.p, 1))) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_site_energies(dmdgpu_ctx* c, double* h_V0, double* h_S0) {
  dmdgpu_status st = check_ready(c);
  if (st) return st;
  HIPCHK(hipSetDevice(c->device));
  if ((st = run_residual(c))) return st;
  if ((st = d2h_sites(c, h_V0, c->d_V0.p, 1))) return st;
  if ((st = d2h_sites(c, h_S0, c->d_S0.p, 1))) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_residual(dmdgpu_ctx* c, double* h_Rq, double* h_Rs, double* h_Rx,
                              double* h_tot) {
  dmdgpu_status st = check_ready(c);
  if (st) return st;
  HIPCHK(hipSetDevice(c->device));
  if ((st = run_residual(c))) return st;
  if ((st = d2h_sites(c, h_Rq, c->d_Rq.p, 3))) return st;
  if ((st = d2h_sites(c, h_Rs, c->d_Rs.p, 1))) return st;
  if ((st = d2h_sites(c, h_Rx, c->d_Rx.p, 1))) return st;
  if ((st = d2h(c, h_tot, c->d_totals.p, 2))) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_residual_resident(dmdgpu_ctx* c) {
  dmdgpu_status st = check_ready(c);
  if (st) return st;
  return run_residual(c);
}

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
  dmdgpu_status st = drain_timing(c);
  if (st) return st;
  for (int k = 0; k < KID_COUNT; k++) {
    if (h_launches) h_launches[k] = c->launches[k];
    if (h_ms) h_ms[k] = c->ms[k];
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_reset_kernel_times(dmdgpu_ctx* c) {
  if (!c) return fail(DMDGPU_ERR_ARG, "null context");
  dmdgpu_status st = drain_timing(c);
  if (st) return st;
  for (int k = 0; k < KID_COUNT; k++) {
    c->launches[k] = 0;
    c->ms[k] = 0.0;
  }
  return DMDGPU_OK;
}

dmdgpu_status dmdgpu_work_counts(dmdgpu_ctx* c, int64_t* pairs, int64_t* pqe) {
  dmdgpu_status st = check_ready(c);
  if (st) return st;
  if (pairs) *pairs = c->pairs;
  if (pqe) *pqe = 50 * c->pairs;
  return DMDGPU_OK;
}

}  // extern "C"
