/*
 * dmd_oracle.c -- CPU FP64 restatement of the BT-AM mean-field DMD force path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).  Nothing in the
 * product library links or calls this file.  See dmd_oracle.h for the two forms
 * (reference order and factorised) and which reference file each one follows.
 */
#include "dmd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Physical constants, src/Macros.hpp:847-851. */
#define OR_HBAR 6.5821192815E-4
#define OR_UMA 103.4993333E-6
#define OR_SQRT2 1.4142135623730951 /* src/Macros.hpp:40 */
#define OR_SQRT6 2.449489742783178  /* src/Macros.hpp:42 */
#define SPC_H 1
#define SPC_MG 12

/* ------------------------------------------------------------------------- */
/* Splines: src/Numerical/cubic-spline.cpp:26-128                             */
/* ------------------------------------------------------------------------- */

static void spline_alloc(or_spline* s, int n, double dx) {
  s->n = n;
  s->dx = dx;
  double** cols[9] = {&s->a, &s->b, &s->c, &s->d, &s->db, &s->dc, &s->dd, &s->ddc, &s->ddd};
  for (int k = 0; k < 9; k++) *cols[k] = (double*)calloc(n > 0 ? n : 1, sizeof(double));
}

static void spline_free(or_spline* s) {
  free(s->a); free(s->b); free(s->c); free(s->d); free(s->db);
  free(s->dc); free(s->dd); free(s->ddc); free(s->ddd);
}

/* Segment index and offset with the reference clamp (cubic-spline.cpp:87-93):
 * m = min(trunc(x/dx), n-1), p = min(x - m dx, dx).  Negative m (x <= -dx) is
 * undefined behaviour in the reference; it is clamped to 0 here. */
static inline int seg(const or_spline* s, double x, double* p) {
  int m = (int)(x / s->dx);
  if (m > s->n - 1) m = s->n - 1;
  if (m < 0) m = 0;
  double pp = x - m * s->dx;
  if (pp > s->dx) pp = s->dx;
  *p = pp;
  return m;
}

static inline double sp_val(const or_spline* s, double x) {
  double p; int m = seg(s, x, &p);
  return s->a[m] + (s->b[m] + (s->c[m] + s->d[m] * p) * p) * p;
}
static inline double sp_d1(const or_spline* s, double x) {
  double p; int m = seg(s, x, &p);
  return s->db[m] + (s->dc[m] + s->dd[m] * p) * p;
}
static inline double sp_d2(const or_spline* s, double x) {
  double p; int m = seg(s, x, &p);
  return s->ddc[m] + s->ddd[m] * p;
}

/* ------------------------------------------------------------------------- */
/* Table loader: src/ADP/MgHx-ADP.cpp:366-501                                 */
/* ------------------------------------------------------------------------- */

static int read_block(FILE* f, or_spline* s) {
  int n; double dx;
  if (fscanf(f, "%d %lf", &n, &dx) != 2 || n <= 0 || !(dx > 0)) return -1;
  spline_alloc(s, n, dx);
  for (int i = 0; i < n; i++) {
    if (fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", &s->a[i], &s->b[i], &s->c[i],
               &s->d[i], &s->db[i], &s->dc[i], &s->dd[i], &s->ddc[i], &s->ddd[i]) != 9)
      return -1;
  }
  return 0;
}

/* Element file: "mass factor rcut", then rho, embed, pair, u, w blocks.
 * Alloy file: "rcut", then pair, u, w blocks. */
static int read_file(const char* folder, const char* name, int alloy, or_spline* rho,
                     or_spline* embed, or_spline* pair, or_spline* u, or_spline* w,
                     double* mass) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s", folder, name);
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  int rc = 0;
  double factor, rcut;
  if (!alloy) {
    if (fscanf(f, "%lf %lf %lf", mass, &factor, &rcut) != 3) rc = -1;
    if (!rc) rc = read_block(f, rho);
    if (!rc) rc = read_block(f, embed);
  } else {
    if (fscanf(f, "%lf", &rcut) != 1) rc = -1;
  }
  if (!rc) rc = read_block(f, pair);
  if (!rc) rc = read_block(f, u);
  if (!rc) rc = read_block(f, w);
  fclose(f);
  return rc;
}

int or_load_tables(or_tables** out, const char* folder) {
  or_tables* t = (or_tables*)calloc(1, sizeof(or_tables));
  int rc = read_file(folder, "adpmg.dat", 0, &t->rho[0], &t->embed[0], &t->pair[0], &t->u[0],
                     &t->w[0], &t->mass[0]);
  if (!rc)
    rc = read_file(folder, "adph.dat", 0, &t->rho[1], &t->embed[1], &t->pair[1], &t->u[1],
                   &t->w[1], &t->mass[1]);
  if (!rc)
    rc = read_file(folder, "adpmgh.dat", 1, NULL, NULL, &t->pair[2], &t->u[2], &t->w[2], NULL);
  if (rc) {
    or_free_tables(t);
    *out = NULL;
    return rc;
  }
  *out = t;
  return 0;
}

void or_free_tables(or_tables* t) {
  if (!t) return;
  for (int k = 0; k < 2; k++) {
    if (t->rho[k].a) spline_free(&t->rho[k]);
    if (t->embed[k].a) spline_free(&t->embed[k]);
  }
  for (int k = 0; k < 3; k++) {
    if (t->pair[k].a) spline_free(&t->pair[k]);
    if (t->u[k].a) spline_free(&t->u[k]);
    if (t->w[k].a) spline_free(&t->w[k]);
  }
  free(t);
}

static const or_spline* pick(const or_tables* t, int which, int slot) {
  switch (which) {
    case 0: return &t->rho[slot];
    case 1: return &t->embed[slot];
    case 2: return &t->pair[slot];
    case 3: return &t->u[slot];
    default: return &t->w[slot];
  }
}

double or_spline_eval(const or_tables* t, int which, int slot, int deriv, double x) {
  const or_spline* s = pick(t, which, slot);
  return deriv == 0 ? sp_val(s, x) : deriv == 1 ? sp_d1(s, x) : sp_d2(s, x);
}

int or_spline_size(const or_tables* t, int which, int slot, double* dx) {
  const or_spline* s = pick(t, which, slot);
  if (dx) *dx = s->dx;
  return s->n;
}

/* Species helpers (MgHx-ADP.cpp:538-543, 881-888). */
static inline int sslot(int spc) { return spc == SPC_H ? 1 : 0; }
static inline int pslot(int si, int sj) {
  if (si == SPC_MG && sj == SPC_MG) return 0;
  if (si == SPC_H && sj == SPC_H) return 1;
  return 2;
}

/* ------------------------------------------------------------------------- */
/* Geometry                                                                   */
/* ------------------------------------------------------------------------- */

/* Minimum-image separation q_i - q_j on periodic axes. */
static inline void sep(const double* lo, const double* hi, const int* per, const double* qi,
                       const double* qj, double* d) {
  for (int a = 0; a < 3; a++) {
    double v = qi[a] - qj[a];
    if (per[a]) {
      double L = hi[a] - lo[a];
      v -= L * rint(v / L);
    }
    d[a] = v;
  }
}

/* Brute-force neighbour list over mean positions, cut-off on the norm with
 * "<=" (Neighbors.cpp:67-117).  Returns 0, 1 = more than cap neighbours,
 * 2 = a site without neighbours. */
int or_neighbors(int n, const double* q, const double* lo, const double* hi,
                 const int* periodic, double rc, int cap, int* nnb, int* nbr) {
  int over = 0, empty = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : over, empty)
  for (int i = 0; i < n; i++) {
    int c = 0;
    for (int j = 0; j < n; j++) {
      if (j == i) continue;
      double d[3];
      sep(lo, hi, periodic, q + 3 * i, q + 3 * j, d);
      double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (r <= rc) {
        if (c == cap) { over = 1; break; }
        nbr[(long)i * cap + c] = j;
        c++;
      }
    }
    nnb[i] = c;
    if (c == 0) empty = 1;
  }
  return over ? 1 : (empty ? 2 : 0);
}

/* Image-consistent position of neighbour j seen from site i. */
static inline void img(const or_system* s, int i, int j, double* qj) {
  double d[3];
  sep(s->lo, s->hi, s->periodic, s->q + 3 * i, s->q + 3 * j, d);
  for (int a = 0; a < 3; a++) qj[a] = s->q[3 * i + a] - d[a];
}

/* ------------------------------------------------------------------------- */
/* ADP pair / triplet functions: src/ADP/MgHx-ADP.cpp                          */
/* Arguments follow potential_function (src/Macros.hpp:252-278): n = site     */
/* occupancies, q = site positions (3 per site), spc = species.               */
/* ------------------------------------------------------------------------- */

typedef struct {
  void (*F)(const or_tables*, double*, const double*, const double*, const int*);
  void (*dF_dq)(const or_tables*, int, double*, const double*, const double*, const int*);
  void (*dF_dn)(const or_tables*, int, double*, const double*, const double*, const int*);
  void (*d2F_dq2)(const or_tables*, int, double*, const double*, const double*, const int*);
} or_fn;

static inline double rvec(const double* q, int a, int b, double* r) {
  double r2 = 0;
  for (int k = 0; k < 3; k++) {
    r[k] = q[3 * a + k] - q[3 * b + k];
    r2 += r[k] * r[k];
  }
  return r2;
}

/* rho_ij (:526-553), d_rho_ij_dq (:557-601), d_rho_ij_dn (:815-844) */
static void f_rho(const or_tables* T, double* out, const double* n, const double* q, const int* spc) {
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  *out = n[1] * sp_val(&T->rho[sslot(spc[1])], nr);
}
static void f_rho_dq(const or_tables* T, int dir, double* out, const double* n, const double* q,
                     const int* spc) {
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  double c = n[1] * sp_d1(&T->rho[sslot(spc[1])], nr) / nr;
  for (int k = 0; k < 3; k++) out[k] = dir == 0 ? c * r[k] : dir == 1 ? -c * r[k] : 0.0;
}
static void f_rho_dn(const or_tables* T, int dir, double* out, const double* n, const double* q,
                     const int* spc) {
  (void)n;
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  *out = dir == 1 ? sp_val(&T->rho[sslot(spc[1])], nr) : 0.0;
}

/* V_pair_ij (:869-900), dV_pair_ij_dq (:904-949), dV_pair_ij_dn (:1170-1209) */
static void f_pair(const or_tables* T, double* out, const double* n, const double* q, const int* spc) {
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  *out = 0.5 * n[0] * n[1] * sp_val(&T->pair[pslot(spc[0], spc[1])], nr);
}
static void f_pair_dq(const or_tables* T, int dir, double* out, const double* n, const double* q,
                      const int* spc) {
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  double c = 0.5 * n[0] * n[1] * sp_d1(&T->pair[pslot(spc[0], spc[1])], nr) / nr;
  for (int k = 0; k < 3; k++) out[k] = dir == 0 ? c * r[k] : dir == 1 ? -c * r[k] : 0.0;
}
static void f_pair_dn(const or_tables* T, int dir, double* out, const double* n, const double* q,
                      const int* spc) {
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  double phi = sp_val(&T->pair[pslot(spc[0], spc[1])], nr);
  *out = dir == 0 ? 0.5 * n[1] * phi : dir == 1 ? 0.5 * n[0] * phi : 0.0;
}

/* Shared triplet geometry for the dipole and quadrupole terms. */
typedef struct {
  double r1[3], r2[3], r1s, r2s, n1, n2, dot;
} trip;
static inline void trip_geo(const double* q, trip* g) {
  g->r1s = rvec(q, 0, 1, g->r1);
  g->r2s = rvec(q, 0, 2, g->r2);
  g->n1 = sqrt(g->r1s);
  g->n2 = sqrt(g->r2s);
  g->dot = g->r1[0] * g->r2[0] + g->r1[1] * g->r2[1] + g->r1[2] * g->r2[2];
}

/* V_dipole_ij1j2 (:1233-1281): 1/2 (n_i n_j1 u1)(n_i n_j2 u2)(r1 . r2) */
static void f_dip(const or_tables* T, double* out, const double* n, const double* q, const int* spc) {
  trip g; trip_geo(q, &g);
  double a1 = n[0] * n[1] * sp_val(&T->u[pslot(spc[0], spc[1])], g.n1);
  double a2 = n[0] * n[2] * sp_val(&T->u[pslot(spc[0], spc[2])], g.n2);
  *out = 0.5 * a1 * a2 * g.dot;
}
/* dV_dipole_ij1j2_dq (:1285-1371) */
static void f_dip_dq(const or_tables* T, int dir, double* out, const double* n, const double* q,
                     const int* spc) {
  trip g; trip_geo(q, &g);
  const or_spline* s1 = &T->u[pslot(spc[0], spc[1])];
  const or_spline* s2 = &T->u[pslot(spc[0], spc[2])];
  double a1 = n[0] * n[1] * sp_val(s1, g.n1), a2 = n[0] * n[2] * sp_val(s2, g.n2);
  double d1 = n[0] * n[1] * sp_d1(s1, g.n1), d2 = n[0] * n[2] * sp_d1(s2, g.n2);
  for (int k = 0; k < 3; k++) {
    double t1 = d1 * a2 * g.dot / g.n1 * g.r1[k];
    double t2 = a1 * d2 * g.dot / g.n2 * g.r2[k];
    if (dir == 0) out[k] = 0.5 * (t1 + t2 + a1 * a2 * (g.r1[k] + g.r2[k]));
    else if (dir == 1) out[k] = -0.5 * (t1 + a1 * a2 * g.r2[k]);
    else out[k] = -0.5 * (t2 + a1 * a2 * g.r1[k]);
  }
}
/* dV_dipole_ij1j2_dn (:1683-1758) */
static void f_dip_dn(const or_tables* T, int dir, double* out, const double* n, const double* q,
                     const int* spc) {
  trip g; trip_geo(q, &g);
  double u1 = sp_val(&T->u[pslot(spc[0], spc[1])], g.n1);
  double u2 = sp_val(&T->u[pslot(spc[0], spc[2])], g.n2);
  double a1 = n[0] * n[1] * u1, a2 = n[0] * n[2] * u2;
  if (dir == 0) *out = 0.5 * (n[1] * u1) * a2 * g.dot + 0.5 * a1 * (n[2] * u2) * g.dot;
  else if (dir == 1) *out = 0.5 * (n[0] * u1) * a2 * g.dot;
  else *out = 0.5 * a1 * (n[0] * u2) * g.dot;
}

/* V_quadrupole_ij1j2 (:1781-1831): 1/2 w1 w2 (r1.r2)^2 - 1/6 w1 w2 r1^2 r2^2 */
static void f_quad(const or_tables* T, double* out, const double* n, const double* q, const int* spc) {
  trip g; trip_geo(q, &g);
  double b1 = n[0] * n[1] * sp_val(&T->w[pslot(spc[0], spc[1])], g.n1);
  double b2 = n[0] * n[2] * sp_val(&T->w[pslot(spc[0], spc[2])], g.n2);
  *out = 0.5 * b1 * b2 * g.dot * g.dot - (1.0 / 6.0) * b1 * b2 * g.r1s * g.r2s;
}
/* dV_quadrupole_ij1j2_dq (:1835-1935) */
static void f_quad_dq(const or_tables* T, int dir, double* out, const double* n, const double* q,
                      const int* spc) {
  trip g; trip_geo(q, &g);
  const or_spline* s1 = &T->w[pslot(spc[0], spc[1])];
  const or_spline* s2 = &T->w[pslot(spc[0], spc[2])];
  double b1 = n[0] * n[1] * sp_val(s1, g.n1), b2 = n[0] * n[2] * sp_val(s2, g.n2);
  double d1 = n[0] * n[1] * sp_d1(s1, g.n1), d2 = n[0] * n[2] * sp_d1(s2, g.n2);
  double dd = g.dot * g.dot;
  for (int k = 0; k < 3; k++) {
    double h1 = d1 * b2 * dd / g.n1 * g.r1[k];
    double h2 = b1 * d2 * dd / g.n2 * g.r2[k];
    double e1 = d1 * b2 * g.n1 * g.r2s * g.r1[k];
    double e2 = b1 * d2 * g.r1s * g.n2 * g.r2[k];
    if (dir == 0)
      out[k] = 0.5 * (h1 + h2 + b1 * b2 * 2 * g.dot * (g.r1[k] + g.r2[k])) -
               (1.0 / 6.0) * (e1 + e2 + b1 * b2 * 2 * g.r2s * g.r1[k] + b1 * b2 * 2 * g.r1s * g.r2[k]);
    else if (dir == 1)
      out[k] = -0.5 * (h1 + b1 * b2 * 2 * g.dot * g.r2[k]) +
               (1.0 / 6.0) * (e1 + b1 * b2 * 2 * g.r2s * g.r1[k]);
    else
      out[k] = -0.5 * (h2 + b1 * b2 * 2 * g.dot * g.r1[k]) +
               (1.0 / 6.0) * (e2 + b1 * b2 * 2 * g.r1s * g.r2[k]);
  }
}
/* dV_quadrupole_ij1j2_dn (:2352-2434) */
static void f_quad_dn(const or_tables* T, int dir, double* out, const double* n, const double* q,
                      const int* spc) {
  trip g; trip_geo(q, &g);
  double w1 = sp_val(&T->w[pslot(spc[0], spc[1])], g.n1);
  double w2 = sp_val(&T->w[pslot(spc[0], spc[2])], g.n2);
  double b1 = n[0] * n[1] * w1, b2 = n[0] * n[2] * w2;
  double dd = g.dot * g.dot, rr = g.r1s * g.r2s;
  if (dir == 0) {
    double c1 = n[1] * w1, c2 = n[2] * w2;
    *out = 0.5 * c1 * b2 * dd + 0.5 * b1 * c2 * dd - (1.0 / 6.0) * c1 * b2 * rr -
           (1.0 / 6.0) * b1 * c2 * rr;
  } else if (dir == 1) {
    double c1 = n[0] * w1;
    *out = 0.5 * c1 * b2 * dd - (1.0 / 6.0) * c1 * b2 * rr;
  } else {
    double c2 = n[0] * w2;
    *out = 0.5 * b1 * c2 * dd - (1.0 / 6.0) * b1 * c2 * rr;
  }
}

/* Second derivatives.  code = ns * a + b is d2F / dq_a dq_b (3 x 3, row
 * index from q_a); the reference's Hessians are d2_rho_ij_dq2 (:664-731),
 * d2V_pair_ij_dq2 (:1013-1087), dV2_dipole_ij1j2_dq2 (:1435-1601) and
 * dV2_quadrupole_ij1j2_dq2 (:2000-2266), bound at :514, :855, :1220, :1769.  Written here as products of radial
 * factors: for c(|r|), d2c/dr2 = c'' rr^T/r^2 + c' (I - rr^T/r^2)/r. */
static inline void rad_hess(double d1, double d2, const double* r, double nr, double k,
                            double* H) {
  double inv = 1.0 / nr;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double tt = r[a] * r[b] * inv * inv;
      H[3 * a + b] += k * (d2 * tt + d1 * ((a == b) - tt) * inv);
    }
}
/* Pair functions depend on r = q_0 - q_1 only: codes 0 and 3 carry +, 1 and 2 -. */
static inline double pair_code_sign(int code) { return (code == 0 || code == 3) ? 1.0 : -1.0; }
static void f_rho_d2q(const or_tables* T, int code, double* out, const double* n, const double* q,
                      const int* spc) {
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  const or_spline* s = &T->rho[sslot(spc[1])];
  memset(out, 0, 9 * sizeof(double));
  rad_hess(sp_d1(s, nr), sp_d2(s, nr), r, nr, pair_code_sign(code) * n[1], out);
}
static void f_pair_d2q(const or_tables* T, int code, double* out, const double* n, const double* q,
                       const int* spc) {
  double r[3]; double nr = sqrt(rvec(q, 0, 1, r));
  const or_spline* s = &T->pair[pslot(spc[0], spc[1])];
  memset(out, 0, 9 * sizeof(double));
  rad_hess(sp_d1(s, nr), sp_d2(s, nr), r, nr, pair_code_sign(code) * 0.5 * n[0] * n[1], out);
}

/* Triplet functions f = c1(|r1|) c2(|r2|) G(r1, r2), r_k = q_0 - q_k.  With
 * e_k = dc_k/dr_k, E_k = d2c_k/dr_k2, g_k = dG/dr_k and G_kl = d2G/dr_k dr_l:
 *   H11 = c2 (G E1 + e1 g1^T + g1 e1^T + c1 G11),  H22 likewise,
 *   H12 = G e1 e2^T + c2 e1 g2^T + c1 g1 e2^T + c1 c2 G12   (= d2f/dr1 dr2),
 * H21 = H12^T.  Since d/dq_0 = d/dr1 + d/dr2 and d/dq_k = -d/dr_k, code
 * 3 a + b is sum_kl w_ak w_bl H_kl with w_0 = (1, 1), w_1 = (-1, 0),
 * w_2 = (0, -1): code 0 = H11 + H12 + H21 + H22, code 4 = H11, code 8 = H22. */
typedef struct { double G, g1[3], g2[3], G11[9], G22[9], G12[9]; } bil;
static void trip_d2(const trip* t, const or_spline* s1, const or_spline* s2, double k1, double k2,
                    const bil* b, int code, double* out) {
  double c1 = k1 * sp_val(s1, t->n1), d1 = k1 * sp_d1(s1, t->n1), dd1 = k1 * sp_d2(s1, t->n1);
  double c2 = k2 * sp_val(s2, t->n2), d2 = k2 * sp_d1(s2, t->n2), dd2 = k2 * sp_d2(s2, t->n2);
  double e1[3], e2[3], E1[9] = {0}, E2[9] = {0}, H11[9], H22[9], H12[9];
  for (int a = 0; a < 3; a++) { e1[a] = d1 * t->r1[a] / t->n1; e2[a] = d2 * t->r2[a] / t->n2; }
  rad_hess(d1, dd1, t->r1, t->n1, 1.0, E1);
  rad_hess(d2, dd2, t->r2, t->n2, 1.0, E2);
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++) {
      int ac = 3 * a + c;
      H11[ac] = c2 * (b->G * E1[ac] + e1[a] * b->g1[c] + b->g1[a] * e1[c] + c1 * b->G11[ac]);
      H22[ac] = c1 * (b->G * E2[ac] + e2[a] * b->g2[c] + b->g2[a] * e2[c] + c2 * b->G22[ac]);
      H12[ac] = b->G * e1[a] * e2[c] + c2 * e1[a] * b->g2[c] + c1 * b->g1[a] * e2[c] +
                c1 * c2 * b->G12[ac];
    }
  /* d/dq_0 = d/dr1 + d/dr2, d/dq_1 = -d/dr1, d/dq_2 = -d/dr2 */
  static const double wt[3][2] = {{1, 1}, {-1, 0}, {0, -1}};
  int sa = code / 3, sb = code % 3;
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++)
      out[3 * a + c] = wt[sa][0] * wt[sb][0] * H11[3 * a + c] +
                       wt[sa][0] * wt[sb][1] * H12[3 * a + c] +
                       wt[sa][1] * wt[sb][0] * H12[3 * c + a] +
                       wt[sa][1] * wt[sb][1] * H22[3 * a + c];
}
/* dipole: G = 1/2 r1.r2 */
static void f_dip_d2q(const or_tables* T, int code, double* out, const double* n, const double* q,
                      const int* spc) {
  trip g; trip_geo(q, &g);
  bil b;
  memset(&b, 0, sizeof(b));
  b.G = 0.5 * g.dot;
  for (int a = 0; a < 3; a++) {
    b.g1[a] = 0.5 * g.r2[a];
    b.g2[a] = 0.5 * g.r1[a];
    b.G12[4 * a] = 0.5;
  }
  trip_d2(&g, &T->u[pslot(spc[0], spc[1])], &T->u[pslot(spc[0], spc[2])], n[0] * n[1],
          n[0] * n[2], &b, code, out);
}
/* quadrupole: G = 1/2 (r1.r2)^2 - 1/6 r1^2 r2^2 */
static void f_quad_d2q(const or_tables* T, int code, double* out, const double* n, const double* q,
                       const int* spc) {
  trip g; trip_geo(q, &g);
  bil b;
  b.G = 0.5 * g.dot * g.dot - g.r1s * g.r2s / 6.0;
  for (int a = 0; a < 3; a++) {
    b.g1[a] = g.dot * g.r2[a] - g.r2s * g.r1[a] / 3.0;
    b.g2[a] = g.dot * g.r1[a] - g.r1s * g.r2[a] / 3.0;
    for (int c = 0; c < 3; c++) {
      double id = (a == c);
      b.G11[3 * a + c] = g.r2[a] * g.r2[c] - g.r2s * id / 3.0;
      b.G22[3 * a + c] = g.r1[a] * g.r1[c] - g.r1s * id / 3.0;
      b.G12[3 * a + c] = g.r2[a] * g.r1[c] + g.dot * id - 2.0 * g.r1[a] * g.r2[c] / 3.0;
    }
  }
  trip_d2(&g, &T->w[pslot(spc[0], spc[1])], &T->w[pslot(spc[0], spc[2])], n[0] * n[1],
          n[0] * n[2], &b, code, out);
}

static const or_fn FN_RHO = {f_rho, f_rho_dq, f_rho_dn, f_rho_d2q};
static const or_fn FN_PAIR = {f_pair, f_pair_dq, f_pair_dn, f_pair_d2q};
static const or_fn FN_DIP = {f_dip, f_dip_dq, f_dip_dn, f_dip_d2q};
static const or_fn FN_QUAD = {f_quad, f_quad_dq, f_quad_dn, f_quad_d2q};

/* ------------------------------------------------------------------------- */
/* Gaussian measure: src/Numerical/Quadrature-Measure.cpp:30-110              */
/* ------------------------------------------------------------------------- */

typedef struct {
  int ns;          /* sites: 2 (pair) or 3 (triplet)        */
  int D;           /* integral dimension                     */
  int board[9][9]; /* board[row][3*site+axis] in {0,1}       */
  double q[9], sig[3], n[3];
  int spc[3];
} measure;

/* Removes sites whose dof-table row duplicates an earlier one (they move
 * together) and builds the D x 3ns point board, D = 3 x (distinct rows). */
static void fill_measure(measure* m, const int* dof, int ns) {
  int aux[9];
  memcpy(aux, dof, sizeof(int) * ns * ns);
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++)
      if (j != i)
        for (int k = 0; k < ns; k++)
          if (aux[j * ns + k] == aux[i * ns + k]) aux[j * ns + k] = 0;
  memset(m->board, 0, sizeof(m->board));
  int rows = 0;
  for (int i = 0; i < ns; i++) {
    int cnt = 0;
    for (int j = 0; j < ns; j++) cnt += aux[i * ns + j];
    if (cnt <= 0) continue;
    for (int j = 0; j < ns; j++)
      if (aux[i * ns + j])
        for (int k = 0; k < 3; k++) m->board[3 * rows + k][3 * j + k] = 1;
    rows++;
  }
  m->ns = ns;
  m->D = 3 * rows;
}

/* ------------------------------------------------------------------------- */
/* Stroud degree-3 rule: src/Numerical/Quadrature-Hermitian-3th.cpp           */
/* Points q +/- sqrt(2) sigma zeta e_l with zeta = sqrt(D/2), weight 1/(2D).  */
/* ------------------------------------------------------------------------- */

static inline void gh3_point(const measure* m, int l, double sgn, double* ql) {
  double zeta = sqrt((double)m->D) / OR_SQRT2;
  for (int s = 0; s < m->ns; s++)
    for (int a = 0; a < 3; a++)
      ql[3 * s + a] = m->q[3 * s + a] + sgn * OR_SQRT2 * m->sig[s] * zeta * m->board[l][3 * s + a];
}

/* meanfield_integral_gh3th (:36-110) */
static double gh3(const or_tables* T, const measure* m, const or_fn* f) {
  double W = 1.0 / (2.0 * m->D), acc = 0.0, ql[9], v;
  for (int sg = 0; sg < 2; sg++)
    for (int l = 0; l < m->D; l++) {
      gh3_point(m, l, sg ? -1.0 : 1.0, ql);
      f->F(T, &v, m->n, ql, m->spc);
      acc += v * W;
    }
  return acc;
}
/* meanfield_integral_gh_3th_dmq (:212-296) */
static void gh3_dmq(const or_tables* T, const measure* m, const or_fn* f, int dir, double* out) {
  double W = 1.0 / (2.0 * m->D), ql[9], g[3];
  out[0] = out[1] = out[2] = 0.0;
  for (int sg = 0; sg < 2; sg++)
    for (int l = 0; l < m->D; l++) {
      gh3_point(m, l, sg ? -1.0 : 1.0, ql);
      f->dF_dq(T, dir, g, m->n, ql, m->spc);
      for (int a = 0; a < 3; a++) out[a] += g[a] * W;
    }
}
/* meanfield_integral_gh3th_dsq (:114-208) */
static double gh3_dsq(const or_tables* T, const measure* m, const or_fn* f, int dir) {
  double W = 1.0 / (2.0 * m->D), ql[9], g[3], acc = 0.0;
  double zeta = sqrt((double)m->D) / OR_SQRT2;
  for (int sg = 0; sg < 2; sg++)
    for (int l = 0; l < m->D; l++) {
      double sgn = sg ? -1.0 : 1.0;
      gh3_point(m, l, sgn, ql);
      f->dF_dq(T, dir, g, m->n, ql, m->spc);
      for (int a = 0; a < 3; a++) acc += g[a] * (sgn * OR_SQRT2 * zeta * m->board[l][3 * dir + a]) * W;
    }
  return acc;
}
/* meanfield_integral_gh_3th_dxi (:300-378) */
static double gh3_dxi(const or_tables* T, const measure* m, const or_fn* f, int dir) {
  double W = 1.0 / (2.0 * m->D), ql[9], v, acc = 0.0;
  for (int sg = 0; sg < 2; sg++)
    for (int l = 0; l < m->D; l++) {
      gh3_point(m, l, sg ? -1.0 : 1.0, ql);
      f->dF_dn(T, dir, &v, m->n, ql, m->spc);
      acc += v * W;
    }
  return acc;
}

/* One pair / triplet function at explicit positions (for tests): which 0 rho,
 * 1 pair, 2 dipole, 3 quadrupole; deriv 0 value (out[0]), 1 dF/dq_code (3),
 * 2 d2F for code (3 x 3).  Species are slots here (0 Mg, 1 H). */
int or_fn_eval(const or_tables* T, int which, int deriv, int code, const double* n,
               const double* q, const int* slots, double* out) {
  static const or_fn* fns[4] = {&FN_RHO, &FN_PAIR, &FN_DIP, &FN_QUAD};
  if (which < 0 || which > 3) return 1;
  int spc[3];
  for (int k = 0; k < (which < 2 ? 2 : 3); k++) spc[k] = slots[k] ? SPC_H : SPC_MG;
  const or_fn* f = fns[which];
  if (deriv == 0) f->F(T, out, n, q, spc);
  else if (deriv == 1) f->dF_dq(T, code, out, n, q, spc);
  else f->d2F_dq2(T, code, out, n, q, spc);
  return 0;
}

/* meanfield_integral_gh_3th_d2mq (:382-463): the rule applied to d2F/dq2. */
static void gh3_d2mq(const or_tables* T, const measure* m, const or_fn* f, int code, double* out) {
  double W = 1.0 / (2.0 * m->D), ql[9], h[9];
  for (int a = 0; a < 9; a++) out[a] = 0.0;
  for (int sg = 0; sg < 2; sg++)
    for (int l = 0; l < m->D; l++) {
      gh3_point(m, l, sg ? -1.0 : 1.0, ql);
      f->d2F_dq2(T, code, h, m->n, ql, m->spc);
      for (int a = 0; a < 9; a++) out[a] += h[a] * W;
    }
}

/* Measure set-up for the pair (i, j) and the triplet (i, j1, j2), with
 * image-consistent neighbour positions (MgHx-mf-V-bulk.cpp:760-776, 894-916). */
static void pair_measure(const or_system* s, int i, int j, measure* m) {
  static const int dof[4] = {1, 0, 0, 1};
  fill_measure(m, dof, 2);
  for (int a = 0; a < 3; a++) m->q[a] = s->q[3 * i + a];
  img(s, i, j, m->q + 3);
  m->sig[0] = s->sig[i]; m->sig[1] = s->sig[j];
  m->n[0] = s->x[i]; m->n[1] = s->x[j];
  m->spc[0] = s->spc[i]; m->spc[1] = s->spc[j];
}
static void trip_measure(const or_system* s, int i, int j1, int j2, measure* m) {
  int dof[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (j1 == j2) { dof[5] = 1; dof[7] = 1; }
  fill_measure(m, dof, 3);
  for (int a = 0; a < 3; a++) m->q[a] = s->q[3 * i + a];
  img(s, i, j1, m->q + 3);
  img(s, i, j2, m->q + 6);
  m->sig[0] = s->sig[i]; m->sig[1] = s->sig[j1]; m->sig[2] = s->sig[j2];
  m->n[0] = s->x[i]; m->n[1] = s->x[j1]; m->n[2] = s->x[j2];
  m->spc[0] = s->spc[i]; m->spc[1] = s->spc[j1]; m->spc[2] = s->spc[j2];
}

/* ------------------------------------------------------------------------- */
/* Per-site evaluators in reference order: src/ADP/MgHx-mf-V-bulk.cpp          */
/* ------------------------------------------------------------------------- */

#define NB(s, i, k) ((s)->nbr[(long)(i) * (s)->cap + (k)])

/* Zero-temperature density, src/ADP/density.cpp:44-104 (masks any species). */
double or_ref_rho0_i(const or_system* s, const or_tables* T, int i) {
  if (s->x[i] < OR_MIN_OCCUPANCY) return 0.0;
  double acc = 0.0;
  for (int k = 0; k < s->nnb[i]; k++) {
    int j = NB(s, i, k);
    if (s->x[j] < OR_MIN_OCCUPANCY) continue;
    double q[6], n[2] = {s->x[i], s->x[j]}, v;
    int spc[2] = {s->spc[i], s->spc[j]};
    for (int a = 0; a < 3; a++) q[a] = s->q[3 * i + a];
    img(s, i, j, q + 3);
    f_rho(T, &v, n, q, spc);
    acc += v;
  }
  return acc;
}

/* evaluate_mf_rho_i_adp_MgHx (:708-787) */
double or_ref_mf_rho_i(const or_system* s, const or_tables* T, int i) {
  if (s->spc[i] == SPC_H && s->x[i] < OR_MIN_OCCUPANCY) return 0.0;
  double acc = 0.0;
  for (int k = 0; k < s->nnb[i]; k++) {
    int j = NB(s, i, k);
    if (s->spc[j] == SPC_H && s->x[j] < OR_MIN_OCCUPANCY) continue;
    measure m;
    pair_measure(s, i, j, &m);
    acc += gh3(T, &m, &FN_RHO);
  }
  return acc;
}

/* Pair + angular part of V0_i (:842-936); embedding added by the callers. */
static double ref_V0_nonembed(const or_system* s, const or_tables* T, int i) {
  double vp = 0.0, vd = 0.0, vq = 0.0;
  for (int k1 = 0; k1 < s->nnb[i]; k1++) {
    int j1 = NB(s, i, k1);
    if (s->spc[j1] == SPC_H && s->x[j1] < OR_MIN_OCCUPANCY) continue;
    measure m;
    pair_measure(s, i, j1, &m);
    vp += gh3(T, &m, &FN_PAIR);
    for (int k2 = k1; k2 < s->nnb[i]; k2++) {
      int j2 = NB(s, i, k2);
      if (s->x[j2] < OR_MIN_OCCUPANCY) continue;
      double f = (k2 == k1) ? 1.0 : 2.0;
      measure mt;
      trip_measure(s, i, j1, j2, &mt);
      vd += f * gh3(T, &mt, &FN_DIP);
      vq += f * gh3(T, &mt, &FN_QUAD);
    }
  }
  return vp + vd + vq;
}

/* evaluate_V0_i_adp_MgHx (:791-953) */
double or_ref_V0_i(const or_system* s, const or_tables* T, int i, const double* mf_rho) {
  if (s->spc[i] == SPC_H && s->x[i] < OR_MIN_OCCUPANCY) return 0.0;
  double rest = ref_V0_nonembed(s, T, i);
  return s->x[i] * sp_val(&T->embed[sslot(s->spc[i])], mf_rho[i]) + rest;
}

/* evaluate_S0_i_adp_MgHx (:957-1142) */
double or_ref_S0_i(const or_system* s, const or_tables* T, int i, const double* mf_rho) {
  double xi = s->x[i], b = s->beta[i];
  if (xi < OR_MIN_OCCUPANCY) return 0.0;
  double mi = OR_UMA * xi * T->mass[sslot(s->spc[i])];
  double V0 = xi * sp_val(&T->embed[sslot(s->spc[i])], mf_rho[i]) + ref_V0_nonembed(s, T, i);
  double H0 = 1.0 / (2.0 * b) + V0;
  double logZ = 3.0 * log((s->sig[i] * sqrt(mi / b)) / OR_HBAR);
  if (xi < 1.0) logZ += log(1.0 / (1.0 - xi));
  double gx = 0.0;
  if (s->spc[i] == SPC_H) gx = s->gamma[i] * (xi > 0.9999 ? 0.9999 : xi);
  return -logZ + b * H0 - gx;
}

/* evaluate_DV0_i_Dq_u_MgHx (:1823-2098): [dV0_i/dq_u (3), dV0_i/dsigma_u]. */
void or_ref_DV0_i_Dq_u(const or_system* s, const or_tables* T, int u, int i,
                       const double* mf_rho, double out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (s->x[u] < OR_MIN_OCCUPANCY) return;
  if (s->x[i] < OR_MIN_OCCUPANCY) return;
  double drho[3] = {0, 0, 0}, drho_s = 0.0, dq[3] = {0, 0, 0}, ds = 0.0, g[3];
  for (int k1 = 0; k1 < s->nnb[i]; k1++) {
    int j1 = NB(s, i, k1);
    if (s->x[j1] < OR_MIN_OCCUPANCY) continue;
    measure m;
    pair_measure(s, i, j1, &m);
    int sites2[2] = {i, j1};
    for (int dir = 0; dir < 2; dir++) {
      if (sites2[dir] != u) continue;
      gh3_dmq(T, &m, &FN_RHO, dir, g);
      for (int a = 0; a < 3; a++) drho[a] += g[a];
      gh3_dmq(T, &m, &FN_PAIR, dir, g);
      for (int a = 0; a < 3; a++) dq[a] += g[a];
      drho_s += gh3_dsq(T, &m, &FN_RHO, dir);
      ds += gh3_dsq(T, &m, &FN_PAIR, dir);
    }
    for (int k2 = k1; k2 < s->nnb[i]; k2++) {
      int j2 = NB(s, i, k2);
      if (s->x[j2] <= OR_MIN_OCCUPANCY) continue; /* "<=" as in :1986 */
      double f = (k1 == k2) ? 1.0 : 2.0;
      int sites3[3] = {i, j1, j2};
      if (sites3[0] != u && sites3[1] != u && sites3[2] != u) continue;
      measure mt;
      trip_measure(s, i, j1, j2, &mt);
      for (int dir = 0; dir < 3; dir++) {
        if (sites3[dir] != u) continue;
        gh3_dmq(T, &mt, &FN_DIP, dir, g);
        for (int a = 0; a < 3; a++) dq[a] += f * g[a];
        gh3_dmq(T, &mt, &FN_QUAD, dir, g);
        for (int a = 0; a < 3; a++) dq[a] += f * g[a];
        ds += f * gh3_dsq(T, &mt, &FN_DIP, dir);
        ds += f * gh3_dsq(T, &mt, &FN_QUAD, dir);
      }
    }
  }
  double Fp = s->x[i] * sp_d1(&T->embed[sslot(s->spc[i])], mf_rho[i]);
  for (int a = 0; a < 3; a++) out[a] = Fp * drho[a] + dq[a];
  out[3] = Fp * drho_s + ds;
}

/* evaluate_DV0_i_Dxi_u_adp_MgHx (:2102-2325) */
double or_ref_DV0_i_Dxi_u(const or_system* s, const or_tables* T, int u, int i,
                          const double* mf_rho) {
  if (s->x[u] < OR_MIN_OCCUPANCY) return 0.0;
  if (s->x[i] < OR_MIN_OCCUPANCY) return 0.0;
  double drho = 0.0, dv = 0.0;
  for (int k1 = 0; k1 < s->nnb[i]; k1++) {
    int j1 = NB(s, i, k1);
    if (s->x[j1] < OR_MIN_OCCUPANCY) continue;
    measure m;
    pair_measure(s, i, j1, &m);
    int sites2[2] = {i, j1};
    for (int dir = 0; dir < 2; dir++) {
      if (sites2[dir] != u) continue;
      drho += gh3_dxi(T, &m, &FN_RHO, dir);
      dv += gh3_dxi(T, &m, &FN_PAIR, dir);
    }
    for (int k2 = k1; k2 < s->nnb[i]; k2++) {
      int j2 = NB(s, i, k2);
      if (s->x[j2] < OR_MIN_OCCUPANCY) continue;
      double f = (k1 == k2) ? 1.0 : 2.0;
      int sites3[3] = {i, j1, j2};
      if (sites3[0] != u && sites3[1] != u && sites3[2] != u) continue;
      measure mt;
      trip_measure(s, i, j1, j2, &mt);
      for (int dir = 0; dir < 3; dir++) {
        if (sites3[dir] != u) continue;
        dv += f * gh3_dxi(T, &mt, &FN_DIP, dir);
        dv += f * gh3_dxi(T, &mt, &FN_QUAD, dir);
      }
    }
  }
  const or_spline* emb = &T->embed[sslot(s->spc[i])];
  double de = 0.0;
  if (s->spc[i] == SPC_H && i == u) de = sp_val(emb, mf_rho[i]);
  if (s->spc[i] == SPC_H && i != u) de = s->x[i] * sp_d1(emb, mf_rho[i]) * drho;
  if (s->spc[i] == SPC_MG) de = sp_d1(emb, mf_rho[i]) * drho; /* (:2311-2315) */
  return de + dv;
}

/* evaluate_D2V0_i_Dmeanq2_u_MgHx (:1366-1599): d2 V0_i / dq_u2 (3 x 3).  Only
 * the diagonal codes of each measure enter (2 dir + dir on pairs, 4 dir on
 * triplets), so on the coincident triplet (i, u, u) the two codes of u are
 * summed and the mixed codes 5, 7 are not; masks are the H-only ones of that
 * function (:1400, :1425, :1439 "<=", :1513 "<="). */
void or_ref_D2V0_i_Dq2_u(const or_system* s, const or_tables* T, int u, int i,
                         const double* mf_rho, double out[9]) {
  for (int a = 0; a < 9; a++) out[a] = 0.0;
  if (s->spc[u] == SPC_H && s->x[u] < OR_MIN_OCCUPANCY) return;
  if (s->spc[i] == SPC_H && s->x[i] < OR_MIN_OCCUPANCY) return;
  double drho[3] = {0, 0, 0}, d2rho[9] = {0}, hv[9] = {0}, g[3], h[9];
  for (int k1 = 0; k1 < s->nnb[i]; k1++) {
    int j1 = NB(s, i, k1);
    if (s->spc[j1] == SPC_H && s->x[j1] <= OR_MIN_OCCUPANCY) continue;
    measure m;
    pair_measure(s, i, j1, &m);
    int sites2[2] = {i, j1};
    for (int dir = 0; dir < 2; dir++) {
      if (sites2[dir] != u) continue;
      gh3_dmq(T, &m, &FN_RHO, dir, g);
      for (int a = 0; a < 3; a++) drho[a] += g[a];
      gh3_d2mq(T, &m, &FN_RHO, 3 * dir, h);
      for (int a = 0; a < 9; a++) d2rho[a] += h[a];
      gh3_d2mq(T, &m, &FN_PAIR, 3 * dir, h);
      for (int a = 0; a < 9; a++) hv[a] += h[a];
    }
    for (int k2 = k1; k2 < s->nnb[i]; k2++) {
      int j2 = NB(s, i, k2);
      if (s->spc[j2] == SPC_H && s->x[j2] <= OR_MIN_OCCUPANCY) continue;
      int sites3[3] = {i, j1, j2};
      if (sites3[0] != u && sites3[1] != u && sites3[2] != u) continue;
      double f = (k1 == k2) ? 1.0 : 2.0;
      measure mt;
      trip_measure(s, i, j1, j2, &mt);
      for (int dir = 0; dir < 3; dir++) {
        if (sites3[dir] != u) continue;
        gh3_d2mq(T, &mt, &FN_DIP, 4 * dir, h);
        for (int a = 0; a < 9; a++) hv[a] += f * h[a];
        gh3_d2mq(T, &mt, &FN_QUAD, 4 * dir, h);
        for (int a = 0; a < 9; a++) hv[a] += f * h[a];
      }
    }
  }
  const or_spline* emb = &T->embed[sslot(s->spc[i])];
  double xi = s->x[i], Fp = sp_d1(emb, mf_rho[i]), Fpp = sp_d2(emb, mf_rho[i]);
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++)
      out[3 * a + c] = xi * Fpp * drho[a] * drho[c] + xi * Fp * d2rho[3 * a + c] + hv[3 * a + c];
}

/* Diagonal blocks of the force-residual Jacobian as the reference's per-site
 * Hessians form them: D_u = sum_{i in {u} U N(u)} beta_i d2V0_i/dq_u2 (the
 * beta_i weights are those of the residual R_q(u) = sum_i beta_i dV0_i/dq_u).
 * No reference function forms this sum; a-21 is its per-(u, i) term. */
void or_ref_diag_blocks(const or_system* s, const or_tables* T, int nsites, const int* sites,
                        const double* mf_rho, double* D) {
  int cnt = sites ? nsites : s->n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < cnt; k++) {
    int u = sites ? sites[k] : k;
    double acc[9] = {0}, h[9];
    for (int c = -1; c < s->nnb[u]; c++) {
      int i = c < 0 ? u : NB(s, u, c);
      or_ref_D2V0_i_Dq2_u(s, T, u, i, mf_rho, h);
      for (int a = 0; a < 9; a++) acc[a] += s->beta[i] * h[a];
    }
    for (int a = 0; a < 9; a++) D[9 * (long)k + a] = acc[a];
  }
}

/* Zero-temperature gradient and Hessian: the functions of the mean-field
 * versions evaluated at the mean positions (no measure).
 * evaluate_DV_i_Dq_u_adp_MgHx (:147-329): H-only masks, "<" throughout
 * (:163, :185, :200, :254); evaluate_D2V_i_Dq2_u_MgHx (:333-526): masks on x
 * for any species, "<" for u and i (:350, :374), "<=" for j1, j2 (:387, :450),
 * diagonal codes only as in a-21.  rho = the zero-temperature density. */
static void pair_at(const or_system* s, int i, int j, double* q, double* n, int* spc) {
  for (int a = 0; a < 3; a++) q[a] = s->q[3 * i + a];
  img(s, i, j, q + 3);
  n[0] = s->x[i]; n[1] = s->x[j];
  spc[0] = s->spc[i]; spc[1] = s->spc[j];
}
static void trip_at(const or_system* s, int i, int j1, int j2, double* q, double* n, int* spc) {
  for (int a = 0; a < 3; a++) q[a] = s->q[3 * i + a];
  img(s, i, j1, q + 3);
  img(s, i, j2, q + 6);
  n[0] = s->x[i]; n[1] = s->x[j1]; n[2] = s->x[j2];
  spc[0] = s->spc[i]; spc[1] = s->spc[j1]; spc[2] = s->spc[j2];
}

void or_ref_DV_i_Dq_u(const or_system* s, const or_tables* T, int u, int i, const double* rho,
                      double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  if (s->spc[u] == SPC_H && s->x[u] < OR_MIN_OCCUPANCY) return;
  if (s->spc[i] == SPC_H && s->x[i] < OR_MIN_OCCUPANCY) return;
  double drho[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, g[3], q[9], n[3];
  int spc[3];
  for (int k1 = 0; k1 < s->nnb[i]; k1++) {
    int j1 = NB(s, i, k1);
    if (s->spc[j1] == SPC_H && s->x[j1] < OR_MIN_OCCUPANCY) continue;
    pair_at(s, i, j1, q, n, spc);
    int sites2[2] = {i, j1};
    for (int dir = 0; dir < 2; dir++) {
      if (sites2[dir] != u) continue;
      f_rho_dq(T, dir, g, n, q, spc);
      for (int a = 0; a < 3; a++) drho[a] += g[a];
      f_pair_dq(T, dir, g, n, q, spc);
      for (int a = 0; a < 3; a++) dv[a] += g[a];
    }
    for (int k2 = k1; k2 < s->nnb[i]; k2++) {
      int j2 = NB(s, i, k2);
      if (s->spc[j2] == SPC_H && s->x[j2] < OR_MIN_OCCUPANCY) continue;
      int sites3[3] = {i, j1, j2};
      double f = (k1 == k2) ? 1.0 : 2.0;
      trip_at(s, i, j1, j2, q, n, spc);
      for (int dir = 0; dir < 3; dir++) {
        if (sites3[dir] != u) continue;
        f_dip_dq(T, dir, g, n, q, spc);
        for (int a = 0; a < 3; a++) dv[a] += f * g[a];
        f_quad_dq(T, dir, g, n, q, spc);
        for (int a = 0; a < 3; a++) dv[a] += f * g[a];
      }
    }
  }
  double Fp = sp_d1(&T->embed[sslot(s->spc[i])], rho[i]);
  for (int a = 0; a < 3; a++) out[a] = s->x[i] * Fp * drho[a] + dv[a];
}

void or_ref_D2V_i_Dq2_u(const or_system* s, const or_tables* T, int u, int i, const double* rho,
                        double out[9]) {
  for (int a = 0; a < 9; a++) out[a] = 0.0;
  if (s->x[u] < OR_MIN_OCCUPANCY) return;
  if (s->x[i] < OR_MIN_OCCUPANCY) return;
  double drho[3] = {0, 0, 0}, d2rho[9] = {0}, hv[9] = {0}, g[3], h[9], q[9], n[3];
  int spc[3];
  for (int k1 = 0; k1 < s->nnb[i]; k1++) {
    int j1 = NB(s, i, k1);
    if (s->x[j1] <= OR_MIN_OCCUPANCY) continue;
    pair_at(s, i, j1, q, n, spc);
    int sites2[2] = {i, j1};
    for (int dir = 0; dir < 2; dir++) {
      if (sites2[dir] != u) continue;
      f_rho_dq(T, dir, g, n, q, spc);
      for (int a = 0; a < 3; a++) drho[a] += g[a];
      f_rho_d2q(T, 3 * dir, h, n, q, spc);
      for (int a = 0; a < 9; a++) d2rho[a] += h[a];
      f_pair_d2q(T, 3 * dir, h, n, q, spc);
      for (int a = 0; a < 9; a++) hv[a] += h[a];
    }
    for (int k2 = k1; k2 < s->nnb[i]; k2++) {
      int j2 = NB(s, i, k2);
      if (s->x[j2] <= OR_MIN_OCCUPANCY) continue;
      int sites3[3] = {i, j1, j2};
      if (sites3[0] != u && sites3[1] != u && sites3[2] != u) continue;
      double f = (k1 == k2) ? 1.0 : 2.0;
      trip_at(s, i, j1, j2, q, n, spc);
      for (int dir = 0; dir < 3; dir++) {
        if (sites3[dir] != u) continue;
        f_dip_d2q(T, 4 * dir, h, n, q, spc);
        for (int a = 0; a < 9; a++) hv[a] += f * h[a];
        f_quad_d2q(T, 4 * dir, h, n, q, spc);
        for (int a = 0; a < 9; a++) hv[a] += f * h[a];
      }
    }
  }
  const or_spline* emb = &T->embed[sslot(s->spc[i])];
  double xi = s->x[i], Fp = sp_d1(emb, rho[i]), Fpp = sp_d2(emb, rho[i]);
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++)
      out[3 * a + c] = xi * Fpp * drho[a] * drho[c] + xi * Fp * d2rho[3 * a + c] + hv[3 * a + c];
}

/* Zero-temperature force and diagonal blocks with the residual's weights:
 * F_u = sum_{i in {u} U N(u)} beta_i dV_i/dq_u (n x 3), D_u likewise (n x 9). */
void or_ref_zero_t(const or_system* s, const or_tables* T, int nsites, const int* sites,
                   const double* rho, double* F, double* D) {
  int cnt = sites ? nsites : s->n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < cnt; k++) {
    int u = sites ? sites[k] : k;
    double f[3] = {0, 0, 0}, d[9] = {0}, g[3], h[9];
    for (int c = -1; c < s->nnb[u]; c++) {
      int i = c < 0 ? u : NB(s, u, c);
      or_ref_DV_i_Dq_u(s, T, u, i, rho, g);
      or_ref_D2V_i_Dq2_u(s, T, u, i, rho, h);
      for (int a = 0; a < 3; a++) f[a] += s->beta[i] * g[a];
      for (int a = 0; a < 9; a++) d[a] += s->beta[i] * h[a];
    }
    for (int a = 0; a < 3; a++) F[3 * (long)k + a] = f[a];
    for (int a = 0; a < 9; a++) D[9 * (long)k + a] = d[a];
  }
}

/* d(-log Z0_u)/d sigma_u and d(-log Z0_u - gamma min(x,0.9999))/dx_u for the
 * site functional S0 (:1124-1139). */
static inline double ent_dsig(const or_system* s, int u) { return -3.0 / s->sig[u]; }
static inline double ent_dx(const or_system* s, int u) {
  double xu = s->x[u], v = -1.5 / xu;
  if (xu < 1.0) v -= 1.0 / (1.0 - xu);
  if (s->spc[u] == SPC_H && xu < 0.9999) v -= s->gamma[u];
  return v;
}

/* Residual of sum_i S0_i, composed from the per-(u, i) evaluators:
 * R_q(u) = sum_{i in {u} U N(u)} beta_i dV0_i/dq_u, R_sigma adds -3/sigma_u,
 * R_x adds the entropy derivative.  No reference function forms this sum. */
void or_ref_residual(const or_system* s, const or_tables* T, int nsites, const int* sites,
                     const double* mf_rho, double* Rq, double* Rs, double* Rx) {
  int cnt = sites ? nsites : s->n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < cnt; k++) {
    int u = sites ? sites[k] : k;
    double q[3] = {0, 0, 0}, sg = 0.0, xx = 0.0, o[4];
    if (s->x[u] >= OR_MIN_OCCUPANCY) {
      for (int c = -1; c < s->nnb[u]; c++) {
        int i = c < 0 ? u : NB(s, u, c);
        double b = s->beta[i];
        or_ref_DV0_i_Dq_u(s, T, u, i, mf_rho, o);
        for (int a = 0; a < 3; a++) q[a] += b * o[a];
        sg += b * o[3];
        xx += b * or_ref_DV0_i_Dxi_u(s, T, u, i, mf_rho);
      }
      sg += ent_dsig(s, u);
      xx += ent_dx(s, u);
    }
    for (int a = 0; a < 3; a++) Rq[3 * k + a] = q[a];
    Rs[k] = sg;
    Rx[k] = xx;
  }
}

/* ------------------------------------------------------------------------- */
/* Factorised O(nb) form                                                      */
/*                                                                            */
/* For site i the dipole and quadrupole terms are bilinear in the per-        */
/* neighbour feature psi_j(r) = n_j (u(|r|) r, w(|r|) r r^T), with the form   */
/* <X,Y> = 1/2 a.a' + 1/2 B:B' - 1/6 trB trB' (MgHx-ADP.cpp:1233-1281,        */
/* 1781-1831).  The reference sums over ordered j1 <= j2 with factor 2 off the */
/* diagonal (MgHx-mf-V-bulk.cpp:879-936).  Off the diagonal the 9D rule moves */
/* one site at a time by 3 sigma, so                                          */
/*   sum_{j1!=j2} Q9 = 1/18 [ sum_ls (<G_ls,G_ls> - sum_j <psi_j^ls,psi_j^ls>) */
/*                          + 2 <U, G0> - 2 sum_j <T_j, psi_j^0> ]            */
/* with G_ls = sum_j psi_j(r_j + 3 s sigma_i e_l), G0 = sum_j psi_j(r_j),      */
/* T_j = sum_ls psi_j(r_j + 3 s sigma_j e_l), U = sum_j T_j.  The diagonal   */
/* (coincident j1 = j2) uses the 6D rule with sqrt(6) sigma shifts of i or of */
/* j, weight 1/12 -- the same points as the pair and density integrals.       */
/* Feature vectors are stored as 9 doubles: a_x a_y a_z B_xx B_yy B_zz B_xy   */
/* B_xz B_yz.                                                                 */
/* ------------------------------------------------------------------------- */

typedef struct { double v[9]; } f9;

static inline double dotM(const double* x, const double* y) {
  double trx = x[3] + x[4] + x[5], try_ = y[3] + y[4] + y[5];
  return 0.5 * (x[0] * y[0] + x[1] * y[1] + x[2] * y[2]) +
         0.5 * (x[3] * y[3] + x[4] * y[4] + x[5] * y[5]) +
         (x[6] * y[6] + x[7] * y[7] + x[8] * y[8]) - (1.0 / 6.0) * trx * try_;
}
/* out = c * M x */
static inline void applyM(double c, const double* x, double* out) {
  double tr = x[3] + x[4] + x[5];
  out[0] = 0.5 * c * x[0]; out[1] = 0.5 * c * x[1]; out[2] = 0.5 * c * x[2];
  out[3] = c * (0.5 * x[3] - tr / 6.0);
  out[4] = c * (0.5 * x[4] - tr / 6.0);
  out[5] = c * (0.5 * x[5] - tr / 6.0);
  out[6] = c * x[6]; out[7] = c * x[7]; out[8] = c * x[8];
}
static inline double dot9(const double* x, const double* y) {
  double a = 0.0;
  for (int k = 0; k < 9; k++) a += x[k] * y[k];
  return a;
}

/* psi without the occupancy factor at point p for pair table slot t. */
static inline void psi_at(const or_tables* T, int t, const double* p, double* o) {
  double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  double u = sp_val(&T->u[t], r), w = sp_val(&T->w[t], r);
  o[0] = u * p[0]; o[1] = u * p[1]; o[2] = u * p[2];
  o[3] = w * p[0] * p[0]; o[4] = w * p[1] * p[1]; o[5] = w * p[2] * p[2];
  o[6] = w * p[0] * p[1]; o[7] = w * p[0] * p[2]; o[8] = w * p[1] * p[2];
}

/* Gradient in p of L . psi(p) (no occupancy factor), also returning psi(p). */
static inline void grad_at(const or_tables* T, int t, const double* p, const double* L,
                           double* g) {
  double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const or_spline *su = &T->u[t], *sw = &T->w[t];
  double u = sp_val(su, r), du = sp_d1(su, r), w = sp_val(sw, r), dw = sp_d1(sw, r);
  double la = L[0] * p[0] + L[1] * p[1] + L[2] * p[2];
  double gQ[3] = {2 * L[3] * p[0] + L[6] * p[1] + L[7] * p[2],
                  L[6] * p[0] + 2 * L[4] * p[1] + L[8] * p[2],
                  L[7] * p[0] + L[8] * p[1] + 2 * L[5] * p[2]};
  double Q = 0.5 * (p[0] * gQ[0] + p[1] * gQ[1] + p[2] * gQ[2]);
  double c = (du * la + dw * Q) / r;
  for (int a = 0; a < 3; a++) g[a] = c * p[a] + u * L[a] + w * gQ[a];
}

static inline int act(const or_system* s, int j) { return s->x[j] >= OR_MIN_OCCUPANCY; }

/* Pair (and coincident-triplet) rule of the factorised form.  Point t of the
 * pair (i, j): p = r_ij + delta, weight w, and dsi = d delta / d sigma_i.
 *   rule 3: the reference's 6D Stroud rule (Quadrature-Hermitian-3th.cpp:
 *           36-110): sqrt(6) sigma shifts of site i (t < 6) or of site j,
 *           one axis at a time, weight 1/12 each (12 points);
 *   rule 5: 3-point Gauss-Hermite per axis on the relative coordinate
 *           r ~ N(r_ij, (sigma_i^2 + sigma_j^2) I): nodes 0, +-sqrt(3)
 *           sigma_rel with weights 2/3, 1/6 (27 points, exact to degree 5);
 *   rule 44: the reference's 44-point 6D degree-5 rule (below);
 *   rule 6: the 6-point 3D Stroud rule on the relative coordinate
 *           r ~ N(r_ij, (sigma_i^2 + sigma_j^2) I): r +- sqrt(3) sigma_rel e_l,
 *           weight 1/6 (degree 3).  With sigma_i = sigma_j it is the
 *           reference's 12-point rule (both put r +- sqrt(6) sigma e_l). */
#define OR_SQRT3 1.7320508075688772
/* rule 44: moment_1th_Vij_gaussian_measure_5th (Quadrature-Hermitian-5th.cpp:
 * 31-157), the 6D degree-5 rule with 22 generators in zeta units and their
 * negatives: (1,...,1) weight 1/128; sqrt(2) e_k weight 1/16; (1,...,1) with
 * -1 at positions a < b (lexicographic) weight 1/128; site displacement
 * sqrt(2) sigma zeta.  The reference scales the weights by pi^(3/2) and the
 * sum by pi^(-3/2) (:50, :156); here they cancel exactly. */
static inline double rule44_gen(int t, double* z) {
  const int g = t % 22;
  const double sg = t < 22 ? 1.0 : -1.0;
  if (g == 0) {
    for (int k = 0; k < 6; k++) z[k] = sg;
    return 1.0 / 128.0;
  }
  if (g <= 6) {
    for (int k = 0; k < 6; k++) z[k] = (k == g - 1) ? sg * OR_SQRT2 : 0.0;
    return 1.0 / 16.0;
  }
  int idx = g - 7, a = 0;
  while (idx >= 5 - a) {
    idx -= 5 - a;
    a++;
  }
  const int b = a + 1 + idx;
  for (int k = 0; k < 6; k++) z[k] = (k == a || k == b) ? -sg : sg;
  return 1.0 / 128.0;
}
static inline int pair_rule_points(int rule) {
  return rule == 5 ? 27 : rule == 44 ? 44 : rule == 6 ? 6 : 12;
}
static inline double pair_rule_point(int rule, int t, const double* r, double si, double sj,
                                     double* p, double* dsi) {
  p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
  if (rule == 44) {
    double z[6];
    const double w = rule44_gen(t, z);
    for (int k = 0; k < 3; k++) {
      p[k] += OR_SQRT2 * (si * z[k] - sj * z[3 + k]);
      dsi[k] = OR_SQRT2 * z[k];
    }
    return w;
  }
  if (rule == 5) {
    static const double node[3] = {0.0, OR_SQRT3, -OR_SQRT3};
    static const double wn[3] = {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0};
    const int a[3] = {t % 3, (t / 3) % 3, t / 9};
    const double srel = sqrt(si * si + sj * sj);
    double w = 1.0;
    for (int k = 0; k < 3; k++) {
      const double xi = node[a[k]];
      p[k] += srel * xi;
      dsi[k] = si / srel * xi;
      w *= wn[a[k]];
    }
    return w;
  }
  if (rule == 6) {
    const int l = t >> 1;
    const double sg = (t & 1) ? -OR_SQRT3 : OR_SQRT3;
    const double srel = sqrt(si * si + sj * sj);
    p[l] += sg * srel;
    dsi[0] = dsi[1] = dsi[2] = 0.0;
    dsi[l] = sg * si / srel;
    return 1.0 / 6.0;
  }
  const int wsel = t / 6, l = (t >> 1) % 3;
  const double d = (t & 1) ? -OR_SQRT6 : OR_SQRT6;
  p[l] += d * (wsel ? sj : si);
  dsi[0] = dsi[1] = dsi[2] = 0.0;
  if (!wsel) dsi[l] = d;
  return 1.0 / 12.0;
}

/* Phase-1 record layout (OR_P1_STRIDE doubles per site). */
enum { P1_LA = 0, P1_LU = 54, P1_LG0 = 63, P1_FP = 72, P1_DVDN = 73, P1_RHO = 74 };

static void phase1_site(const or_system* s, const or_tables* T, int i, double* rho_out,
                        double* V0_out, double* rec) {
  memset(rec, 0, sizeof(double) * OR_P1_STRIDE);
  if (!act(s, i)) { *rho_out = 0.0; *V0_out = 0.0; return; }
  const double* qi = s->q + 3 * i;
  double si = s->sig[i], ni = s->x[i];
  int spi = s->spc[i];
  double rho = 0, Ep = 0, C = 0, selfA = 0, TPsi = 0;
  double G0[9] = {0}, U[9] = {0}, GA[6][9];
  memset(GA, 0, sizeof(GA));
  for (int k = 0; k < s->nnb[i]; k++) {
    int j = NB(s, i, k);
    if (!act(s, j)) continue;
    double nj = s->x[j], sj = s->sig[j], r[3], p[3], ps[9], T9[9] = {0}, psi0[9];
    int t = pslot(spi, s->spc[j]);
    const or_spline* srho = &T->rho[sslot(s->spc[j])];
    const or_spline* sphi = &T->pair[t];
    sep(s->lo, s->hi, s->periodic, qi, s->q + 3 * j, r);
    psi_at(T, t, r, psi0);
    for (int c = 0; c < 9; c++) G0[c] += nj * psi0[c];
    for (int l = 0; l < 3; l++)
      for (int sg = 0; sg < 2; sg++) {
        double sh = sg ? -3.0 : 3.0;
        p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
        p[l] += sh * si;
        psi_at(T, t, p, ps);
        for (int c = 0; c < 9; c++) GA[2 * l + sg][c] += nj * ps[c];
        selfA += nj * nj * dotM(ps, ps);
        p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
        p[l] += sh * sj;
        psi_at(T, t, p, ps);
        for (int c = 0; c < 9; c++) T9[c] += ps[c];
      }
    for (int c = 0; c < 9; c++) U[c] += nj * T9[c];
    TPsi += nj * nj * dotM(T9, psi0);
    double Phi = 0, R = 0, Cp = 0, dsi[3];
    for (int pt = 0; pt < pair_rule_points(s->rule); pt++) {
      const double wt = pair_rule_point(s->rule, pt, r, si, sj, p, dsi);
      psi_at(T, t, p, ps);
      Cp += wt * dotM(ps, ps);
      double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      Phi += wt * sp_val(sphi, pr);
      R += wt * sp_val(srho, pr);
    }
    C += nj * nj * Cp;
    rho += nj * R;
    Ep += nj * Phi;
  }
  double GG = 0;
  for (int ls = 0; ls < 6; ls++) GG += dotM(GA[ls], GA[ls]);
  double At = (GG - selfA + 2.0 * dotM(U, G0) - 2.0 * TPsi) / 18.0 + C;
  const or_spline* emb = &T->embed[sslot(spi)];
  double F = sp_val(emb, rho);
  *rho_out = rho;
  *V0_out = ni * F + 0.5 * ni * Ep + ni * ni * At;
  for (int ls = 0; ls < 6; ls++) applyM(1.0 / 9.0, GA[ls], rec + P1_LA + 9 * ls);
  applyM(1.0 / 9.0, U, rec + P1_LU);
  applyM(1.0 / 9.0, G0, rec + P1_LG0);
  rec[P1_FP] = sp_d1(emb, rho);
  rec[P1_DVDN] = F + 0.5 * Ep + 2.0 * ni * At;
  rec[P1_RHO] = rho;
}

void or_fact_phase1(const or_system* s, const or_tables* T, int nthreads, double* mf_rho,
                    double* V0, double* p1) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < s->n; i++)
    phase1_site(s, T, i, mf_rho + i, V0 + i, p1 + (long)i * OR_P1_STRIDE);
}

/* flip(v): negate the dipole (odd) part, used for psi(-p) = flip(psi(p)). */
static inline void flip(const double* v, double* o) {
  o[0] = -v[0]; o[1] = -v[1]; o[2] = -v[2];
  for (int c = 3; c < 9; c++) o[c] = v[c];
}

/* Phase 2 for site u: gradient of sum_c beta_c V0_c with respect to q_u,
 * sigma_u and x_u.  Role (a) is V0_u itself (u is the centre), role (b) is
 * V0_k of each neighbour k (u is a neighbour).  Both roles use the same 25
 * points p = r_uk + shift; the adjoints are combined before one gradient. */
static void phase2_site(const or_system* s, const or_tables* T, const double* p1, int u,
                        double* gq, double* gs, double* gx) {
  gq[0] = gq[1] = gq[2] = 0.0; *gs = 0.0; *gx = 0.0;
  if (!act(s, u)) return;
  const double* qu = s->q + 3 * u;
  const double* Ru = p1 + (long)u * OR_P1_STRIDE;
  double su = s->sig[u], nu = s->x[u], bu = s->beta[u];
  int spu = s->spc[u];
  double q0 = 0, q1 = 0, q2 = 0, sacc = 0, xacc = 0;
  for (int kk = 0; kk < s->nnb[u]; kk++) {
    int k = NB(s, u, kk);
    if (!act(s, k)) continue;
    const double* Rk = p1 + (long)k * OR_P1_STRIDE;
    double nk = s->x[k], sk = s->sig[k], bk = s->beta[k];
    double wa = bu * nu * nu * nk, wb = bk * nk * nk * nu;
    int t = pslot(spu, s->spc[k]);
    double r[3], p[3], ps[9], psi0[9], Tuk[9] = {0}, SA[9] = {0}, la[9], lb[9], L[9], g[3], tmp[9];
    sep(s->lo, s->hi, s->periodic, qu, s->q + 3 * k, r);
    psi_at(T, t, r, psi0);
    /* j-shift points of centre u (shift sigma_k) = i-shift points (l,-s) of k */
    for (int l = 0; l < 3; l++)
      for (int sg = 0; sg < 2; sg++) {
        double sh = sg ? -3.0 : 3.0;
        p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
        p[l] += sh * sk;
        psi_at(T, t, p, ps);
        for (int c = 0; c < 9; c++) Tuk[c] += ps[c];
        applyM(-nk / 9.0, psi0, tmp);
        for (int c = 0; c < 9; c++) la[c] = Ru[P1_LG0 + c] + tmp[c];
        flip(Rk + P1_LA + 9 * (2 * l + (1 - sg)), lb);
        applyM(-nu / 9.0, ps, tmp);
        for (int c = 0; c < 9; c++) lb[c] += tmp[c];
        for (int c = 0; c < 9; c++) L[c] = wa * la[c] + wb * lb[c];
        grad_at(T, t, p, L, g);
        q0 += g[0]; q1 += g[1]; q2 += g[2];
        xacc += bk * nk * nk * dot9(lb, ps);
      }
    /* i-shift points of centre u (shift sigma_u) = j-shift points of k */
    for (int l = 0; l < 3; l++)
      for (int sg = 0; sg < 2; sg++) {
        double sh = sg ? -3.0 : 3.0;
        p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
        p[l] += sh * su;
        psi_at(T, t, p, ps);
        for (int c = 0; c < 9; c++) SA[c] += ps[c];
        applyM(-nk / 9.0, ps, tmp);
        for (int c = 0; c < 9; c++) la[c] = Ru[P1_LA + 9 * (2 * l + sg) + c] + tmp[c];
        flip(Rk + P1_LG0, lb);
        applyM(-nu / 9.0, psi0, tmp);
        for (int c = 0; c < 9; c++) lb[c] += tmp[c];
        for (int c = 0; c < 9; c++) L[c] = wa * la[c] + wb * lb[c];
        grad_at(T, t, p, L, g);
        q0 += g[0]; q1 += g[1]; q2 += g[2];
        sacc += sh * g[l];
        xacc += bk * nk * nk * dot9(lb, ps);
      }
    /* unshifted point */
    applyM(-nk / 9.0, Tuk, tmp);
    for (int c = 0; c < 9; c++) la[c] = Ru[P1_LU + c] + tmp[c];
    flip(Rk + P1_LU, lb);
    applyM(-nu / 9.0, SA, tmp);
    for (int c = 0; c < 9; c++) lb[c] += tmp[c];
    for (int c = 0; c < 9; c++) L[c] = wa * la[c] + wb * lb[c];
    grad_at(T, t, r, L, g);
    q0 += g[0]; q1 += g[1]; q2 += g[2];
    xacc += bk * nk * nk * dot9(lb, psi0);
    /* sqrt(6) sigma points: coincident angular term, pair and density */
    const or_spline* sphi = &T->pair[t];
    const or_spline* srk = &T->rho[sslot(s->spc[k])];
    const or_spline* sru = &T->rho[sslot(spu)];
    double cphi = 0.5 * nu * nk * (bu + bk);
    double cra = bu * nu * Ru[P1_FP] * nk;
    double crb = bk * nk * Rk[P1_FP] * nu;
    double cang = 2.0 * (wa * nk + wb * nu);
    double dsu[3];
    for (int pt = 0; pt < pair_rule_points(s->rule); pt++) {
      const double wt = pair_rule_point(s->rule, pt, r, su, sk, p, dsu);
      psi_at(T, t, p, ps);
      applyM(wt * cang, ps, L);
      grad_at(T, t, p, L, g);
      double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      double cs = wt * (cphi * sp_d1(sphi, pr) + cra * sp_d1(srk, pr) + crb * sp_d1(sru, pr)) / pr;
      for (int a = 0; a < 3; a++) g[a] += cs * p[a];
      q0 += g[0]; q1 += g[1]; q2 += g[2];
      sacc += g[0] * dsu[0] + g[1] * dsu[1] + g[2] * dsu[2];
      xacc += wt * (bk * nk * nk * (2.0 * nu) * dotM(ps, ps) +
                    bk * (nk * Rk[P1_FP] * sp_val(sru, pr) + 0.5 * nk * sp_val(sphi, pr)));
    }
  }
  gq[0] = q0; gq[1] = q1; gq[2] = q2;
  *gs = sacc;
  *gx = bu * Ru[P1_DVDN] + xacc;
}

void or_fact_residual(const or_system* s, const or_tables* T, int nthreads, double* Rq,
                      double* Rs, double* Rx, double* S0, double* V0, double* mf_rho,
                      double* totals) {
  double* p1 = (double*)malloc(sizeof(double) * OR_P1_STRIDE * (size_t)s->n);
  or_fact_phase1(s, T, nthreads, mf_rho, V0, p1);
  double tS = 0.0, tV = 0.0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : tS, tV)
  for (int u = 0; u < s->n; u++) {
    double gq[3], gs, gx;
    phase2_site(s, T, p1, u, gq, &gs, &gx);
    if (act(s, u)) {
      double xu = s->x[u], b = s->beta[u];
      double mi = OR_UMA * xu * T->mass[sslot(s->spc[u])];
      double logZ = 3.0 * log((s->sig[u] * sqrt(mi / b)) / OR_HBAR);
      if (xu < 1.0) logZ += log(1.0 / (1.0 - xu));
      double gxx = s->spc[u] == SPC_H ? s->gamma[u] * (xu > 0.9999 ? 0.9999 : xu) : 0.0;
      S0[u] = -logZ + b * (1.0 / (2.0 * b) + V0[u]) - gxx;
      Rq[3 * u] = gq[0]; Rq[3 * u + 1] = gq[1]; Rq[3 * u + 2] = gq[2];
      Rs[u] = gs + ent_dsig(s, u);
      Rx[u] = gx + ent_dx(s, u);
      tS += S0[u];
      tV += V0[u];
    } else {
      S0[u] = 0.0;
      Rq[3 * u] = Rq[3 * u + 1] = Rq[3 * u + 2] = 0.0;
      Rs[u] = 0.0;
      Rx[u] = 0.0;
    }
  }
  totals[0] = tS;
  totals[1] = tV;
  free(p1);
}

/* ------------------------------------------------------------------------- */
/* Factorised diagonal Hessian blocks                                         */
/*                                                                            */
/* d2/dq_u2 of sum_c beta_c V0_c.  Every term that is linear in one feature   */
/* psi(p) has the form L . psi(p) with the same adjoint L as the gradient in  */
/* phase2_site, so its Hessian is sum_a L_a d2psi_a/dp2 (hess_at).  What is  */
/* quadratic in features of the same centre u adds J^T M J terms, J = dpsi/dp */
/* summed over the neighbours: 1/9 [sum_ls JG_ls^T M JG_ls + JU^T M JG0 +     */
/* JG0^T M JU], less the j1 = j2 parts of those sums, plus 2 n_k^2 J^T M J on */
/* the coincident points.  For a neighbour centre k, only psi_u of k moves,   */
/* and the coincident triplet (k, u, u) contributes codes 4 + 8 only (no J^T  */
/* M J), as evaluate_D2V0_i_Dmeanq2_u_MgHx sums it (:1540-1570); exact = 1    */
/* adds its mixed codes 5 + 7 (2 n_k^2 n_u^2 J^T M J), which makes D_u the    */
/* exact diagonal block of the residual's Jacobian.  Embedding:               */
/* x F''(rho) g g^T + x F'(rho) d2rho, with g the density gradient of the     */
/* centre with respect to q_u.                                                 */
/* ------------------------------------------------------------------------- */

/* J[3 a + g] = d psi_a / dp_g (no occupancy factor). */
static inline void jac_at(const or_tables* T, int t, const double* p, double* J) {
  static const int qi[6] = {0, 1, 2, 0, 0, 1}, qj[6] = {0, 1, 2, 1, 2, 2};
  double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  double u = sp_val(&T->u[t], r), du = sp_d1(&T->u[t], r) / r;
  double w = sp_val(&T->w[t], r), dw = sp_d1(&T->w[t], r) / r;
  for (int a = 0; a < 3; a++)
    for (int g = 0; g < 3; g++) J[3 * a + g] = du * p[a] * p[g] + (a == g ? u : 0.0);
  for (int m = 0; m < 6; m++) {
    int i = qi[m], j = qj[m];
    for (int g = 0; g < 3; g++)
      J[3 * (3 + m) + g] = dw * p[g] * p[i] * p[j] + w * ((i == g ? p[j] : 0.0) + (j == g ? p[i] : 0.0));
  }
}

/* H += k sum_a L_a d2psi_a/dp2.  L . psi = u(r) (La . p) + w(r) Q(p), Q the
 * quadratic form of L[3..8]. */
static inline void hess_at(const or_tables* T, int t, const double* p, const double* L, double k,
                           double* H) {
  double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const or_spline *su = &T->u[t], *sw = &T->w[t];
  double u1 = sp_d1(su, r), u2 = sp_d2(su, r), w = sp_val(sw, r), w1 = sp_d1(sw, r),
         w2 = sp_d2(sw, r);
  double la = L[0] * p[0] + L[1] * p[1] + L[2] * p[2];
  double HQ[9] = {2 * L[3], L[6], L[7], L[6], 2 * L[4], L[8], L[7], L[8], 2 * L[5]};
  double gQ[3];
  for (int a = 0; a < 3; a++) gQ[a] = HQ[3 * a] * p[0] + HQ[3 * a + 1] * p[1] + HQ[3 * a + 2] * p[2];
  double Q = 0.5 * (p[0] * gQ[0] + p[1] * gQ[1] + p[2] * gQ[2]);
  double R[9] = {0};
  rad_hess(u1, u2, p, r, la, R);
  rad_hess(w1, w2, p, r, Q, R);
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++)
      H[3 * a + c] += k * (R[3 * a + c] + u1 / r * (p[a] * L[c] + L[a] * p[c]) +
                           w1 / r * (p[a] * gQ[c] + gQ[a] * p[c]) + w * HQ[3 * a + c]);
}

/* H += k A^T M B for 9 x 3 Jacobians. */
static inline void jtmj(const double* A, const double* B, double k, double* H) {
  for (int c = 0; c < 3; c++) {
    double col[9], mc[9];
    for (int a = 0; a < 9; a++) col[a] = B[3 * a + c];
    applyM(1.0, col, mc);
    for (int b = 0; b < 3; b++) {
      double acc = 0.0;
      for (int a = 0; a < 9; a++) acc += A[3 * a + b] * mc[a];
      H[3 * b + c] += k * acc;
    }
  }
}

static void diag_site(const or_system* s, const or_tables* T, const double* p1, int u, int exact,
                      double* D) {
  for (int a = 0; a < 9; a++) D[a] = 0.0;
  if (!act(s, u)) return;
  const double* qu = s->q + 3 * u;
  const double* Ru = p1 + (long)u * OR_P1_STRIDE;
  double su = s->sig[u], nu = s->x[u], bu = s->beta[u];
  int spu = s->spc[u];
  const or_spline* sru = &T->rho[sslot(spu)];
  double H[9] = {0}, HJ[9] = {0}, gu[3] = {0, 0, 0};
  double JG[6][27], JU[27] = {0}, JG0[27] = {0};
  memset(JG, 0, sizeof(JG));
  for (int kk = 0; kk < s->nnb[u]; kk++) {
    int k = NB(s, u, kk);
    if (!act(s, k)) continue;
    const double* Rk = p1 + (long)k * OR_P1_STRIDE;
    double nk = s->x[k], sk = s->sig[k], bk = s->beta[k];
    double wa = bu * nu * nu * nk, wb = bk * nk * nk * nu;
    int t = pslot(spu, s->spc[k]);
    double r[3], p[3], ps[9], psi0[9], Tuk[9] = {0}, SA[9] = {0}, la[9], lb[9], L[9], tmp[9];
    double J0[27], J[27], JT[27] = {0};
    sep(s->lo, s->hi, s->periodic, qu, s->q + 3 * k, r);
    psi_at(T, t, r, psi0);
    jac_at(T, t, r, J0);
    for (int c = 0; c < 27; c++) JG0[c] += nk * J0[c];
    /* j-shift points of centre u (shift sigma_k) = i-shift points (l,-s) of k */
    for (int l = 0; l < 3; l++)
      for (int sg = 0; sg < 2; sg++) {
        double sh = sg ? -3.0 : 3.0;
        p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
        p[l] += sh * sk;
        psi_at(T, t, p, ps);
        jac_at(T, t, p, J);
        for (int c = 0; c < 9; c++) Tuk[c] += ps[c];
        for (int c = 0; c < 27; c++) JT[c] += J[c];
        applyM(-nk / 9.0, psi0, tmp);
        for (int c = 0; c < 9; c++) la[c] = Ru[P1_LG0 + c] + tmp[c];
        flip(Rk + P1_LA + 9 * (2 * l + (1 - sg)), lb);
        applyM(-nu / 9.0, ps, tmp);
        for (int c = 0; c < 9; c++) lb[c] += tmp[c];
        for (int c = 0; c < 9; c++) L[c] = wa * la[c] + wb * lb[c];
        hess_at(T, t, p, L, 1.0, H);
      }
    for (int c = 0; c < 27; c++) JU[c] += nk * JT[c];
    jtmj(JT, J0, -nk * nk / 9.0, HJ);
    jtmj(J0, JT, -nk * nk / 9.0, HJ);
    /* i-shift points of centre u (shift sigma_u) = j-shift points of k */
    for (int l = 0; l < 3; l++)
      for (int sg = 0; sg < 2; sg++) {
        double sh = sg ? -3.0 : 3.0;
        p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
        p[l] += sh * su;
        psi_at(T, t, p, ps);
        jac_at(T, t, p, J);
        for (int c = 0; c < 9; c++) SA[c] += ps[c];
        for (int c = 0; c < 27; c++) JG[2 * l + sg][c] += nk * J[c];
        jtmj(J, J, -nk * nk / 9.0, HJ);
        applyM(-nk / 9.0, ps, tmp);
        for (int c = 0; c < 9; c++) la[c] = Ru[P1_LA + 9 * (2 * l + sg) + c] + tmp[c];
        flip(Rk + P1_LG0, lb);
        applyM(-nu / 9.0, psi0, tmp);
        for (int c = 0; c < 9; c++) lb[c] += tmp[c];
        for (int c = 0; c < 9; c++) L[c] = wa * la[c] + wb * lb[c];
        hess_at(T, t, p, L, 1.0, H);
      }
    /* unshifted point */
    applyM(-nk / 9.0, Tuk, tmp);
    for (int c = 0; c < 9; c++) la[c] = Ru[P1_LU + c] + tmp[c];
    flip(Rk + P1_LU, lb);
    applyM(-nu / 9.0, SA, tmp);
    for (int c = 0; c < 9; c++) lb[c] += tmp[c];
    for (int c = 0; c < 9; c++) L[c] = wa * la[c] + wb * lb[c];
    hess_at(T, t, r, L, 1.0, H);
    /* pair-rule points: coincident angular term, pair and density */
    const or_spline* sphi = &T->pair[t];
    const or_spline* srk = &T->rho[sslot(s->spc[k])];
    const or_spline* emk = &T->embed[sslot(s->spc[k])];
    double cphi = 0.5 * nu * nk * (bu + bk);
    double cra = bu * nu * Ru[P1_FP] * nk;
    double crb = bk * nk * Rk[P1_FP] * nu;
    double cang = 2.0 * (wa * nk + wb * nu);
    double gk[3] = {0, 0, 0}, dsu[3];
    for (int pt = 0; pt < pair_rule_points(s->rule); pt++) {
      const double wt = pair_rule_point(s->rule, pt, r, su, sk, p, dsu);
      psi_at(T, t, p, ps);
      jac_at(T, t, p, J);
      applyM(wt * cang, ps, L);
      hess_at(T, t, p, L, 1.0, H);
      jtmj(J, J, 2.0 * wt * nk * nk, HJ);
      if (exact) jtmj(J, J, 2.0 * wt * bk * nk * nk * nu * nu, H); /* codes 5 + 7 of (k, u, u) */
      double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      rad_hess(sp_d1(sphi, pr), sp_d2(sphi, pr), p, pr, wt * cphi, H);
      rad_hess(sp_d1(srk, pr), sp_d2(srk, pr), p, pr, wt * cra, H);
      rad_hess(sp_d1(sru, pr), sp_d2(sru, pr), p, pr, wt * crb, H);
      double a1 = wt * nk * sp_d1(srk, pr) / pr, b1 = wt * nu * sp_d1(sru, pr) / pr;
      for (int a = 0; a < 3; a++) { gu[a] += a1 * p[a]; gk[a] += b1 * p[a]; }
    }
    double ck = bk * nk * sp_d2(emk, Rk[P1_RHO]);
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) H[3 * a + c] += ck * gk[a] * gk[c];
  }
  for (int ls = 0; ls < 6; ls++) jtmj(JG[ls], JG[ls], 1.0 / 9.0, HJ);
  jtmj(JU, JG0, 1.0 / 9.0, HJ);
  jtmj(JG0, JU, 1.0 / 9.0, HJ);
  double cu = bu * nu * sp_d2(&T->embed[sslot(spu)], Ru[P1_RHO]);
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++)
      D[3 * a + c] = H[3 * a + c] + bu * nu * nu * HJ[3 * a + c] + cu * gu[a] * gu[c];
}

void or_fact_diag_blocks(const or_system* s, const or_tables* T, int nthreads, int exact,
                         double* D) {
  double* p1 = (double*)malloc(sizeof(double) * OR_P1_STRIDE * (size_t)s->n);
  double* rho = (double*)malloc(sizeof(double) * (size_t)s->n);
  double* V0 = (double*)malloc(sizeof(double) * (size_t)s->n);
  or_fact_phase1(s, T, nthreads, rho, V0, p1);
#pragma omp parallel for schedule(dynamic, 16)
  for (int u = 0; u < s->n; u++) diag_site(s, T, p1, u, exact, D + 9 * (long)u);
  free(p1);
  free(rho);
  free(V0);
}

/* ------------------------------------------------------------------------- */
/* Diffusion step                                                             */
/* ------------------------------------------------------------------------- */

int or_diff_neighbors(int n, const double* q, const int* spc, const double* lo,
                      const double* hi, const int* periodic, double rc, int cap, int* cnt,
                      int* lst) {
  int over = 0, empty = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : over, empty)
  for (int i = 0; i < n; i++) {
    int c = 0;
    if (spc[i] == 1) {
      for (int j = 0; j < n; j++) {
        if (j == i || spc[j] != 1) continue;
        double d[3];
        sep(lo, hi, periodic, q + 3 * i, q + 3 * j, d);
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (r <= rc) {
          if (c == cap) { over = 1; break; }
          lst[(long)i * cap + c] = j;
          c++;
        }
      }
      if (c == 0) empty = 1;
    }
    cnt[i] = c;
  }
  return over ? 1 : (empty ? 2 : 0);
}

void or_diff_mu(int n, const double* x, const int* spc, const double* gamma, const double* Rx,
                double* mu) {
  for (int i = 0; i < n; i++) {
    if (!(x[i] >= OR_MIN_OCCUPANCY)) {
      mu[i] = 0.0;
      continue;
    }
    double own = -1.5 / x[i];
    if (x[i] < 1.0) own -= 1.0 / (1.0 - x[i]);
    if (spc[i] == 1 && x[i] < 0.9999) own -= gamma[i];
    mu[i] = Rx[i] - own;
  }
}

/* Logarithmic mean of x over l(x) = ln(x / (1 - x)): (b - a) / (l(b) - l(a)),
 * a (1 - a) in the limit b -> a; 0 unless both sites are active (act) and
 * 0 < x < 1, so that a masked site exchanges by the Fick term only. */
static inline double log_mean_mobility(double a, double b, int act) {
  if (!act || !(a > 0.0 && a < 1.0 && b > 0.0 && b < 1.0)) return 0.0;
  const double d = log(b / (1.0 - b)) - log(a / (1.0 - a));
  if (fabs(d) < 1e-8) {
    const double m = 0.5 * (a + b);
    return m * (1.0 - m);
  }
  return (b - a) / d;
}

void or_diff_rates(int n, const double* q, const double* x, const double* xa, const double* mu,
                   const double* beta, const double* lo, const double* hi, const int* periodic,
                   const int* cnt, const int* lst, int cap, double d0, double eb, double* xdot) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double f = 0.0;
    for (int k = 0; k < cnt[i]; k++) {
      const int j = lst[(long)i * cap + k];
      double d[3];
      sep(lo, hi, periodic, q + 3 * i, q + 3 * j, d);
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double g = d0 / r2 * exp(-(0.5 * (beta[i] + beta[j])) * eb);
      const int act = xa[i] >= OR_MIN_OCCUPANCY && xa[j] >= OR_MIN_OCCUPANCY;
      f += g * ((x[j] - x[i]) + log_mean_mobility(x[i], x[j], act) * (mu[j] - mu[i]));
    }
    xdot[i] = f;
  }
}
