/*
 * dmd_oracle.h -- CPU FP64 restatement of the BT-AM mean-field DMD force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product library (libdmdgpu.so) never links or
 * calls it.
 *
 * Two independent restatements of the same mathematics are provided:
 *
 *  - "reference order" (or_ref_*): the per-site evaluators of
 *    src/ADP/MgHx-mf-V-bulk.cpp restated loop for loop.  They use the generic
 *    Gaussian measure of src/Numerical/Quadrature-Measure.cpp:30-110, the 3rd
 *    order Stroud rule of src/Numerical/Quadrature-Hermitian-3th.cpp:36-464 and
 *    the pair / triplet ADP functions of src/ADP/MgHx-ADP.cpp.  The cost is
 *    O(nb^2) per site and O(nb^3) per residual entry, so this form is used only
 *    on small systems.
 *
 *  - "factorised" (or_fact_*): the O(nb) two-phase form that the GPU kernels
 *    implement (per-atom dipole / quadrupole sums, then a gather of adjoints).
 *    It is the CPU baseline that is timed.  Tests pin it against the
 *    reference-order form and against finite differences of the energy.
 *
 * Species codes follow enum AtomicSpecie (src/Macros.hpp:56-175): H = 1, Mg = 12.
 */
#ifndef DMD_ORACLE_H
#define DMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXNEIGH 250          /* src/Macros.hpp:862 */
#define OR_MIN_OCCUPANCY 5e-4    /* src/Macros.hpp:871 */

/* Cubic spline with n segments of width dx (src/Macros.hpp:286-329). */
typedef struct {
  int n;
  double dx;
  double *a, *b, *c, *d, *db, *dc, *dd, *ddc, *ddd;
} or_spline;

/* ADP tables for the Mg-H system (src/Macros.hpp:486-533).
 * Species slot 0 = Mg, 1 = H.  Pair slot 0 = MgMg, 1 = HH, 2 = MgH. */
typedef struct {
  or_spline rho[2], embed[2];
  or_spline pair[3], u[3], w[3];
  double mass[2];
} or_tables;

/* Atom arrays (DMSwarm-equivalent SoA) plus the neighbour lists. */
typedef struct {
  int n;
  const double* q;      /* n x 3 mean positions            */
  const double* sig;    /* n     widths (stdv-q)             */
  const double* x;      /* n     occupancies (molar-fraction) */
  const int* spc;       /* n     AtomicSpecie codes          */
  const double* beta;   /* n     1 / (k_B T)                 */
  const double* gamma;  /* n     chemical multiplier         */
  double lo[3], hi[3];
  int periodic[3];
  int cap;              /* row stride of nbr                 */
  const int* nnb;       /* n                                 */
  const int* nbr;       /* n x cap                           */
  int rule;             /* pair/coincident rule of the factorised form:
                           3 (default, 0 too) = the reference's 12-point 6D
                           Stroud rule; 5 = 27-point tensor Gauss-Hermite on
                           the relative coordinate (degree 5); 44 = the
                           reference's 44-point 6D degree-5 pair rule;
                           6 = 6-point Stroud rule on the relative
                           coordinate (degree 3)                          */
} or_system;

/* Tables ------------------------------------------------------------------ */
int or_load_tables(or_tables** out, const char* folder);
void or_free_tables(or_tables* t);
/* which: 0 rho, 1 embed, 2 pair, 3 u, 4 w; slot as above; deriv 0/1/2. */
double or_spline_eval(const or_tables* t, int which, int slot, int deriv, double x);
int or_spline_size(const or_tables* t, int which, int slot, double* dx);

/* One pair (which 0 rho, 1 pair) or triplet (2 dipole, 3 quadrupole) function
 * at explicit positions q (3 per site) and occupancies n; slots 0 Mg / 1 H.
 * deriv 0: value; 1: gradient with respect to site `code`; 2: Hessian block
 * `code` = ns * a + b.  For tests. */
int or_fn_eval(const or_tables* t, int which, int deriv, int code, const double* n,
               const double* q, const int* slots, double* out);

/* Neighbours (brute force, src/Atoms/Neighbors.cpp:39-134) ----------------- */
int or_neighbors(int n, const double* q, const double* lo, const double* hi,
                 const int* periodic, double rc, int cap, int* nnb, int* nbr);

/* Reference order (src/ADP/MgHx-mf-V-bulk.cpp) ------------------------------ */
double or_ref_rho0_i(const or_system* s, const or_tables* t, int i);
double or_ref_mf_rho_i(const or_system* s, const or_tables* t, int i);
double or_ref_V0_i(const or_system* s, const or_tables* t, int i, const double* mf_rho);
double or_ref_S0_i(const or_system* s, const or_tables* t, int i, const double* mf_rho);
void or_ref_DV0_i_Dq_u(const or_system* s, const or_tables* t, int u, int i,
                       const double* mf_rho, double out[4]);
double or_ref_DV0_i_Dxi_u(const or_system* s, const or_tables* t, int u, int i,
                          const double* mf_rho);
/* d2 V0_i / dq_u2 as a 3 x 3 row-major block (evaluate_D2V0_i_Dmeanq2_u_MgHx). */
void or_ref_D2V0_i_Dq2_u(const or_system* s, const or_tables* t, int u, int i,
                         const double* mf_rho, double out[9]);
/* D_u = sum_{i in {u} U N(u)} beta_i d2V0_i/dq_u2 for the listed sites (n x 9). */
void or_ref_diag_blocks(const or_system* s, const or_tables* t, int nsites, const int* sites,
                        const double* mf_rho, double* D);
/* Zero temperature (src/ADP/MgHx-mf-V-bulk.cpp:147-329 and :333-526): dV_i/dq_u
 * and d2V_i/dq_u2 at the mean positions, rho = or_ref_rho0_i of every site. */
void or_ref_DV_i_Dq_u(const or_system* s, const or_tables* t, int u, int i, const double* rho,
                      double out[3]);
void or_ref_D2V_i_Dq2_u(const or_system* s, const or_tables* t, int u, int i, const double* rho,
                        double out[9]);
/* F_u = sum_i beta_i dV_i/dq_u (cnt x 3) and D_u = sum_i beta_i d2V_i/dq_u2 (cnt x 9). */
void or_ref_zero_t(const or_system* s, const or_tables* t, int nsites, const int* sites,
                   const double* rho, double* F, double* D);
/* Residual of sum_i S0_i for the listed sites u (all sites when sites==NULL). */
void or_ref_residual(const or_system* s, const or_tables* t, int nsites, const int* sites,
                     const double* mf_rho, double* Rq, double* Rs, double* Rx);

/* Factorised O(nb) form ------------------------------------------------------ */
/* Phase 1 per site: mf_rho, V0 and the per-atom adjoint data (stride 80). */
#define OR_P1_STRIDE 80
void or_fact_phase1(const or_system* s, const or_tables* t, int nthreads,
                    double* mf_rho, double* V0, double* p1);
/* Full residual.  Outputs Rq (n x 3), Rs, Rx, S0, V0, mf_rho per site and the
 * totals [sum S0, sum V0] over active sites. */
void or_fact_residual(const or_system* s, const or_tables* t, int nthreads,
                      double* Rq, double* Rs, double* Rx, double* S0, double* V0,
                      double* mf_rho, double* totals);
/* Diagonal Hessian blocks D_u (n x 9), same definition as or_ref_diag_blocks
 * (exact = 0), or the exact diagonal blocks of dR_q/dq (exact = 1: adds the
 * mixed codes of the coincident triplets (i, u, u) that the reference omits). */
void or_fact_diag_blocks(const or_system* s, const or_tables* t, int nthreads, int exact,
                         double* D);

/* Diffusion step (build-defined rate law; the reference declares only the
 * topology, src/Atoms/Neighbors.cpp:139-277, and the parameters,
 * src/Variables.hpp:66-90) ---------------------------------------------------
 * Diffusive neighbours: H sites (species code 1) within rc of an H site,
 * brute force over all pairs with the minimum image; at most cap per site.
 * Returns 0, 1 when a site has more than cap, 2 when an H site has none. */
int or_diff_neighbors(int n, const double* q, const int* spc, const double* lo,
                      const double* hi, const int* periodic, double rc, int cap, int* cnt,
                      int* lst);
/* Interaction chemical potential (units of k_B T) from the occupancy residual:
 * mu = Rx - (-3/(2x) - [x<1]/(1-x) - [H, x<0.9999] gamma) on active sites,
 * 0 on masked ones (x < 5e-4). */
void or_diff_mu(int n, const double* x, const int* spc, const double* gamma, const double* Rx,
                double* mu);
/* Lattice-gas rates in Onsager form with the ideal (Fermi-Dirac) site entropy
 * l(x) = ln(x/(1-x)) and a logarithmic-mean mobility:
 *   xdot_i = sum_j G_ij [ (x_j - x_i) + L(x_i, x_j) (mu_j - mu_i) ],
 *   L(a, b) = (b - a) / (l(b) - l(a)), 0 unless both sites are active
 *             (xa >= 5e-4: the activity is frozen at the start of a step),
 *   G_ij = d0 / r_ij^2 exp(-eb (beta_i + beta_j)/2)      (d0 in A^2/s, eb in eV),
 * i.e. G L (mu_tot_j - mu_tot_i) with mu_tot = mu + l(x). */
void or_diff_rates(int n, const double* q, const double* x, const double* xa, const double* mu,
                   const double* beta, const double* lo, const double* hi, const int* periodic,
                   const int* cnt, const int* lst, int cap, double d0, double eb, double* xdot);

#ifdef __cplusplus
}
#endif
#endif
