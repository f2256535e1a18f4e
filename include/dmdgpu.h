/*
 * dmdgpu.h -- C ABI of the MI355X-native mean-field DMD force path.
 *
 * The library replaces the per-site CPU evaluators of BT-AM
 * (src/ADP/MgHx-mf-V-bulk.cpp, reached through the dmd_equations table of
 * src/Macros.hpp:540-798) and the O(N^2) neighbour search
 * (src/Atoms/Neighbors.cpp:39-134) with HIP kernels for gfx950.  A PETSc SNES
 * FormFunction (the reference's evaluate_RHS,
 * src/Mechanical-eqs/Mechanical-Relaxation-bulk.cpp:441) calls
 * dmdgpu_residual() in place of its OpenMP site loops.
 *
 * Conventions
 *  - Every function returns a dmdgpu_status; 0 is success (PetscErrorCode
 *    style, src/Macros.hpp:28-38).  dmdgpu_last_error() gives a message.
 *  - Plain pointers and sizes only.  Pointers named h_* are host memory; the
 *    caller keeps ownership.  Device buffers belong to the context until
 *    dmdgpu_destroy().
 *  - Arrays are structure-of-arrays over n sites.  Positions are n x 3
 *    row-major (x, y, z per site).  Species codes follow enum AtomicSpecie
 *    (src/Macros.hpp:56-175): H = 1, Mg = 12; other species are rejected.
 *  - One context per stream; calls on one context are serialised by the
 *    caller.  Distinct contexts may be used from distinct threads.
 *  - All arithmetic of this interface is FP64 unless the precision option
 *    selects the split FP32 evaluation (dmdgpu_set_precision).
 */
#ifndef DMDGPU_H
#define DMDGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  DMDGPU_OK = 0,
  DMDGPU_ERR_ARG = 1,          /* bad argument / size                          */
  DMDGPU_ERR_HIP = 2,          /* HIP runtime error (no device, OOM, launch)    */
  DMDGPU_ERR_IO = 3,           /* table file missing or malformed              */
  DMDGPU_ERR_SPECIES = 4,      /* species other than H / Mg, or Mg with x<5e-4  */
  DMDGPU_ERR_MAXNEIGH = 5,     /* a site has more than 250 neighbours
                                  (Neighbors.cpp:99-107)                        */
  DMDGPU_ERR_NONEIGH = 6,      /* a site has no neighbours (Neighbors.cpp:110-116) */
  DMDGPU_ERR_BOX = 7,          /* periodic box edge < 2 r_c (Atom.cpp:230-255) */
  DMDGPU_ERR_STATE = 8,        /* call order: tables / state / neighbours unset */
  DMDGPU_ERR_TABLES = 9,       /* r-tables do not share one (n, dx) grid        */
  DMDGPU_ERR_BUDGET = 10,      /* instance does not fit the device-memory budget */
  DMDGPU_ERR_DOMAIN = 11       /* a spline argument left the table's domain: a
                                  density rho_bar <= -drho of the embedding table,
                                  where the reference's cubic_spline indexes a
                                  negative segment (src/Numerical/cubic-spline.cpp:
                                  84-93), or a non-finite density.  The outputs
                                  of that evaluation are still written.          */
} dmdgpu_status;

typedef struct dmdgpu_ctx dmdgpu_ctx;

/* Which table file init_adp_MgHx reads (src/ADP/MgHx-ADP.cpp:366-501). */
typedef enum { DMDGPU_MgMg = 0, DMDGPU_HH = 1, DMDGPU_MgH = 2 } dmdgpu_species_comb;

/* Context ------------------------------------------------------------------ */

/* Creates a context on HIP device `device` with its own stream.  No
 * counterpart in the reference (the swarm and the tables are globals there). */
dmdgpu_status dmdgpu_create(int device, dmdgpu_ctx** out);
dmdgpu_status dmdgpu_destroy(dmdgpu_ctx* ctx);
const char* dmdgpu_last_error(void);

/* The stream all kernels of this context are launched on (a hipStream_t). */
void* dmdgpu_stream(dmdgpu_ctx* ctx);
dmdgpu_status dmdgpu_sync(dmdgpu_ctx* ctx);

/* Tables --------------------------------------------------------------------
 * Replaces init_adp_MgHx(adpPotential*, species_comb_MgH, const char* folder)
 * (src/ADP/MgHx-ADP.cpp:366-501): reads adpmg.dat / adph.dat / adpmgh.dat in
 * the reference layout from `folder` and uploads them.  All three files must be
 * loaded before dmdgpu_set_state.  Tables loaded later (any grid) take effect
 * at the next evaluation, relaxation tick or diffusion step of every resident
 * instance. */
dmdgpu_status dmdgpu_init_adp_MgHx(dmdgpu_ctx* ctx, int species_comb, const char* folder);

/* Neighbour cut-off on mean distances; default r_cutoff_ADP_MgHx =
 * 6.2934034034 (src/Macros.hpp:857). */
dmdgpu_status dmdgpu_set_cutoff(dmdgpu_ctx* ctx, double rc);

/* State ----------------------------------------------------------------------
 * Copies the DMSwarm-equivalent fields (src/Atoms/Atom.cpp:149-203:
 * coor, stdv-q, molar-fraction, specie, thermalmultp = beta,
 * chemicalmultp = gamma) and the box (dump-input.cpp:76-125) to device SoA
 * buffers.  box = {xlo, ylo, zlo, xhi, yhi, zhi}; periodic[3] = 1 for "pp".
 * At most 6,710,784 sites per instance (DMDGPU_ERR_ARG beyond: the gradient
 * kernel addresses an instance's adjoint records, 640 B per site, with 32-bit
 * byte offsets). */
dmdgpu_status dmdgpu_set_state(dmdgpu_ctx* ctx, int n, const double* h_mean_q,
                               const double* h_stdv_q, const double* h_xi,
                               const int32_t* h_specie, const double* h_beta,
                               const double* h_gamma, const double* h_box,
                               const int32_t* h_periodic);

/* Replaces the unknowns only (positions, widths, occupancies), keeping the
 * neighbour lists: the SNES iterate update.  NULL leaves a field unchanged. */
dmdgpu_status dmdgpu_update_state(dmdgpu_ctx* ctx, const double* h_mean_q,
                                  const double* h_stdv_q, const double* h_xi);

/* Neighbours -----------------------------------------------------------------
 * Replaces neighbors(DMD*) (src/Atoms/Neighbors.cpp:39-134): every site j != i
 * with |q_i - q_j| <= r_c at mean positions, minimum image on periodic axes,
 * built on the GPU with a cell list.  Rows are sorted by site index, as the
 * reference's O(N^2) scan produces them. */
dmdgpu_status dmdgpu_neighbors(dmdgpu_ctx* ctx);

/* Copies the lists out: h_numneigh[n], h_list[n * cap] (row stride cap >= 250).
 * Mirrors AtomTopology {numneigh, mech_neighs_ptr} (src/Macros.hpp:238-244). */
dmdgpu_status dmdgpu_get_neighbors(dmdgpu_ctx* ctx, int32_t* h_numneigh, int32_t* h_list,
                                   int cap);

/* Evaluation -----------------------------------------------------------------
 * Mean-field density of every site, the loop of main-BT-AM.cpp:197-207 with
 * evaluate_mf_rho_i_adp_MgHx (MgHx-mf-V-bulk.cpp:708-787) as the body. */
dmdgpu_status dmdgpu_mf_rho(dmdgpu_ctx* ctx, double* h_mf_rho);

/* Per-site V0_i and S0_i of every site: evaluate_V0_i_adp_MgHx (:791-953) and
 * evaluate_S0_i_adp_MgHx (:957-1142) with mf_rho from the same call. */
dmdgpu_status dmdgpu_site_energies(dmdgpu_ctx* ctx, double* h_V0, double* h_S0);

/* Per-(u, i) parity hooks at the reference's own granularity: for m queries
 * (h_u[k], h_i[k]) in the caller's site numbering,
 *   dmdgpu_site_DV0_Dq:  h_out[4 k .. 4 k + 3] = [dV0_i/dq_u (3), dV0_i/dsigma_u],
 *                        evaluate_DV0_i_Dq_u_MgHx (MgHx-mf-V-bulk.cpp:1823-2098);
 *   dmdgpu_site_DV0_Dxi: h_out[k] = dV0_i/dx_u,
 *                        evaluate_DV0_i_Dxi_u_adp_MgHx (:2102-2325), with its
 *                        embedding form for an Mg centre i (F'(rho_i) drho_i/dx_u
 *                        without the factor x_i, and 0 for u = i, :2300-2315).
 * Zeros when u or i is inactive or u is neither i nor a neighbour of i, as
 * the reference's loops give.  FP64 at the current state, with the context's
 * pair rule (the reference defines them for the 3rd-order rule).  Unweighted:
 * R_q(u) = sum_{i in {u} U N(u)} beta_i dV0_i/dq_u. */
dmdgpu_status dmdgpu_site_DV0_Dq(dmdgpu_ctx* ctx, int m, const int32_t* h_u, const int32_t* h_i,
                                 double* h_out);
dmdgpu_status dmdgpu_site_DV0_Dxi(dmdgpu_ctx* ctx, int m, const int32_t* h_u, const int32_t* h_i,
                                  double* h_out);

/* The SNES FormFunction body.  Gradient of the free-entropy functional
 * sum_i S0_i (MgHx-mf-V-bulk.cpp:957-1142) with respect to the mean positions,
 * widths and occupancies of every site:
 *   R_q(u)     = sum_{i in {u} U N(u)} beta_i dV0_i/dq_u         (n x 3)
 *   R_sigma(u) = sum_i beta_i dV0_i/dsigma_u - 3 / sigma_u        (n)
 *   R_x(u)     = sum_i beta_i dV0_i/dx_u + d(-log Z0_u)/dx_u - gamma_u [H] (n)
 * The per-(u, i) terms are evaluate_DV0_i_Dq_u_MgHx (:1823-2098) and
 * evaluate_DV0_i_Dxi_u_adp_MgHx (:2102-2325).  Inactive sites (x < 5e-4) get
 * zeros.  totals[2] = {sum_i S0_i, sum_i V0_i}.  Any output may be NULL.
 * DMDGPU_ERR_DOMAIN when a density left the embedding table's domain (this
 * and dmdgpu_mf_rho, dmdgpu_site_energies, dmdgpu_diag_blocks,
 * dmdgpu_instance_residual). */
dmdgpu_status dmdgpu_residual(dmdgpu_ctx* ctx, double* h_Rq, double* h_Rsigma, double* h_Rx,
                              double* h_totals);

/* The same residual in the packed vector form of a SNES FormFunction
 * (evaluate_RHS, Mechanical-Relaxation-bulk.cpp:441, with X and Y the
 * solution and function Vecs): X = [q (3n) | sigma (n)] sets the positions
 * and widths first (NULL: evaluate at the current state), R = [R_q (3n) |
 * R_sigma (n) | R_x (n)] receives the residual (may be NULL); caller's site
 * order.  on_device = 0: X and R are host memory (staged through page-locked
 * buffers); 1: device memory of the context's device (no host copies; the
 * reordering runs on the device).  totals[2] as dmdgpu_residual (host, may be
 * NULL).  Occupancies are set with dmdgpu_update_state. */
dmdgpu_status dmdgpu_residual_packed(dmdgpu_ctx* ctx, const double* X, double* R,
                                     double* h_totals, int on_device);

/* Same evaluation on the device-resident state, leaving the results in device
 * buffers (no host copies); used by resident solvers and for timing. */
dmdgpu_status dmdgpu_residual_resident(dmdgpu_ctx* ctx);

/* Jacobian-vector product of the (q, sigma) residual for a MatShell MatMult:
 * the reference builds its Jacobian by finite-difference colouring
 * (SNESComputeJacobianDefaultColor + MatFDColoring,
 * Mechanical-Relaxation-bulk.cpp:282-319); this is the matrix-free form.
 * h_X = [q (3n) | sigma (n)], h_V likewise, h_JV = [dR_q (3n) | dR_sigma (n)]:
 * central difference of two residual evaluations at X +- eps V (RMS
 * displacement 1e-5 A).  Occupancies fixed; the state is left at X. */
dmdgpu_status dmdgpu_jvp(dmdgpu_ctx* ctx, const double* h_X, const double* h_V, double* h_JV);
/* The same product with X, V and JV in host memory (on_device = 0, staged
 * through page-locked buffers) or in device memory of the context's device
 * (on_device = 1: a MatShell over device Vecs; no host copies of the vectors,
 * |V| is reduced on the device).  dmdgpu_jvp is the on_device = 0 form;
 * both give bit for bit the same JV. */
dmdgpu_status dmdgpu_jvp_packed(dmdgpu_ctx* ctx, const double* X, const double* V, double* JV,
                                int on_device);

/* Diagonal 3x3 blocks of the force residual's Jacobian, for a block-Jacobi
 * preconditioner: h_D[9 u + 3 a + b] = D_u[a][b] with
 *   D_u = sum_{i in {u} U N(u)} beta_i d2V0_i/dq_u2,
 * the per-(u, i) Hessians of evaluate_D2V0_i_Dmeanq2_u_MgHx
 * (MgHx-mf-V-bulk.cpp:1366-1599) summed with the weights of R_q(u).
 * exact = 0: as that function assembles them (on the coincident triplet
 * (i, u, u) it sums the codes 4 and 8 of u, not the mixed codes 5 and 7);
 * exact = 1: the exact d R_q(u) / d q_u, those mixed codes included.  The
 * pair rule is the context's (dmdgpu_set_rule); the reference defines the
 * blocks for the 3rd-order rule.  Evaluated in FP64 whatever the precision
 * setting; inactive sites (x < 5e-4) get zeros. */
dmdgpu_status dmdgpu_diag_blocks(dmdgpu_ctx* ctx, int exact, double* h_D);
/* The same blocks for resident instance `id` at its current state (caller's
 * site order); DMDGPU_ERR_STATE while a solve of that instance is running. */
dmdgpu_status dmdgpu_instance_diag_blocks(dmdgpu_ctx* ctx, int id, int exact, double* h_D);
/* The blocks of dmdgpu_diag_blocks written into DEVICE memory d_D (n x 9
 * doubles, caller's site order), bit for bit the same values, with no device
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The blocks are staged in a fixed device array of 65,536
 * sites' blocks (4.5 MiB, held by each device the library runs on until the
 * process ends); an instance with more sites is refused with DMDGPU_ERR_ARG.
 * The array is shared by every call on that device and keeps the last call's
 * contents: calls that overlap on one device (from different contexts or
 * threads) race.  Needs dmdgpu_neighbors first, and tables loaded since the
 * last evaluation to have been used by an allocating call (DMDGPU_ERR_STATE
 * otherwise: their upload may allocate); returns after the stream has
 * finished (with DMDGPU_ERR_DOMAIN as dmdgpu_diag_blocks). */
dmdgpu_status dmdgpu_diag_blocks_noalloc(dmdgpu_ctx* ctx, int exact, double* d_D);

/* Kernel tuning, for both precisions: k_phase1 = lanes of a wavefront that
 * share one site in phase 1 (4, 8 or 16); k_phase2 = blocks per 256-site chunk
 * of the pair kernels (4, 8, 16, 32 or 64: on average that many lanes per site);
 * unroll2 = phase 2 advances two shifted or three pair-rule points per loop
 * iteration (0: one).
 * The tuner picks them per workload.  No counterpart in the reference. */
dmdgpu_status dmdgpu_set_tuning(dmdgpu_ctx* ctx, int k_phase1, int k_phase2, int unroll2);
/* The same for one precision only (DMDGPU_FP64 or DMDGPU_SPLIT32): batched
 * evaluations run both, each with its own tuning. */
dmdgpu_status dmdgpu_set_tuning_precision(dmdgpu_ctx* ctx, int precision, int k_phase1,
                                          int k_phase2, int unroll2);

/* Precision of the quadrature-point evaluation.  DMDGPU_FP64: everything FP64
 * (the reference's arithmetic).  DMDGPU_SPLIT32: splines, feature vectors and
 * point gradients in FP32, with pair distances, spline offsets, per-site sums,
 * records, embedding, entropy terms and totals in FP64.  No counterpart in the
 * reference. */
enum { DMDGPU_FP64 = 0, DMDGPU_SPLIT32 = 1 };
dmdgpu_status dmdgpu_set_precision(dmdgpu_ctx* ctx, int precision);

/* Quadrature of the pair terms (density, pair potential) and of the
 * coincident angular terms.  3 (default): the reference's 12-point 6D Stroud
 * rule (src/Numerical/Quadrature-Hermitian-3th.cpp:36-110).  5: 27-point tensor
 * Gauss-Hermite rule on the relative coordinate r ~ N(r_ij, (s_i^2+s_j^2) I)
 * (nodes 0, +-sqrt(3) s_rel; weights 2/3, 1/6 per axis; exact to degree 5).
 * 44: the reference's own 44-point 6D degree-5 pair rule
 * (src/Numerical/Quadrature-Hermitian-5th.cpp:31-157, which its evaluators do
 * not call).  6: the 6-point Stroud rule on the relative coordinate
 * (r +- sqrt(3) s_rel e_l, weight 1/6, degree 3): the reference's 12-point
 * rule when the two widths are equal, and otherwise exact in the fourth
 * moment along an axis where that rule is not.  Distinct angular triplets keep the 3rd-order 9D rule under all
 * three: the reference has no higher-order 9D rule. */
dmdgpu_status dmdgpu_set_rule(dmdgpu_ctx* ctx, int order);

/* Resident instances -------------------------------------------------------
 * No counterpart in the reference (it holds one swarm per process).  A context
 * keeps any number of independent instances (systems) resident in one device
 * arena; one launch of the residual kernels evaluates all of them and every
 * instance's result is the same as when it is evaluated alone.
 * dmdgpu_set_state is the one-instance form: it replaces all instances. */

/* Copies an instance in (same fields as dmdgpu_set_state), builds its neighbour
 * lists and returns its id.  DMDGPU_ERR_BUDGET when the arena would exceed the
 * budget set with dmdgpu_set_budget; DMDGPU_ERR_ARG for more than 6,710,784
 * sites (as dmdgpu_set_state); DMDGPU_ERR_HIP when the device cannot grow the
 * arena, which leaves the resident instances as they were (should even their
 * arrays not be restorable, every instance is removed and the message says
 * so). */
dmdgpu_status dmdgpu_add_instance(dmdgpu_ctx* ctx, int n, const double* h_mean_q,
                                  const double* h_stdv_q, const double* h_xi,
                                  const int32_t* h_specie, const double* h_beta,
                                  const double* h_gamma, const double* h_box,
                                  const int32_t* h_periodic, int* id);
dmdgpu_status dmdgpu_remove_instance(dmdgpu_ctx* ctx, int id);
/* Sites, active ordered pairs of an instance, and the arena extent (sites). */
dmdgpu_status dmdgpu_instance_info(dmdgpu_ctx* ctx, int id, int* n, int64_t* pairs,
                                   int64_t* arena_sites);
/* Device-memory budget in bytes (0 = unlimited): every allocation of the
 * context -- tables, arena, cell grid, metadata -- must stay within it, also
 * while the arena grows (an instance that would take it past the budget is
 * refused with DMDGPU_ERR_BUDGET).  dmdgpu_bytes_per_site: bytes of one arena
 * site (state, lists, records, outputs, solver vectors); dmdgpu_device_bytes:
 * the context's device allocations now. */
dmdgpu_status dmdgpu_set_budget(dmdgpu_ctx* ctx, int64_t bytes);
int64_t dmdgpu_bytes_per_site(void);
dmdgpu_status dmdgpu_device_bytes(dmdgpu_ctx* ctx, int64_t* bytes);
/* One residual evaluation of every instance (default precision). */
dmdgpu_status dmdgpu_evaluate(dmdgpu_ctx* ctx);
/* Results of the last evaluation of an instance, in its caller's site order. */
dmdgpu_status dmdgpu_instance_residual(dmdgpu_ctx* ctx, int id, double* h_Rq, double* h_Rsigma,
                                       double* h_Rx, double* h_totals);
/* Current iterate (mean positions, widths, occupancies) of an instance. */
dmdgpu_status dmdgpu_instance_state(dmdgpu_ctx* ctx, int id, double* h_mean_q, double* h_stdv_q,
                                    double* h_xi);
/* Replaces the unknowns of an instance (NULL leaves a field unchanged). */
dmdgpu_status dmdgpu_instance_set_unknowns(dmdgpu_ctx* ctx, int id, const double* h_mean_q,
                                           const double* h_stdv_q, const double* h_xi);

/* Relaxation -------------------------------------------------------------------
 * Stands in for the reference's SNES solve (SNESSolve,
 * Mechanical-Relaxation-bulk.cpp:346, NGMRES + line search per
 * main-BT-AM.cpp:64-67): an inexact Newton-Krylov iteration on the mean
 * positions and widths (selected by dof_mask) of the active sites of each
 * instance, occupancies fixed, until the 2-norm of the selected residual
 * components is <= atol.  Hessian actions are
 * finite differences of the residual; with `mixed`, those probes use the split
 * FP32 evaluation; while |R|_2 > 1e-3 sqrt(ndof) (a thousand times the split
 * residual's noise) the Newton and line-search residuals are split
 * evaluations too, and below it every accepted residual and the stopping test
 * are FP64.  Instances advance independently: one tick gives every running
 * instance exactly one residual evaluation (all instances in one launch).
 * A relaxation tick's evaluation forms only what the solve needs: no R_x;
 * when dof_mask selects no H unknowns, no pairs with two H ends and no R_q,
 * R_sigma of the H sites; with the residual-norm line search (merit = 0), no
 * energies (the pair-rule points give the densities only).  Every relaxed
 * component is the same bit for bit.  dmdgpu_instance_residual after such a
 * tick evaluates again, at the same point and in the same precision, for the
 * whole residual and the totals. */
typedef struct {
  double atol;     /* stop when |(R_q, R_sigma)|_2 <= atol                */
  int max_newton;  /* Newton steps                                        */
  int max_cg;      /* CG iterations per Newton step                       */
  double eta;      /* CG stops at |r| <= eta |R|                          */
  int mixed;       /* 1: split-precision Hessian probes                   */
  int max_evals;   /* residual evaluations (0 = no limit)                 */
  int dof_mask;    /* relaxed unknowns, DMDGPU_DOF_* bits; 0 = Mg q + sigma */
  double probe;    /* RMS displacement of a Hessian probe (A); 0 = default
                      (1e-7 with FP64 probes, 5e-5 with split-precision ones) */
  int precond;     /* k > 0: block-Jacobi preconditioned CG: the exact
                      diagonal blocks of dmdgpu_diag_blocks are inverted per
                      site (3x3 on q; 1/(3/sigma^2 + tr D_u) on sigma) at
                      Newton steps 0, k, 2k, ... and kept in between;
                      0: plain CG */
  int merit;       /* line-search merit: 0 = the residual norm |R|_2
                      (sufficient decrease); 1 = the free energy sum_i S0_i,
                      of which R is the gradient (Armijo on the energy, with a
                      step also accepted when the energy change is below its
                      rounding level and |R| decreases): monotone descent
                      where the landscape is not convex (free surfaces,
                      interstitial H positions) */
  int solver;      /* 0: the inexact Newton-CG above; 1: NGMRES(m) with a
                      critical-point line search, the reference's SNES
                      configuration (main-BT-AM.cpp:63-66: snes_type ngmres,
                      ngmres_m 3, linesearch cp): per iteration a step
                      X - lambda R with lambda from secant iterations on
                      R(X - lambda R) . R = 0 (at most 3, steps bounded as
                      above), then the combination of the last m iterates
                      that minimises the linearised residual, kept when it
                      passes the Oosterlee-Washio selection tests (restart
                      after 2 failures); FP64 evaluations only (mixed and
                      precond must be 0); max_newton bounds its iterations */
  int ngmres_m;    /* NGMRES history depth 1..3 (0 = 3) */
} dmdgpu_relax_opts;
/* Relaxed unknowns.  Interstitial H at fixed occupancy has a soft (and at low
 * occupancy width-unbounded: sigma_eq^2 ~ 1/(beta k x)) landscape, so the
 * default relaxes the Mg lattice and keeps H sites fixed.  With H positions
 * (DMDGPU_DOF_H_Q) periodic bulk systems converge; with free surfaces the
 * structure can reconstruct far from the start, where the residual-norm line
 * search stalls and the energy merit (merit = 1) keeps descending. */
enum {
  DMDGPU_DOF_MG_Q = 1,
  DMDGPU_DOF_MG_SIGMA = 2,
  DMDGPU_DOF_H_Q = 4,
  DMDGPU_DOF_H_SIGMA = 8
};
enum {
  DMDGPU_RELAX_IDLE = 0,
  DMDGPU_RELAX_RUNNING = 1,
  DMDGPU_RELAX_CONVERGED = 2,
  DMDGPU_RELAX_MAXITER = 3,
  DMDGPU_RELAX_STALLED = 4,  /* line search found no decrease */
  DMDGPU_RELAX_DOMAIN = 5    /* an evaluation left the embedding table's domain
                                (see DMDGPU_ERR_DOMAIN) */
};
dmdgpu_status dmdgpu_relax_start(dmdgpu_ctx* ctx, int id, const dmdgpu_relax_opts* opts);
/* One tick of every running instance; *running = instances still running. */
dmdgpu_status dmdgpu_relax_tick(dmdgpu_ctx* ctx, int* running);
dmdgpu_status dmdgpu_relax_status(dmdgpu_ctx* ctx, int id, int* state, double* rnorm,
                                  int* newton_steps, int* residual_evals);
/* Residual evaluations of a solve by precision (FP64, split), and the active
 * ordered pairs of an instance (all; those with an Mg end, the pairs a tick
 * relaxing the Mg lattice evaluates in its gradient phase): work counts for
 * throughput figures. */
dmdgpu_status dmdgpu_relax_evals(dmdgpu_ctx* ctx, int id, int* evals_fp64, int* evals_split);
dmdgpu_status dmdgpu_instance_pairs(dmdgpu_ctx* ctx, int id, int64_t* pairs, int64_t* pairs_mg);
/* Free energy sum_i S0_i at the solve's current iterate: with merit = 1 as
 * the solve last evaluated it; with merit = 0 (whose ticks form no energies)
 * evaluated in FP64 at the final iterate once the solve has ended
 * (DMDGPU_ERR_STATE while it runs). */
dmdgpu_status dmdgpu_relax_energy(dmdgpu_ctx* ctx, int id, double* energy);
/* Starts every instance and ticks until all are done (or max_ticks). */
dmdgpu_status dmdgpu_relax(dmdgpu_ctx* ctx, const dmdgpu_relax_opts* opts, int max_ticks);
/* 1 (default): a tick's device work is captured once as a hipGraph and
 * replayed; recaptured when the arena or the tuning changes. */
dmdgpu_status dmdgpu_set_graph(dmdgpu_ctx* ctx, int on);

/* Diffusion step -------------------------------------------------------------
 * One implicit (backward Euler) update of the H occupancies of an instance,
 * x = x_n + dt xdot(x), positions and widths fixed, solved by Picard sweeps
 * (one residual evaluation each) until max |x_k+1 - x_k| <= tol.  Diffusive
 * sites are the H sites; their neighbours are the H sites within 2.6 A
 * (list_of_diffusive_neighs, src/Atoms/Neighbors.cpp:139-277; r_cutoff_Eb and
 * max_chemical_neighs = 50, src/Macros.hpp:860,863; parameters declared in
 * src/Variables.hpp:66-90).  The reference has no rate law; this one is a
 * lattice-gas gradient flow in Onsager form with the ideal site entropy
 * l(x) = ln(x/(1-x)) and a logarithmic-mean mobility:
 *   xdot_i = sum_j G_ij [(x_j - x_i) + L(x_i, x_j) (mu_j - mu_i)],
 *   L(a, b) = (b - a) / (l(b) - l(a))   (a (1 - a) when a = b),
 *   G_ij = D0 / r_ij^2 exp(-Eb (beta_i + beta_j)/2),
 * mu_i = the occupancy residual R_x(i) without the site's own entropy terms
 * (in k_B T; 0 on masked sites): Fick's law with D = D0 exp(-Eb/k_B T) when
 * mu = 0, antisymmetric pair fluxes (mass conservation), and stationary states
 * with mu + l(x) uniform.  A masked site (x < 5e-4) exchanges by the Fick term
 * only; which sites are masked is frozen at the start of the step.
 * Advanced by dmdgpu_relax_tick together with relaxations; dmdgpu_relax_status
 * reports the sweeps (newton_steps) and the last max |dx| (rnorm).
 * Errors: DMDGPU_ERR_MAXNEIGH (> 50 diffusive neighbours),
 * DMDGPU_ERR_NONEIGH (an H site without one), as the reference. */
typedef struct {
  double dt;     /* time step (s)                                  */
  double D0;     /* diffusivity prefactor (m^2/s)                  */
  double Eb;     /* migration barrier (eV)                         */
  double tol;    /* stop when max |x_k+1 - x_k| <= tol             */
  int max_iter;  /* Picard sweeps                                  */
  int mixed;     /* 1: split-precision sweeps until max |dx| <= 1e-7,
                    then FP64 sweeps to tol                        */
} dmdgpu_diffuse_opts;
dmdgpu_status dmdgpu_diffuse_start(dmdgpu_ctx* ctx, int id, const dmdgpu_diffuse_opts* opts);
/* xdot of an instance at its current state (caller's order; 0 on Mg sites),
 * with the D0 and Eb of opts.  Not while solves are running. */
dmdgpu_status dmdgpu_diffusion_rates(dmdgpu_ctx* ctx, int id, const dmdgpu_diffuse_opts* opts,
                                     double* h_xdot);
/* Diffusive neighbour lists of an instance in the caller's numbering:
 * h_counts (n), h_lists (n x 50, row order of the neighbour list). */
dmdgpu_status dmdgpu_diffusion_neighbors(dmdgpu_ctx* ctx, int id, int* h_counts, int* h_lists);

/* Instrumentation: when on, every kernel launch is bracketed by HIP events on
 * the context's stream.  dmdgpu_kernel_times returns, per kernel id (see
 * dmdgpu_kernel_name), the launch count and the summed milliseconds. */
dmdgpu_status dmdgpu_set_timing(dmdgpu_ctx* ctx, int on);
int dmdgpu_kernel_count(void);
const char* dmdgpu_kernel_name(int id);
dmdgpu_status dmdgpu_kernel_times(dmdgpu_ctx* ctx, int64_t* h_launches, double* h_ms);
dmdgpu_status dmdgpu_reset_kernel_times(dmdgpu_ctx* ctx);

/* Counts for the throughput metric: active ordered pairs of the current
 * lists (P), and pair-quadrature evaluations per residual (50 P). */
dmdgpu_status dmdgpu_work_counts(dmdgpu_ctx* ctx, int64_t* pairs, int64_t* pqe_per_residual);

#ifdef __cplusplus
}
#endif
#endif
