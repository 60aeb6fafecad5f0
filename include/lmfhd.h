/*
 * lmfhd.h -- C ABI of the MI355X-native implicit staggered-grid Stokes solver.
 *
 * Drop-in for LowMachFHD's Stokes solve (staggered_grid/src_gmres): the entry points below
 * replace the Fortran routines named beside each one, with the same argument meaning.  A Fortran
 * host binds them through iso_c_binding; array arguments are FBoxLib fab views
 * (pointer + box bounds + ghost width, column-major), so a caller passes dataptr(mf, box) as is.
 *
 * Conventions
 *   - Single level.  The caller's multifabs may be split into boxes (FBoxLib's boxarray, e.g.
 *     boxarray_maxsize at main_driver.f90:185): lmfhd_desc.nboxes / box_lo / box_hi list them.
 *     Every array argument then holds one fab per box, and a multi-component argument (dm face
 *     components, edge arrays, ...) holds nboxes fabs per component, component-major:
 *     fab[c * nboxes + b] is box b of component c.  With nboxes = 0 one box covers the domain,
 *     each array is one fab, and its lo / hi are taken as labels (only the extents are checked).
 *   - Array bounds: lo/hi are the VALID index bounds of the fab, including the nodal extension
 *     (a face array in direction d has hi[d] = box hi[d] + 1); ng ghost layers surround them in
 *     the first dm directions.  Ghost values on input are ignored (the library applies the
 *     reference's periodic and wall rules itself); on output only the valid region is written.
 *     A nodal point on a boundary between two boxes (a face or edge both hold) is read from the
 *     first box in the list that holds it and written to every box that holds it.
 *   - Boundary codes are the reference's physical codes (src_common/bc.f90): -1 periodic,
 *     100..103 no-slip walls, 200..203 slip walls.
 *   - Every function returns LMFHD_OK or a negative error code; the library never aborts.
 *   - One host thread per context; each context owns one HIP stream and its device memory.
 */
#ifndef LMFHD_H
#define LMFHD_H
#ifdef __cplusplus
extern "C" {
#endif

#define LMFHD_OK 0
#define LMFHD_E_ARG -1        /* bad argument / shape mismatch */
#define LMFHD_E_NOMEM -2      /* device allocation failed */
#define LMFHD_E_BREAKDOWN -3  /* Arnoldi breakdown H(i+1,i) = 0 (gmres.f90:390-392 aborts) */
#define LMFHD_E_NAN -4        /* non-finite value met in a norm */
#define LMFHD_E_HIP -5        /* HIP runtime error */
#define LMFHD_E_STATE -6      /* coefficients not set */
#define LMFHD_E_BUDGET -7     /* serving: request exceeds the HBM budget */

#define LMFHD_FP64 0          /* every step in FP64 (reference arithmetic) */
#define LMFHD_MIXED 1         /* FP32 multigrid preconditioner, FP64 Krylov */

#define LMFHD_GMRES_LEFT 0    /* left-preconditioned GMRES of gmres.f90 (reference criterion) */
#define LMFHD_FGMRES 1        /* right-preconditioned flexible GMRES (true-residual criterion) */

/* FBoxLib fab view of host memory. */
typedef struct lmfhd_fab {
  double* data;   /* element (lo-ng, lo-ng, lo-ng) */
  int lo[3];
  int hi[3];
  int ng;
} lmfhd_fab;

/* Problem and solver description.  Field names follow probin_common.f90 / probin_gmres.f90. */
typedef struct lmfhd_desc {
  int dm;                 /* 2 or 3 */
  int n_cells[3];
  double dx;              /* dx = dy = dz (stag_applyop.f90:118-128 requires it) */
  double dz2d;            /* 2-D cell depth for the stochastic variance (default 1) */
  int bc_lo[3], bc_hi[3];
  int visc_type;          /* -3..-1, 1..3 (probin_common.f90:142-149) */
  int precision;          /* LMFHD_FP64 or LMFHD_MIXED */
  int krylov;             /* LMFHD_GMRES_LEFT or LMFHD_FGMRES */
  int precon_type;        /* 1 projection (-1: the expensive pressure update), +-2 lower and
                             +-3 upper triangular, +-4 block diagonal, +-5 Uzawa-type; a negative
                             sign flips the pressure block (probin_gmres.f90:26-36,
                             apply_precon.f90:91-426) */
  int gmres_max_inner, gmres_max_outer, gmres_max_iter, gmres_min_iter;
  double gmres_rel_tol, gmres_abs_tol, p_norm_weight;
  int stag_mg_nsmooths_down, stag_mg_nsmooths_up, stag_mg_nsmooths_bottom;
  int stag_mg_minwidth, stag_mg_max_vcycles, stag_mg_smoother;
  double stag_mg_omega, stag_mg_rel_tol;
  int mg_nsmooths_down, mg_nsmooths_up, mg_nsmooths_bottom, mg_minwidth, mg_max_vcycles;
  int device;             /* HIP device ordinal */
  int gmres_spatial_order; /* 2, or 4: apply_matrix with stag_applyop_4o + compute_grad_4o
                              (apply_matrix.f90:123-135; visc_type 1, periodic boxes) */
  int visc_schur_approx;   /* 0, or 1: the pressure block adds visc_schur_complement
                              (apply_precon.f90:538) built from two MAC solves with face
                              coefficient 1/diag(L) (inverse_diag_lap); |visc_type| 1 or 2 */
  int filtering_width;     /* 0, or 1, 2, 4: stochastic_m_fluxdiv filters the scaled stresses
                              (filter_stoch_m_flux, stochastic_m_fluxdiv.f90:604-778; periodic
                              boxes with n_cells >= filtering_width) */
  int graphs;              /* 1: each Arnoldi step of gmres (operator, preconditioner, Gram-Schmidt,
                              Givens update) and each preconditioner application is captured as a
                              HIP graph on first use and replayed; 0: kernels launched one by one
                              (same results either way); -1 (the default): the size class's tuning
                              record decides (lmfhd_tune.graphs; off without a record) */
  int batch;               /* 1: the preconditioner applications of concurrent solves of one size class
                              (contexts created with batch = 1 and the same grid, boundary
                              conditions and solver form, solving at the same time) run in lock step
                              as one batched launch sequence: the k-th launch of every solve in one
                              launch, one grid slice per solve.  Results equal separate solves bit
                              for bit.  Applies with precon_type 1, visc_schur_approx 0 and one
                              V-cycle per application (the defaults); otherwise ignored.  0: off */
  int nboxes;              /* 0: one box covers the domain; >= 1: the boxes of the caller's boxarray,
                              which must tile the domain without overlap */
  const int* box_lo;       /* 3 * nboxes cell indices (domain cells are 0 .. n_cells - 1), box b at
                              [3b, 3b + 3); the third index is 0 in 2-D.  Copied by lmfhd_ctx_create */
  const int* box_hi;
} lmfhd_desc;

typedef struct lmfhd_stats {
  int iters;              /* inner iterations (Krylov vectors built), gmres.f90 total_iter */
  int outer_iters;
  int converged;          /* 0 reproduces the reference's warning-only non-convergence */
  int n_precon, n_matrix; /* operator applications */
  double norm_b, norm_pre_b, norm_init_resid, norm_resid, norm_resid_stokes;
  double true_rel_resid;  /* |b - A x| / |b| of the returned x, FP64 */
  double resid_hist[256]; /* residual estimate after each inner iteration */
} lmfhd_stats;

typedef struct lmfhd_ctx lmfhd_ctx;

/* Defaults of probin_gmres.f90:122-163 and probin_common.f90:272-356. */
void lmfhd_desc_defaults(lmfhd_desc* d);

/* Replaces stag_mg_layout_build (stag_mg_layout.f90:21) and mgt_macproj_precon_build
 * (macproject.f90:364): builds the multigrid hierarchies and the device state. */
int lmfhd_ctx_create(const lmfhd_desc* d, lmfhd_ctx** out);
int lmfhd_ctx_destroy(lmfhd_ctx* ctx);
/* HIP stream (hipStream_t) on which every kernel of this context is launched. */
void* lmfhd_ctx_stream(lmfhd_ctx* ctx);
/* Device bytes held by the context. */
long long lmfhd_ctx_device_bytes(lmfhd_ctx* ctx);

/* Coefficients of (theta alpha - A0): alpha_fc (dm face arrays), beta_cc and gamma_cc (cell),
 * beta_ed (3 edge arrays xy,xz,yz in 3-D; 1 nodal array in 2-D), theta_alpha.  These are the
 * alpha_fc, beta, beta_ed, gamma, theta_alpha arguments of gmres (gmres.f90:32-47). */
int lmfhd_set_coeffs(lmfhd_ctx* ctx, const lmfhd_fab* alpha_fc, const lmfhd_fab* beta_cc,
                     const lmfhd_fab* beta_ed, const lmfhd_fab* gamma_cc, double theta_alpha);

/* stag_applyop_level (stag_applyop.f90:70): Lphi = (theta alpha - A0) phi on the faces of
 * colour `color` (0 = all; 1..2dm: component (color+1)/2, faces with (i+j+k) even for odd
 * colours); Lphi is zero on wall faces.  Uncoloured faces of Lphi are left unchanged. */
int lmfhd_stag_applyop(lmfhd_ctx* ctx, const lmfhd_fab* phi, lmfhd_fab* Lphi, int color);

/* apply_matrix (apply_matrix.f90:24): b_u = (theta alpha - A0) x_u + G x_p, b_p = -D x_u. */
int lmfhd_apply_matrix(lmfhd_ctx* ctx, const lmfhd_fab* x_u, const lmfhd_fab* x_p,
                       lmfhd_fab* b_u, lmfhd_fab* b_p);

/* Inhomogeneous velocity boundary values, as set_inhomogeneous_vel_bcs leaves them
 * (multifab_physbc_stag.f90:2462): vel_bc_n = dm face fabs (only their wall faces are read);
 * vel_bc_t = the tangential values on the walls, 2 node fabs in 2-D (y-velocity on x-walls,
 * x-velocity on y-walls) or 6 edge fabs in 3-D (y, z on x-walls; x, z on y-walls; x, y on
 * z-walls), each nodal in its wall direction and its component; slip-wall entries hold
 * traction / eta, as the reference divides them.  Ghosts: 2 v - interior at no-slip walls,
 * interior -/+ dx v at slip walls (lo / hi side), the wall faces take vel_bc_n.
 * apply_matrix with the optional vel_bc_n, vel_bc_t arguments (apply_matrix.f90:24-40, 83-105). */
int lmfhd_apply_matrix_bc(lmfhd_ctx* ctx, const lmfhd_fab* x_u, const lmfhd_fab* x_p, lmfhd_fab* b_u,
                          lmfhd_fab* b_p, const lmfhd_fab* vel_bc_n, const lmfhd_fab* vel_bc_t);
/* convert_to_homogeneous (src_gmres/convert_to_homogeneous.f90:19): in place, b_u -= L phi and
 * b_p += D phi for phi = 0 carrying those boundary values; the solve of the converted system
 * with homogeneous walls then gives the interior velocity of the inhomogeneous problem. */
int lmfhd_convert_to_homogeneous(lmfhd_ctx* ctx, lmfhd_fab* b_u, lmfhd_fab* b_p, const lmfhd_fab* vel_bc_n,
                                 const lmfhd_fab* vel_bc_t);
/* The plotfile cell fields of lmfhd_plot_fields_dev from fabs; out = one cell fab per field. */
int lmfhd_plot_fields(lmfhd_ctx* ctx, const lmfhd_fab* x_u, const lmfhd_fab* x_p, int flags, lmfhd_fab* out);

/* apply_precon (apply_precon.f90:29): every precon_type of the descriptor (+-1 .. +-5), with
 * visc_schur_complement (:538) when visc_schur_approx = 1. */
int lmfhd_apply_precon(lmfhd_ctx* ctx, const lmfhd_fab* b_u, const lmfhd_fab* b_p,
                       lmfhd_fab* x_u, lmfhd_fab* x_p);

/* stag_mg_solver (stag_mg_solver.f90:37): phi (initial guess in, solution out). */
int lmfhd_stag_mg_solver(lmfhd_ctx* ctx, const lmfhd_fab* rhs, lmfhd_fab* phi);

/* macproject's cell-centred multigrid solve of D(alpha_fc^-1 G) phi = rhs (macproject.f90:34,
 * mac_multigrid :127), phi from zero. */
int lmfhd_mac_mg_solver(lmfhd_ctx* ctx, const lmfhd_fab* rhs, lmfhd_fab* phi);

/* gmres (gmres.f90:32): x_u, x_p initial guess in, solution out; norm_pre_rhs may be NULL. */
int lmfhd_gmres(lmfhd_ctx* ctx, const lmfhd_fab* b_u, const lmfhd_fab* b_p, lmfhd_fab* x_u,
                lmfhd_fab* x_p, double* norm_pre_rhs, lmfhd_stats* stats);

/* compute_div (div_and_grad.f90:884) and compute_grad (:417, pressure BC rules). */
int lmfhd_compute_div(lmfhd_ctx* ctx, const lmfhd_fab* u, lmfhd_fab* div);
int lmfhd_compute_grad(lmfhd_ctx* ctx, const lmfhd_fab* p, lmfhd_fab* gradp);

/* stochastic_m_fluxdiv (stochastic_m_fluxdiv.f90:42), increment = .false., one RNG stage:
 * W_cc: dm cell arrays of N(0,1) noise; W_ed: 3-D 6 edge arrays (xy c1, xy c2, xz c1, xz c2,
 * yz c1, yz c2), 2-D 2 nodal arrays; eta/temperature at cells and edges (T may be NULL = 1);
 * f: dm face arrays.  The caller fills W as fill_m_stochastic does (variance 2 on the diagonal
 * and copied off-diagonals for stoch_stress_form = 1). */
int lmfhd_stochastic_m_fluxdiv(lmfhd_ctx* ctx, const lmfhd_fab* W_cc, const lmfhd_fab* W_ed,
                               const lmfhd_fab* eta_cc, const lmfhd_fab* eta_ed,
                               const lmfhd_fab* T_cc, const lmfhd_fab* T_ed, double dt,
                               double k_B, double variance_coef_mom, lmfhd_fab* f);

/* average_cc_to_face (convert_stag.f90:132) and average_cc_to_edge / _node (:612, :395);
 * cc must carry one ghost layer holding the boundary values. */
int lmfhd_average_cc_to_face(lmfhd_ctx* ctx, const lmfhd_fab* cc, lmfhd_fab* face);
int lmfhd_average_cc_to_edge(lmfhd_ctx* ctx, const lmfhd_fab* cc, lmfhd_fab* edge);

/* ---- device-resident entry points (serving loop, benchmarks) -----------------------------
 * A device Stokes vector is one contiguous FP64 buffer: the dm face arrays then the cell
 * array, each on its valid region, x fastest.  lmfhd_dev_vec_len gives its length. */
long long lmfhd_dev_vec_len(lmfhd_ctx* ctx);
int lmfhd_gmres_dev(lmfhd_ctx* ctx, const double* b, double* x, lmfhd_stats* stats);
int lmfhd_apply_matrix_dev(lmfhd_ctx* ctx, const double* x, double* b);
int lmfhd_apply_precon_dev(lmfhd_ctx* ctx, const double* b, double* x);
/* The cell fields write_plotfile puts in a plotfile (src_lowMach/write_plotfile.f90:358-392)
 * from a Stokes vector, as consecutive cell arrays: flags bit 0 averaged_vel{x,y,z}
 * (average_face_to_cc, src_common/convert_stag.f90:16), bit 1 shifted_vel{x,y,z}
 * (shift_face_to_cc, :865), bit 2 pres; in that order, dm components each. */
int lmfhd_plot_fields_dev(lmfhd_ctx* ctx, const double* x, int flags, double* out);
/* Device forms of lmfhd_apply_matrix_bc / lmfhd_convert_to_homogeneous (same references):
 * vel_bc_n packed like the faces of a device Stokes vector, vel_bc_t the 2 / 6 wall arrays
 * back to back in the order above, lmfhd_dev_vel_bc_t_len doubles in all. */
int lmfhd_apply_matrix_bc_dev(lmfhd_ctx* ctx, const double* x, double* b, const double* vel_bc_n,
                              const double* vel_bc_t);
int lmfhd_convert_to_homogeneous_dev(lmfhd_ctx* ctx, double* b, const double* vel_bc_n, const double* vel_bc_t);
long long lmfhd_dev_vel_bc_t_len(lmfhd_ctx* ctx);
/* Lphi = (theta alpha - A0) phi on all faces, phi/Lphi: dm face arrays packed as above. */
int lmfhd_stag_applyop_dev(lmfhd_ctx* ctx, const double* phi, double* Lphi);
/* Same operator in the context's multigrid precision on the finest multigrid level
 * (FP32 for LMFHD_MIXED): phi/Lphi are packed face arrays of that precision. */
int lmfhd_stag_applyop_mg_dev(lmfhd_ctx* ctx, const void* phi, void* Lphi);
/* Run nsweeps colour Gauss-Seidel sweeps (2dm colours each) on the finest multigrid level,
 * in place on the level's current phi/rhs: the smoothing loop of stag_mg_solver.f90:360-398
 * on its own, for kernel timing. */
int lmfhd_stag_smooth_dev(lmfhd_ctx* ctx, int nsweeps);
/* Synchronise the context stream. */
int lmfhd_sync(lmfhd_ctx* ctx);

/* Coefficients from one packed device buffer of lmfhd_dev_coef_len doubles:
 * alpha_fc (dm face arrays), beta_cc, beta_ed (3 edge arrays; 2-D: 1 node array), gamma_cc,
 * each on its valid region, x fastest.  Same meaning as lmfhd_set_coeffs (the gmres arguments
 * alpha_fc, beta, beta_ed, gamma, theta_alpha, src_gmres/gmres.f90:32-47); asynchronous on the
 * context stream, and alpha_fc is not checked for zeros. */
long long lmfhd_dev_coef_len(lmfhd_ctx* ctx);
int lmfhd_set_coeffs_dev(lmfhd_ctx* ctx, const double* coef, double theta_alpha);

/* stochastic_m_fluxdiv (src_lowMach/stochastic_m_fluxdiv.f90:42) with the noise generated on
 * the device instead of read from filled arrays: every N(0,1) value is a pure function of
 * (seed, stage, array, position) -- Philox4x64-10 at counter (position, array, stage, 0) with
 * key (seed, 0x243F6A8885A308D3), uniform ((w >> 11) + 1/2) 2^-53, normal quantile; array 0
 * is the cell stress (word = component), array 1 + e the edge / node stress e (word =
 * component).  stoch_stress_form 1 doubles the cell variance and uses word 0 for both edge
 * components (fill_m_stochastic, :781-859).  eta_cc / eta_ed (device; edges packed xy, xz,
 * yz) may be null to use the context's beta_cc / beta_ed.  Writes the faces of the device
 * vector f and zeroes its cell part, or with increment != 0 adds to the faces and leaves the
 * rest (the reference's `increment` argument); asynchronous on the context stream. */
int lmfhd_stochastic_m_fluxdiv_dev(lmfhd_ctx* ctx, unsigned long long seed, unsigned long long stage,
                                   int stoch_stress_form, const double* eta_cc, const double* eta_ed, double dt,
                                   double k_B, double variance_coef_mom, int increment, double* f);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or time-step loops): lmfhd_stochastic_m_fluxdiv_dev, same
 * arguments and same results, except that with filtering_width > 0 the filtered stresses are
 * staged in a fixed 64 MiB device array of the library (one per GPU, kept until the process
 * ends) instead of memory the context allocates on first use; no call allocates.  Limit: the
 * staging must fit, (dm + 2 * n_edge_arrays + 1) slots of the largest cell or edge array
 * (3-D grids up to about 90^3 with filtering); larger grids return LMFHD_E_ARG.  Calls that
 * overlap on one GPU share that array and race. */
int lmfhd_stochastic_m_fluxdiv_dev_noalloc(lmfhd_ctx* ctx, unsigned long long seed, unsigned long long stage,
                                           int stoch_stress_form, const double* eta_cc, const double* eta_ed,
                                           double dt, double k_B, double variance_coef_mom, int increment,
                                           double* f);

/* ---- momentum time steps for one species ----------------------------------------------------
 * The momentum part of advance_timestep_inertial (src_lowMach/advance_timestep_inertial.f90:64)
 * and advance_timestep_overdamped (src_lowMach/advance_timestep_overdamped.f90:50) for a fluid
 * whose density, viscosities and temperature are constant in time, without gravity and with
 * homogeneous velocity boundaries: the mass-flux, scalar-advection and charge parts of those
 * routines vanish there.  x (device Stokes vector) holds v, pi on entry and the new state on
 * return; wall-normal faces of v are set to zero first (multifab_physbc_domainvel).  The
 * context's coefficients must be those of the step's two Stokes solves, as the reference passes
 * them to gmres:
 *   inertial:   alpha_fc = rho on faces, theta_alpha = 1/dt, beta = eta/2, beta_ed = eta_ed/2,
 *               gamma = kappa/2 (the Crank-Nicolson halving of :452-459);
 *   overdamped: theta_alpha = 0, beta = eta, beta_ed = eta_ed, gamma = kappa.
 * f1, f2 (device Stokes vectors, faces used, may be null) are stochastic flux divergences made by
 * lmfhd_stochastic_m_fluxdiv(_dev):
 *   inertial:   f1 = div Sigma^n (weights 1, at dt) for the predictor; the corrector adds
 *               1/2 (f1 + f2) with f2 = div Sigma^n' evaluated with the updated eta (null: f1, as
 *               for constant eta);
 *   overdamped: f1 = div Sigma^(1) for the predictor (weights (1, 0), at dt/2) and f2 for the
 *               corrector (weights (1/sqrt 2, 1/sqrt 2), at dt).
 * The corrector's absolute GMRES tolerance is max(gmres_abs_tol, norm_pre_rhs * gmres_rel_tol)
 * of the predictor (:493-496); stats[0], stats[1] (stats may be null) are the two solves. */
int lmfhd_advance_inertial_dev(lmfhd_ctx* ctx, double dt, double* x, const double* f1, const double* f2,
                               lmfhd_stats* stats);
int lmfhd_advance_overdamped_dev(lmfhd_ctx* ctx, double* x, const double* f1, const double* f2, lmfhd_stats* stats);
/* mk_advective_m_fluxdiv (src_lowMach/mk_advective_m_fluxdiv.f90:18): m_update = -div(m u) with
 * m = rho u (centred fluxes), or m_update minus it with increment; zero on wall faces.  Host
 * fabs: u and rho (face density) as for stag_applyop. */
int lmfhd_advective_m_fluxdiv(lmfhd_ctx* ctx, const lmfhd_fab* u, const lmfhd_fab* rho, lmfhd_fab* m_update,
                              int increment);

/* ---- concurrent solves (serving loop, ensembles) --------------------------------------------
 * Contexts are independent: solves on different contexts run concurrently, each on its own
 * stream, and give the same bits as when run alone (all reductions have a fixed order).
 * The reference solves one system per call (src_gmres/gmres.f90:32); these entry points only
 * schedule several such calls. */
/* Start lmfhd_gmres_dev on a host thread owned by the context and return at once; b and x
 * must stay valid until lmfhd_wait.  LMFHD_E_STATE if a solve is still running there. */
int lmfhd_gmres_dev_async(lmfhd_ctx* ctx, const double* b, double* x);
/* 1 when the context has no asynchronous solve running, else 0. */
int lmfhd_poll(lmfhd_ctx* ctx);
/* Wait for the context's asynchronous solve; its status and statistics. */
int lmfhd_wait(lmfhd_ctx* ctx, lmfhd_stats* stats);
/* n solves on n distinct contexts, at most max_threads at a time (0: 8); stats[q] and
 * rcs[q] (both may be null) per entry; returns the first nonzero status. */
int lmfhd_gmres_dev_batch(lmfhd_ctx** ctxs, int n, const double* const* b, double* const* x, lmfhd_stats* stats,
                          int* rcs, int max_threads);
/* Last error message of this thread. */
const char* lmfhd_last_error(void);

/* ---- launch-parameter autotuner ----------------------------------------------------------------
 * Launch shapes of one size class (grid, boundary conditions, visc_type, precision, krylov,
 * precon_type, visc_schur_approx of a descriptor).  They change how the work is launched, not
 * what is computed: the solutions agree to rounding (the one-workgroup coarse kernels sum the
 * MAC bottom mean in another order) and iteration counts are unchanged.  The reference has no
 * counterpart; its nearest knob is the tiling flag of mfiter_build (stag_applyop.f90:143-201). */
typedef struct lmfhd_tune {
  int coarse_n;        /* staggered multigrid levels with every cell count <= coarse_n run as one
                          one-workgroup kernel per V-cycle (0: none; at most 16) */
  int mac_coarse_n;    /* the same for the MAC pressure multigrid */
  int coarse_threads;  /* threads of that workgroup: 256, 512 or 1024 */
  int red_blocks;      /* workgroups of an inner-product launch: 64 .. 4096 */
  int graphs;          /* 1: HIP-graph replay of Arnoldi steps (taken when lmfhd_desc.graphs = -1) */
  int rows_n;          /* staggered multigrid levels (and stag_applyop) with n_x >= rows_n launch one
                          (j, k) row per 64-lane wave instead of one face per lane (0: never) */
  double ms_per_solve; /* the autotuner's measured time per solve of the winner (0 when set by hand) */
} lmfhd_tune;
/* Time candidate launch parameters on a synthetic problem of the descriptor's size class (reps
 * timed solves per candidate after one warm-up; coordinate search from the defaults), record
 * the fastest for the class and return it.  Contexts created later for the class use it. */
int lmfhd_autotune(const lmfhd_desc* d, int reps, lmfhd_tune* out);
/* Record parameters for the descriptor's size class by hand (e.g. from an earlier run). */
int lmfhd_tune_set(const lmfhd_desc* d, const lmfhd_tune* t);
/* The record of the class in *t (1), or the defaults (0); negative on a bad argument. */
int lmfhd_tune_get(const lmfhd_desc* d, lmfhd_tune* t);
/* The parameters a context runs with. */
int lmfhd_ctx_tune(lmfhd_ctx* ctx, lmfhd_tune* t);

#ifdef __cplusplus
}
#endif
#endif
