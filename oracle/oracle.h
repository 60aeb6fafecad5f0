/*
 * oracle.h -- FP64 CPU restatement of LowMachFHD's implicit staggered Stokes solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path
 * (lowmachfhd_amd/csrc) and the "port" CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * never links or calls it.
 *
 * Data exchanged through this interface uses the "valid-region" layout that the HIP
 * library also uses: every array holds exactly the reference's valid (non-ghost) index
 * range of one box covering the whole domain, x fastest (Fortran order), with nodal
 * directions storing n+1 points (the duplicated periodic face included).  Internally
 * the oracle copies into ghosted fabs and fills ghost cells exactly as the reference
 * does (periodic images, then the physical-boundary rules).
 */
#ifndef LMFHD_ORACLE_H
#define LMFHD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int dm;              /* 2 or 3 */
  int n[3];            /* cells per direction (n[2] ignored in 2-D) */
  double dx;           /* uniform spacing (reference requires dx=dy=dz) */
  double dz2d;         /* depth used for the 2-D stochastic cell volume (prob_hi-prob_lo)/n in z */
  int bc_lo[3], bc_hi[3]; /* physical BC codes: -1 periodic, 100..103 no-slip, 200..203 slip */
  int visc_type;       /* -3..3 except 0 */
  /* staggered velocity multigrid (probin_gmres.f90 defaults in brackets) */
  int stag_nu_down, stag_nu_up, stag_nu_bottom; /* [2,2,8] */
  int stag_minwidth;   /* [2] */
  int stag_max_vcycles;/* [1] */
  int stag_smoother;   /* 0 Jacobi, 1 colour GS [1] */
  double stag_omega;   /* [1] */
  double stag_rel_tol; /* [1e-9] */
  /* cell-centred MAC pressure multigrid (this build's own design) */
  int mac_nu_down, mac_nu_up, mac_nu_bottom, mac_minwidth, mac_max_vcycles;
  /* GMRES */
  int gmres_max_inner, gmres_max_outer, gmres_max_iter, gmres_min_iter;
  double gmres_rel_tol, gmres_abs_tol, p_norm_weight;
  int precon_type;     /* +-1 .. +-5 (apply_precon.f90:91-426, visc_schur_approx = 0) */
  int omp_threads;     /* 0 = OpenMP default */
  int gmres_spatial_order; /* 2, or 4: apply_matrix with stag_applyop_4o + compute_grad_4o */
  int visc_schur_approx;   /* 0, or 1: visc_schur_complement in apply_precon */
  int filtering_width;     /* 0, 1, 2 or 4: filter_stoch_m_flux in stochastic_m_fluxdiv */
} orc_params;

typedef struct {
  int total_iter;        /* inner iterations (Krylov vectors built) */
  int outer_iter;        /* outer (restart) iterations */
  int converged;         /* 1 if the true preconditioned residual met the tolerance */
  int n_precon;          /* apply_precon calls */
  int n_matrix;          /* apply_matrix calls */
  double norm_b, norm_pre_b, norm_init_resid, norm_resid, norm_resid_stokes;
  double resid_hist[256];/* |M^-1 r| estimate after each inner iteration */
} orc_stats;

void orc_default_params(orc_params* p, int dm, int nx, int ny, int nz, double dx);
void* orc_create(const orc_params* p);
void orc_destroy(void* ctx);
int orc_set_coeffs(void* ctx, const double* const* alpha_fc, const double* beta_cc,
                   const double* const* beta_ed, const double* gamma_cc, double theta_alpha);

/* Lphi = (theta*alpha - A0) phi on every face, colour 0..2dm (stag_applyop.f90:20,70) */
int orc_stag_applyop(void* ctx, const double* const* phi, double* const* Lphi, int color);
/* b_u = A x_u + G x_p ; b_p = -D x_u (apply_matrix.f90:24) */
int orc_apply_matrix(void* ctx, const double* const* x_u, const double* x_p,
                     double* const* b_u, double* b_p);
/* apply_matrix with inhomogeneous velocity boundary values vel_bc_n (dm face arrays, read on
 * wall faces) and vel_bc_t (2-D: 2 node arrays, 3-D: 6 edge arrays in the reference's order:
 * y-, z-velocity on x-walls, x-, z-velocity on y-walls, x-, y-velocity on z-walls);
 * gmres_spatial_order 2 only (returns 1 otherwise) */
int orc_apply_matrix_bc(void* ctx, const double* const* x_u, const double* x_p, double* const* b_u, double* b_p,
                        const double* const* vel_bc_n, const double* const* vel_bc_t);
/* convert_to_homogeneous: b_u -= L phi, b_p += D phi for phi = 0 carrying those values, in place */
int orc_convert_to_homogeneous(void* ctx, double* const* b_u, double* b_p, const double* const* vel_bc_n,
                               const double* const* vel_bc_t);
/* x = M^-1 b, projection preconditioner (apply_precon.f90:29) */
int orc_apply_precon(void* ctx, const double* const* b_u, const double* b_p,
                     double* const* x_u, double* x_p);
/* phi <- staggered MG applied to (theta alpha - A0) phi = rhs, phi is the initial guess */
int orc_stag_mg_solve(void* ctx, const double* const* rhs, double* const* phi);
/* phi <- cell-centred MG for D(alpha_fc^-1 G) phi = rhs, phi starts at zero */
int orc_mac_mg_solve(void* ctx, const double* rhs, double* phi);
/* left-preconditioned restarted GMRES (gmres.f90:32); x is initial guess / solution */
int orc_gmres(void* ctx, const double* const* b_u, const double* b_p,
              double* const* x_u, double* x_p, orc_stats* st);
int orc_compute_div(void* ctx, const double* const* u, double* div);
int orc_compute_grad(void* ctx, const double* p, double* const* g);
/* weighted staggered inner product per component + cc inner product (norm_inner_product.f90) */
int orc_inner_prod(void* ctx, const double* const* u1, const double* p1,
                   const double* const* u2, const double* p2, double* out_u, double* out_p);
/* stochastic momentum flux divergence (stochastic_m_fluxdiv.f90:42), increment=0.
 * W_cc: dm cell arrays; W_ed: 3-D: 6 arrays (xy c1,c2, xz c1,c2, yz c1,c2); 2-D: node c1,c2.
 * eta_cc/T_cc cell arrays; eta_ed/T_ed edge (3-D, 3 arrays) or node (2-D, 1 array).
 * T_cc or T_ed may be NULL (temperature 1). */
int orc_stochastic_m_fluxdiv(void* ctx, const double* const* W_cc, const double* const* W_ed,
                             const double* eta_cc, const double* const* eta_ed,
                             const double* T_cc, const double* const* T_ed,
                             double dt, double kT, double variance_coef,
                             double* const* f);
/* cc -> face (2-point mean) and cc -> edge/node (4-point mean); cc includes one ghost layer
 * given as a (n+2)^dm array (ghost values supplied by the caller) (convert_stag.f90:132,395,612) */
int orc_average_cc_to_face(void* ctx, const double* cc_g1, double* const* face);
int orc_average_cc_to_edge(void* ctx, const double* cc_g1, double* const* edge);
int orc_nlevs_mg(void* ctx, int minwidth);
/* multigrid pieces on the context's grid (level 0 coefficients): one smoother colour, residual
 * (or simple) restriction to the grid coarsened by 2, prolongation-add from it */
int orc_stag_smooth_colour(void* ctx, double* const* phi, const double* const* rhs, int color);
int orc_stag_restriction(void* ctx, const double* const* fine, double* const* coarse, int simple);
int orc_stag_prolongation(void* ctx, const double* const* coarse, double* const* fine);

#ifdef __cplusplus
}
#endif
#endif
