! lmfhd_c.f90 -- Fortran binding (iso_c_binding) of the C ABI in lmfhd.h, for a LowMachFHD /
! FBoxLib host.  A maintainer adds it to staggered_grid/src_gmres and links liblmfhd.so; the
! reference's gmres (src_gmres/gmres.f90:32-47), apply_matrix (apply_matrix.f90:24-40),
! apply_precon (apply_precon.f90:29-43), stag_mg_solver (stag_mg_solver.f90:37-50),
! stag_applyop (stag_applyop.f90:20-31) and stochastic_m_fluxdiv (stochastic_m_fluxdiv.f90:42-54)
! then call the entry points below on fab views of their multifabs, with no repacking.
!
! Array arguments: one lmfhd_fab per box (dataptr(mf, i) with the valid bounds of the box,
! nodal +1 included, and the ghost width), and for an argument of several multifabs (the dm
! face components, the edge arrays) nboxes views per multifab, multifab-major -- exactly what
! mf_views below fills.  The box list goes into lmfhd_desc (nboxes, box_lo, box_hi) once, when
! the context is created (lmfhd_ctx_create replaces stag_mg_layout_build and
! mgt_macproj_precon_build).  Single process: the views cover every box of the boxarray.
module lmfhd_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LMFHD_OK = 0, LMFHD_E_ARG = -1, LMFHD_E_NOMEM = -2, &
       LMFHD_E_BREAKDOWN = -3, LMFHD_E_NAN = -4, LMFHD_E_HIP = -5, LMFHD_E_STATE = -6, LMFHD_E_BUDGET = -7
  integer(c_int), parameter :: LMFHD_FP64 = 0, LMFHD_MIXED = 1
  integer(c_int), parameter :: LMFHD_GMRES_LEFT = 0, LMFHD_FGMRES = 1

  type, bind(c) :: lmfhd_fab
     type(c_ptr)    :: data            ! element (lo-ng, lo-ng, lo-ng)
     integer(c_int) :: lo(3), hi(3)    ! valid bounds, nodal +1 included
     integer(c_int) :: ng
  end type lmfhd_fab

  type, bind(c) :: lmfhd_desc
     integer(c_int) :: dm, n_cells(3)
     real(c_double) :: dx, dz2d
     integer(c_int) :: bc_lo(3), bc_hi(3), visc_type, precision, krylov, precon_type
     integer(c_int) :: gmres_max_inner, gmres_max_outer, gmres_max_iter, gmres_min_iter
     real(c_double) :: gmres_rel_tol, gmres_abs_tol, p_norm_weight
     integer(c_int) :: stag_mg_nsmooths_down, stag_mg_nsmooths_up, stag_mg_nsmooths_bottom
     integer(c_int) :: stag_mg_minwidth, stag_mg_max_vcycles, stag_mg_smoother
     real(c_double) :: stag_mg_omega, stag_mg_rel_tol
     integer(c_int) :: mg_nsmooths_down, mg_nsmooths_up, mg_nsmooths_bottom, mg_minwidth, mg_max_vcycles
     integer(c_int) :: device, gmres_spatial_order, visc_schur_approx, filtering_width, graphs, batch
     integer(c_int) :: nboxes
     type(c_ptr)    :: box_lo, box_hi  ! c_loc of integer(c_int) arrays (3, nboxes)
  end type lmfhd_desc

  type, bind(c) :: lmfhd_stats
     integer(c_int) :: iters, outer_iters, converged, n_precon, n_matrix
     real(c_double) :: norm_b, norm_pre_b, norm_init_resid, norm_resid, norm_resid_stokes
     real(c_double) :: true_rel_resid, resid_hist(256)
  end type lmfhd_stats

  type, bind(c) :: lmfhd_tune
     integer(c_int) :: coarse_n, mac_coarse_n, coarse_threads, red_blocks, graphs, rows_n
     real(c_double) :: ms_per_solve
  end type lmfhd_tune

  interface
     subroutine lmfhd_desc_defaults(d) bind(c)
       import; type(lmfhd_desc) :: d
     end subroutine lmfhd_desc_defaults
     integer(c_int) function lmfhd_ctx_create(d, ctx) bind(c)
       import; type(lmfhd_desc), intent(in) :: d; type(c_ptr) :: ctx
     end function lmfhd_ctx_create
     integer(c_int) function lmfhd_ctx_destroy(ctx) bind(c)
       import; type(c_ptr), value :: ctx
     end function lmfhd_ctx_destroy
     integer(c_int) function lmfhd_set_coeffs(ctx, alpha_fc, beta_cc, beta_ed, gamma_cc, theta_alpha) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: alpha_fc(*), beta_cc(*), beta_ed(*), gamma_cc(*)
       real(c_double), value :: theta_alpha
     end function lmfhd_set_coeffs
     integer(c_int) function lmfhd_stag_applyop(ctx, phi, Lphi, color) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: phi(*); type(lmfhd_fab) :: Lphi(*)
       integer(c_int), value :: color
     end function lmfhd_stag_applyop
     integer(c_int) function lmfhd_apply_matrix(ctx, x_u, x_p, b_u, b_p) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: x_u(*), x_p(*); type(lmfhd_fab) :: b_u(*), b_p(*)
     end function lmfhd_apply_matrix
     integer(c_int) function lmfhd_apply_matrix_bc(ctx, x_u, x_p, b_u, b_p, vel_bc_n, vel_bc_t) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: x_u(*), x_p(*), vel_bc_n(*), vel_bc_t(*)
       type(lmfhd_fab) :: b_u(*), b_p(*)
     end function lmfhd_apply_matrix_bc
     integer(c_int) function lmfhd_convert_to_homogeneous(ctx, b_u, b_p, vel_bc_n, vel_bc_t) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab) :: b_u(*), b_p(*)
       type(lmfhd_fab), intent(in) :: vel_bc_n(*), vel_bc_t(*)
     end function lmfhd_convert_to_homogeneous
     integer(c_int) function lmfhd_apply_precon(ctx, b_u, b_p, x_u, x_p) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: b_u(*), b_p(*); type(lmfhd_fab) :: x_u(*), x_p(*)
     end function lmfhd_apply_precon
     integer(c_int) function lmfhd_stag_mg_solver(ctx, rhs, phi) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: rhs(*); type(lmfhd_fab) :: phi(*)
     end function lmfhd_stag_mg_solver
     integer(c_int) function lmfhd_mac_mg_solver(ctx, rhs, phi) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: rhs(*); type(lmfhd_fab) :: phi(*)
     end function lmfhd_mac_mg_solver
     integer(c_int) function lmfhd_gmres(ctx, b_u, b_p, x_u, x_p, norm_pre_rhs, stats) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: b_u(*), b_p(*); type(lmfhd_fab) :: x_u(*), x_p(*)
       real(c_double) :: norm_pre_rhs
       type(lmfhd_stats) :: stats
     end function lmfhd_gmres
     integer(c_int) function lmfhd_compute_div(ctx, u, div) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: u(*); type(lmfhd_fab) :: div(*)
     end function lmfhd_compute_div
     integer(c_int) function lmfhd_compute_grad(ctx, p, gradp) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: p(*); type(lmfhd_fab) :: gradp(*)
     end function lmfhd_compute_grad
     integer(c_int) function lmfhd_stochastic_m_fluxdiv(ctx, W_cc, W_ed, eta_cc, eta_ed, T_cc, T_ed, dt, k_B, &
          variance_coef_mom, f) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: W_cc(*), W_ed(*), eta_cc(*), eta_ed(*)
       type(c_ptr), value :: T_cc, T_ed   ! c_loc of view arrays, or c_null_ptr (temperature 1)
       real(c_double), value :: dt, k_B, variance_coef_mom
       type(lmfhd_fab) :: f(*)
     end function lmfhd_stochastic_m_fluxdiv
     integer(c_int) function lmfhd_average_cc_to_face(ctx, cc, face) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: cc(*); type(lmfhd_fab) :: face(*)
     end function lmfhd_average_cc_to_face
     integer(c_int) function lmfhd_average_cc_to_edge(ctx, cc, edge) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: cc(*); type(lmfhd_fab) :: edge(*)
     end function lmfhd_average_cc_to_edge
     integer(c_int) function lmfhd_advective_m_fluxdiv(ctx, u, rho, m_update, increment) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: u(*), rho(*); type(lmfhd_fab) :: m_update(*)
       integer(c_int), value :: increment
     end function lmfhd_advective_m_fluxdiv
     integer(c_int) function lmfhd_plot_fields(ctx, x_u, x_p, flags, out) bind(c)
       import; type(c_ptr), value :: ctx
       type(lmfhd_fab), intent(in) :: x_u(*), x_p(*); integer(c_int), value :: flags
       type(lmfhd_fab) :: out(*)
     end function lmfhd_plot_fields
     integer(c_int) function lmfhd_autotune(d, reps, out) bind(c)
       import; type(lmfhd_desc), intent(in) :: d; integer(c_int), value :: reps; type(lmfhd_tune) :: out
     end function lmfhd_autotune
     type(c_ptr) function lmfhd_last_error() bind(c)
       import
     end function lmfhd_last_error
  end interface

contains

  ! The box list of a layout into the descriptor (boxes must outlive lmfhd_ctx_create, which
  ! copies them).  FBoxLib cell boxes are 0-based in the domain, as lmfhd_desc expects.
  subroutine desc_boxes(d, la, box_lo, box_hi)
    use layout_module
    use box_module
    type(lmfhd_desc), intent(inout) :: d
    type(layout), intent(in) :: la
    integer(c_int), target, intent(inout) :: box_lo(:,:), box_hi(:,:)   ! (3, nboxes(la))
    integer :: i, dm
    type(box) :: bx
    dm = get_dim(la)
    box_lo = 0; box_hi = 0
    do i = 1, nboxes(la)
       bx = get_box(la, i)
       box_lo(1:dm, i) = lwb(bx); box_hi(1:dm, i) = upb(bx)
    end do
    d%nboxes = nboxes(la)
    d%box_lo = c_loc(box_lo); d%box_hi = c_loc(box_hi)
  end subroutine desc_boxes

  ! Views of every box of multifabs mf(1:nmf) (e.g. the dm face components of a velocity), in
  ! the order the C ABI reads them: views((m - 1) * nb + i) is box i of mf(m).
  subroutine mf_views(mf, views)
    use multifab_module
    type(multifab), intent(in) :: mf(:)
    type(lmfhd_fab), intent(out) :: views(:)
    real(kind=dp_t), pointer :: p(:,:,:,:)
    integer :: m, i, nb, dm
    integer, allocatable :: lo(:), hi(:)
    nb = nfabs(mf(1))
    do m = 1, size(mf)
       dm = get_dim(mf(m))
       allocate(lo(dm), hi(dm))
       do i = 1, nb
          p => dataptr(mf(m), i)
          lo = lwb(get_ibox(mf(m), i)); hi = upb(get_ibox(mf(m), i))
          associate (v => views((m - 1) * nb + i))
            v%data = c_loc(p(lbound(p, 1), lbound(p, 2), lbound(p, 3), 1))
            v%lo = 0; v%hi = 0
            v%lo(1:dm) = lo; v%hi(1:dm) = hi
            v%ng = nghost(mf(m))
          end associate
       end do
       deallocate(lo, hi)
    end do
  end subroutine mf_views

end module lmfhd_c
