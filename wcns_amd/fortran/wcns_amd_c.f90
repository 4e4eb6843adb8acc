! wcns_amd_c.f90 -- ISO_C_BINDING interface of libwcns_amd.so (include/wcns_amd.h)
! for the reference's Fortran host.  One interface per C entry point, with the
! C argument meaning; wcns_params matches the C struct member for member.
module wcns_amd_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WCNS_OK = 0, WCNS_E_ARG = -1, WCNS_E_STATE = -2, WCNS_E_HIP = -3
  integer(c_int), parameter :: WCNS_E_NOMEM = -4, WCNS_E_BC = -5, WCNS_E_JACOBIAN = -6, WCNS_E_ROE = -7
  integer(c_int), parameter :: WCNS_E_REPAIR = -8, WCNS_E_UNSUPPORTED = -9
  integer(c_int), parameter :: WCNS_LHS_FP64 = 0, WCNS_LHS_FP32 = 1
  integer(c_int), parameter :: WCNS_TEC_SURFACE = 0, WCNS_TEC_WALL = 1, WCNS_TEC_VOLUME = 2

  type, bind(C) :: wcns_params
     real(c_double) :: moo, reynolds, attack, sideslip, tref, twall
     real(c_double) :: cfl, timedt_rate, efix, csrv, gama, prl, prt
     integer(c_int) :: nvis, nlamtur, nscheme, nflux, ntmst
     real(c_double) :: sref, lfref, xref, yref, zref
     integer(c_int) :: nwholefield
     real(c_double) :: timedt_turb, kmaxlim, xk, xb, lref
  end type wcns_params

  interface
     integer(c_int) function wcns_create(ctx, device, prm) bind(C, name="wcns_create")
       import :: c_int, c_ptr, wcns_params
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: device
       type(wcns_params), intent(in) :: prm
     end function
     integer(c_int) function wcns_destroy(ctx) bind(C, name="wcns_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     type(c_ptr) function wcns_last_error(ctx) bind(C, name="wcns_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_add_block(ctx, ni, nj, nk, x, y, z, nb) bind(C, name="wcns_add_block")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ni, nj, nk
       real(c_double), intent(in) :: x(*), y(*), z(*)
       integer(c_int), intent(out) :: nb
     end function
     integer(c_int) function wcns_add_region(ctx, nb, s_raw, bctype, t_raw, nbt, ibcwin) &
          bind(C, name="wcns_add_region")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nb, bctype, nbt, ibcwin
       integer(c_int), intent(in) :: s_raw(6), t_raw(6)
     end function
     integer(c_int) function wcns_setup(ctx) bind(C, name="wcns_setup")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_set_lhs_precision(ctx, prec) bind(C, name="wcns_set_lhs_precision")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: prec
     end function
     integer(c_int) function wcns_set_prim(ctx, nb, r, u, v, w, p) bind(C, name="wcns_set_prim")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nb
       real(c_double), intent(in) :: r(*), u(*), v(*), w(*), p(*)
     end function
     integer(c_int) function wcns_get_field(ctx, nb, which, out) bind(C, name="wcns_get_field")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nb, which
       real(c_double), intent(out) :: out(*)
     end function
     integer(c_int) function wcns_set_field(ctx, nb, which, in) bind(C, name="wcns_set_field")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nb, which
       real(c_double), intent(in) :: in(*)
     end function
     integer(c_int) function wcns_boundary_all(ctx) bind(C, name="wcns_boundary_all")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_initial_q_dq_c_t(ctx) bind(C, name="wcns_initial_q_dq_c_t")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_timestep(ctx) bind(C, name="wcns_timestep")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_r_h_s_vis(ctx) bind(C, name="wcns_r_h_s_vis")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_r_h_s_invis(ctx) bind(C, name="wcns_r_h_s_invis")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_communicate_dq(ctx) bind(C, name="wcns_communicate_dq")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_l_h_s(ctx) bind(C, name="wcns_l_h_s")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_communicate_pv(ctx) bind(C, name="wcns_communicate_pv")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_residual(ctx, rms, resmax, loc) bind(C, name="wcns_residual")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: rms, resmax
       integer(c_int), intent(out) :: loc(5)
     end function
     integer(c_int) function wcns_residual_noalloc(ctx, rms, resmax, loc) bind(C, name="wcns_residual_noalloc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: rms, resmax
       integer(c_int), intent(out) :: loc(5)
     end function
     integer(c_int) function wcns_step(ctx, nsteps, rms_hist) bind(C, name="wcns_step")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       type(c_ptr), value :: rms_hist
     end function
     integer(c_int) function wcns_step_graph(ctx, nsteps) bind(C, name="wcns_step_graph")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function
     integer(c_int) function wcns_residual_history(ctx, nmax, rms, resmax, loc, count) &
          bind(C, name="wcns_residual_history")
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: nmax
       real(c_double), intent(out) :: rms(*), resmax(*)
       integer(c_int), intent(out) :: loc(5, *)
       integer(c_long_long), intent(out) :: count
     end function
     integer(c_int) function wcns_autotune(ctx, reps, tj, tk, ms) bind(C, name="wcns_autotune")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: reps
       integer(c_int), intent(out) :: tj, tk
       real(c_double), intent(out) :: ms
     end function
     integer(c_int) function wcns_sweep_layout(ctx, tj, tk, paired) bind(C, name="wcns_sweep_layout")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: tj, tk, paired
     end function
     integer(c_int) function wcns_force(ctx, out) bind(C, name="wcns_force")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(9)
     end function
     integer(c_int) function wcns_force_record(ctx) bind(C, name="wcns_force_record")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_force_history(ctx, nmax, out, count) bind(C, name="wcns_force_history")
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: nmax
       real(c_double), intent(out) :: out(9, *)
       integer(c_long_long), intent(out) :: count
     end function
     integer(c_int) function wcns_save_flow(ctx, flowname, nout, wholetime) bind(C, name="wcns_save_flow")
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: flowname(*)
       integer(c_int), value :: nout
       real(c_double), value :: wholetime
     end function
     ! Tecplot binary output of post_processing's nplot choices (PostProcess.f90:37-55)
     integer(c_int) function wcns_tecplot(ctx, path, kind) bind(C, name="wcns_tecplot")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: kind
     end function
     integer(c_int) function wcns_load_flow(ctx, flowname, t_from_file, nout, wholetime) &
          bind(C, name="wcns_load_flow")
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: flowname(*)
       integer(c_int), value :: t_from_file
       integer(c_int), intent(out) :: nout
       real(c_double), intent(out) :: wholetime
     end function
     integer(c_int) function wcns_get_counter(ctx, which, value) bind(C, name="wcns_get_counter")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: which
       integer(c_long_long), intent(out) :: value
     end function
     integer(c_int) function wcns_sync(ctx) bind(C, name="wcns_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function wcns_timing(ctx, on) bind(C, name="wcns_timing")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function
     integer(c_int) function wcns_timing_get(ctx, slot, total_ms, launches) bind(C, name="wcns_timing_get")
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: slot
       real(c_double), intent(out) :: total_ms
       integer(c_long_long), intent(out) :: launches
     end function
     ! tangent / adjoint: every array argument is a DEVICE address
     integer(c_int) function wcns_adj_residual(ctx, nb, p, x, r) bind(C, name="wcns_adj_residual")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, p, x, r
       integer(c_int), value :: nb
     end function
     integer(c_int) function wcns_adj_jvp(ctx, nb, p, x, dp, dx, r, dr) bind(C, name="wcns_adj_jvp")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, p, x, dp, dx, r, dr
       integer(c_int), value :: nb
     end function
     integer(c_int) function wcns_adj_vjp(ctx, nb, p, x, w, gp, gx) bind(C, name="wcns_adj_vjp")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, p, x, w, gp, gx
       integer(c_int), value :: nb
     end function
     integer(c_int) function wcns_adj_force(ctx, nb, p, x, omega, j, gp, gx) bind(C, name="wcns_adj_force")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, p, x, gp, gx
       integer(c_int), value :: nb
       real(c_double), intent(in) :: omega(6)
       real(c_double), intent(out) :: j
     end function
     integer(c_int) function wcns_region_images(ctx, nb, nr, img, jmg, kmg, box) bind(C, name="wcns_region_images")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nb, nr
       integer(c_int), intent(out) :: img(*), jmg(*), kmg(*), box(6)
     end function
     integer(c_int) function wcns_multipoint(ctx, ngroups, npts, pts, cap) bind(C, name="wcns_multipoint")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: ngroups, npts(*), pts(*)
       integer(c_int), value :: cap
     end function
  end interface

  ! the context the shims of wcns_shims.f90 work on
  type(c_ptr), save :: wcns_ctx_h = c_null_ptr

contains

  ! wcns_last_error as a Fortran string
  function wcns_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: l
    msg = ' '
    p = wcns_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    l = 0
    do while (l < 256)
      if (s(l + 1) == c_null_char) exit
      l = l + 1
      msg(l:l) = s(l)
    end do
  end function wcns_error_text

  ! The reference's convention for fatal conditions: print and stop.
  subroutine wcns_abort(where, rc)
    character(len=*), intent(in) :: where
    integer(c_int), intent(in) :: rc
    write(*, '(a,a,a,i4,a,a)') ' wcns_amd: ', where, ' failed (', rc, '): ', trim(wcns_error_text(wcns_ctx_h))
    stop 1
  end subroutine wcns_abort
end module wcns_amd_c
