! wcns_shims.f90 -- the reference's hot-path subroutines as thin shims over the
! C-ABI, so that the driver loop of src/Main.f90:85-109 calls them unchanged:
!   boundary_all        sub_tgh.f90:123-186      initial_q_dq_c_t  sub_tgh.f90:1-32
!   timestep_tgh        sub_tgh.f90:1270-1348    r_h_s_vis         RHS.f90:2-30
!   communicate_dq_npp  parallel.f90:3020-3164   r_h_s_invis       RHS.f90:35-72
!   l_h_s_tgh           LHS.f90:1-34             communicate_pv_npp parallel.f90:3166-3326
!   residual_curve      PostProcess.f90:1195-1347
! They act on the context wcns_ctx_h of module wcns_amd_c.  residual_curve
! leaves the monitor in module wcns_monitor (the reference keeps it in
! resdual_module: dres_add, resmax_total, n_error).
module wcns_monitor
  use iso_c_binding
  implicit none
  real(c_double), save :: dres_add = 0.0d0, resmax_total = 0.0d0
  integer(c_int), save :: n_error(5) = 0
end module wcns_monitor

subroutine boundary_all
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_boundary_all(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('boundary_all', rc)
end subroutine boundary_all

subroutine initial_q_dq_c_t
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_initial_q_dq_c_t(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('initial_q_dq_c_t', rc)
end subroutine initial_q_dq_c_t

subroutine timestep_tgh
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_timestep(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('timestep_tgh', rc)
end subroutine timestep_tgh

subroutine r_h_s_vis
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_r_h_s_vis(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('r_h_s_vis', rc)
end subroutine r_h_s_vis

subroutine r_h_s_invis
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_r_h_s_invis(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('r_h_s_invis', rc)
end subroutine r_h_s_invis

subroutine communicate_dq_npp
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_communicate_dq(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('communicate_dq_npp', rc)
end subroutine communicate_dq_npp

subroutine l_h_s_tgh
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_l_h_s(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('l_h_s_tgh', rc)
end subroutine l_h_s_tgh

subroutine communicate_pv_npp
  use wcns_amd_c
  integer(c_int) :: rc
  rc = wcns_communicate_pv(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('communicate_pv_npp', rc)
end subroutine communicate_pv_npp

subroutine residual_curve(nbg, ngzone)
  use wcns_amd_c
  use wcns_monitor
  integer, intent(in) :: nbg, ngzone
  integer(c_int) :: rc
  rc = wcns_residual(wcns_ctx_h, dres_add, resmax_total, n_error)
  if (rc /= 0) call wcns_abort('residual_curve', rc)
end subroutine residual_curve
