! ref_sa_bridge.f90 -- TEST INFRASTRUCTURE ONLY, never part of the library.
!
! Loads one block handed over from C into the reference's module variables
! (global_variables, Globalvar.f90:284-423, with the bounds of
! PreProcess.f90:744-807) and runs the reference's own per-block
! turbulence-model step, turbulence(nb) (turblence.f90:15-49), with
! nameturb = 'SA' (ref_sa_step) or 'SST' (ref_sst_step), ncmpcor = 0, ndes = 0,
! ntrans = 0, nwtmax = 0, method = 1.
! oracle/Makefile links it with the reference's turblence.f90, Sub1.f90,
! Sub2.f90, Convect.f90, BL.f90 and Globalvar.f90, compiled from source in the
! reference's serial configuration (PARALLEL undefined), into _ref/libref_sa.so.
! The tests compare the C restatement of the same step against it.
!
! Every field comes in the C oracle's frame: (ni+6)*(nj+6)*(nk+6) doubles,
! i fastest, index 0 = node -2.
module ref_sa_store
  implicit none
  real, target, allocatable, save :: s_r(:,:,:), s_u(:,:,:), s_v(:,:,:), s_w(:,:,:)
  real, target, allocatable, save :: s_qke(:,:,:,:), s_vist(:,:,:), s_visl(:,:,:)
  real, target, allocatable, save :: s_vol(:,:,:), s_dtdt(:,:,:), s_dist(:,:,:)
  real, target, allocatable, save :: s_kcx(:,:,:), s_kcy(:,:,:), s_kcz(:,:,:)
  real, target, allocatable, save :: s_etx(:,:,:), s_ety(:,:,:), s_etz(:,:,:)
  real, target, allocatable, save :: s_ctx(:,:,:), s_cty(:,:,:), s_ctz(:,:,:)
  real, target, allocatable, save :: s_p(:,:,:), s_kct(:,:,:), s_ett(:,:,:), s_ctt(:,:,:)
  real, target, allocatable, save :: s_c(:,:,:)
contains
  subroutine release()
    if (allocated(s_r)) deallocate(s_r, s_u, s_v, s_w, s_qke, s_vist, s_visl, s_vol, &
         s_dtdt, s_dist, s_kcx, s_kcy, s_kcz, s_etx, s_ety, s_etz, s_ctx, s_cty, s_ctz, &
         s_p, s_kct, s_ett, s_ctt, s_c)
  end subroutine release

  ! copy nodes lo..n+hi of a C-frame field into dst(lo:ni+hi, lo:nj+hi, lo:nk+hi)
  subroutine load(dst, src, ni, nj, nk, lo, hi)
    integer, intent(in) :: ni, nj, nk, lo, hi
    real, intent(in) :: src(-2:ni+3, -2:nj+3, -2:nk+3)
    real, intent(out) :: dst(lo:ni+hi, lo:nj+hi, lo:nk+hi)
    dst = src(lo:ni+hi, lo:nj+hi, lo:nk+hi)
  end subroutine load
end module ref_sa_store

! Loads one block: fld = pointers to r, u, v, w, qke(1), vist, visl, vol, dtdt,
! dist, kcx, kcy, kcz, etx, ety, etz, ctx, cty, ctz, p and, for nl = 2, qke(2)
! (C frame); the grid is fixed, so the time metrics kct, ett, ctt are zero.
subroutine ref_turb_load(ni, nj, nk, nl, fld)
  use iso_c_binding
  use global_variables, only: g_ni => ni, g_nj => nj, g_nk => nk, nmax, r, u, v, w, qke, vist, visl, &
       vol, dtdt, ds_turbulence, kcx, kcy, kcz, etx, ety, etz, ctx, cty, ctz, p, kct, ett, ctt, c
  use ref_sa_store
  implicit none
  integer, intent(in) :: ni, nj, nk, nl
  type(c_ptr), intent(in) :: fld(*)
  real(c_double), pointer :: f(:,:,:)

  call release()
  allocate(s_r(-2:ni+3, -2:nj+3, -2:nk+3), s_u(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(s_v(-2:ni+3, -2:nj+3, -2:nk+3), s_w(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(s_qke(-2:ni+3, -2:nj+3, -2:nk+3, nl), s_vist(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(s_visl(-1:ni+1, -1:nj+1, -1:nk+1), s_vol(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(s_dtdt(1:ni, 1:nj, 1:nk), s_dist(1:ni, 1:nj, 1:nk))
  allocate(s_kcx(-1:ni+1, -1:nj+1, -1:nk+1), s_kcy(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(s_kcz(-1:ni+1, -1:nj+1, -1:nk+1), s_etx(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(s_ety(-1:ni+1, -1:nj+1, -1:nk+1), s_etz(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(s_ctx(-1:ni+1, -1:nj+1, -1:nk+1), s_cty(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(s_ctz(-1:ni+1, -1:nj+1, -1:nk+1), s_p(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(s_kct(-1:ni+1, -1:nj+1, -1:nk+1), s_ett(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(s_ctt(-1:ni+1, -1:nj+1, -1:nk+1), s_c(-1:ni+1, -1:nj+1, -1:nk+1))
  s_kct = 0.0; s_ett = 0.0; s_ctt = 0.0; s_c = 0.0

  call c_f_pointer(fld(1), f, [ni + 6, nj + 6, nk + 6]); call load(s_r, f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(2), f, [ni + 6, nj + 6, nk + 6]); call load(s_u, f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(3), f, [ni + 6, nj + 6, nk + 6]); call load(s_v, f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(4), f, [ni + 6, nj + 6, nk + 6]); call load(s_w, f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(5), f, [ni + 6, nj + 6, nk + 6]); call load(s_qke(:,:,:,1), f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(6), f, [ni + 6, nj + 6, nk + 6]); call load(s_vist, f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(7), f, [ni + 6, nj + 6, nk + 6]); call load(s_visl, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(8), f, [ni + 6, nj + 6, nk + 6]); call load(s_vol, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(9), f, [ni + 6, nj + 6, nk + 6]); call load(s_dtdt, f, ni, nj, nk, 1, 0)
  call c_f_pointer(fld(10), f, [ni + 6, nj + 6, nk + 6]); call load(s_dist, f, ni, nj, nk, 1, 0)
  call c_f_pointer(fld(11), f, [ni + 6, nj + 6, nk + 6]); call load(s_kcx, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(12), f, [ni + 6, nj + 6, nk + 6]); call load(s_kcy, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(13), f, [ni + 6, nj + 6, nk + 6]); call load(s_kcz, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(14), f, [ni + 6, nj + 6, nk + 6]); call load(s_etx, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(15), f, [ni + 6, nj + 6, nk + 6]); call load(s_ety, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(16), f, [ni + 6, nj + 6, nk + 6]); call load(s_etz, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(17), f, [ni + 6, nj + 6, nk + 6]); call load(s_ctx, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(18), f, [ni + 6, nj + 6, nk + 6]); call load(s_cty, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(19), f, [ni + 6, nj + 6, nk + 6]); call load(s_ctz, f, ni, nj, nk, -1, 1)
  call c_f_pointer(fld(20), f, [ni + 6, nj + 6, nk + 6]); call load(s_p, f, ni, nj, nk, -2, 3)
  if (nl == 2) then
    call c_f_pointer(fld(21), f, [ni + 6, nj + 6, nk + 6]); call load(s_qke(:,:,:,2), f, ni, nj, nk, -2, 3)
  end if

  ! the module state recast_grid / recast_field (Sub2.f90:324-402) would set
  g_ni = ni; g_nj = nj; g_nk = nk
  nmax = max(ni, nj, nk)
  r => s_r; u => s_u; v => s_v; w => s_w
  qke => s_qke; vist => s_vist; visl => s_visl
  vol => s_vol; dtdt => s_dtdt; ds_turbulence => s_dist
  kcx => s_kcx; kcy => s_kcy; kcz => s_kcz
  etx => s_etx; ety => s_ety; etz => s_etz
  ctx => s_ctx; cty => s_cty; ctz => s_ctz
  p => s_p; kct => s_kct; ett => s_ett; ctt => s_ctt; c => s_c
end subroutine ref_turb_load

! SA: dims = (ni, nj, nk); scal = (reynolds, muoo, kmaxlim, timedt_turb, xk, xb, efix);
! fld as ref_turb_load (20 pointers).  qke_out, vist_out: C frame.  Returns 0.
integer(c_int) function ref_sa_step(dims, scal, fld, qke_out, vist_out) bind(C, name="ref_sa_step")
  use iso_c_binding
  use global_variables, only: method, nlamtur, nameturb, ncmpcor, ndes, ntrans, nwtmax, reynolds, re, &
       muoo, kmaxlim, timedt_turb, xk, xb, efix
  use ref_sa_store
  implicit none
  integer(c_int), intent(in) :: dims(3)
  real(c_double), intent(in) :: scal(7)
  type(c_ptr), intent(in) :: fld(20)
  real(c_double), intent(out) :: qke_out(-2:dims(1)+3, -2:dims(2)+3, -2:dims(3)+3)
  real(c_double), intent(out) :: vist_out(-2:dims(1)+3, -2:dims(2)+3, -2:dims(3)+3)
  external :: turbulence

  call ref_turb_load(dims(1), dims(2), dims(3), 1, fld)
  ! run options of the SA path (param.dat namelists; re as init_inflow sets it,
  ! PreProcess.f90:1106)
  method = 1; nlamtur = 1; nameturb = 'SA'
  ncmpcor = 0; ndes = 0; ntrans = 0; nwtmax = 0
  reynolds = scal(1); re = 1.0 / reynolds
  muoo = scal(2); kmaxlim = scal(3); timedt_turb = scal(4)
  xk = scal(5); xb = scal(6); efix = scal(7)

  call turbulence(1)

  qke_out = s_qke(:,:,:,1)
  vist_out = s_vist
  ref_sa_step = 0
end function ref_sa_step

! SST: scal = (reynolds, koo, omgoo, kmaxlim, timedt_turb, xk, xb, efix); fld as
! ref_turb_load with qke(2) = omega (21 pointers).  Returns 0.
integer(c_int) function ref_sst_step(dims, scal, fld, k_out, w_out, vist_out) bind(C, name="ref_sst_step")
  use iso_c_binding
  use global_variables, only: method, nlamtur, nameturb, ncmpcor, ndes, ntrans, nwtmax, reynolds, re, &
       koo, omgoo, kmaxlim, timedt_turb, xk, xb, efix
  use ref_sa_store
  implicit none
  integer(c_int), intent(in) :: dims(3)
  real(c_double), intent(in) :: scal(8)
  type(c_ptr), intent(in) :: fld(21)
  real(c_double), intent(out) :: k_out(-2:dims(1)+3, -2:dims(2)+3, -2:dims(3)+3)
  real(c_double), intent(out) :: w_out(-2:dims(1)+3, -2:dims(2)+3, -2:dims(3)+3)
  real(c_double), intent(out) :: vist_out(-2:dims(1)+3, -2:dims(2)+3, -2:dims(3)+3)
  external :: turbulence

  call ref_turb_load(dims(1), dims(2), dims(3), 2, fld)
  method = 1; nlamtur = 2; nameturb = 'SST'
  ncmpcor = 0; ndes = 0; ntrans = 0; nwtmax = 0
  reynolds = scal(1); re = 1.0 / reynolds
  koo = scal(2); omgoo = scal(3); kmaxlim = scal(4); timedt_turb = scal(5)
  xk = scal(6); xb = scal(7); efix = scal(8)

  call turbulence(1)

  k_out = s_qke(:,:,:,1)
  w_out = s_qke(:,:,:,2)
  vist_out = s_vist
  ref_sst_step = 0
end function ref_sst_step

! Wall values of the turbulence variables: for each of the npts face points
! (is,js,ks, it,jt,kt, i,j,k in idx, 9 per point; ghost, first interior, face)
! calls the reference's BC_wall_turbulence and BC_face_dif_turbulence
! (turblence.f90:2252-2306) on the given state, as viscid_wall (BC.f90:1244-1409)
! does around its face update.  scal = (reynolds, visc); nl = 1 (SA) or 2 (SST);
! fld = r, t, qke, qkw (ignored for nl = 1), dist (C frame).
integer(c_int) function ref_wall_turb(dims, nl, scal, fld, npts, idx, qke_out, qkw_out) bind(C, name="ref_wall_turb")
  use iso_c_binding
  use global_variables, only: method, nlamtur, nameturb, reynolds, visc, r, t, mb_qke, mb_distance
  use ref_sa_store
  implicit none
  integer(c_int), intent(in) :: dims(3)
  integer(c_int), value :: nl, npts
  real(c_double), intent(in) :: scal(2)
  type(c_ptr), intent(in) :: fld(5)
  integer(c_int), intent(in) :: idx(9, npts)
  real(c_double), intent(out) :: qke_out(-2:dims(1)+3, -2:dims(2)+3, -2:dims(3)+3)
  real(c_double), intent(out) :: qkw_out(-2:dims(1)+3, -2:dims(2)+3, -2:dims(3)+3)
  real(c_double), pointer :: f(:,:,:)
  real, target, allocatable, save :: s_t(:,:,:)
  integer :: ni, nj, nk, m, q
  external :: BC_wall_turbulence, BC_face_dif_turbulence

  ni = dims(1); nj = dims(2); nk = dims(3)
  call release()
  if (allocated(s_t)) deallocate(s_t)
  allocate(s_r(-2:ni+3, -2:nj+3, -2:nk+3), s_t(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(s_qke(-2:ni+3, -2:nj+3, -2:nk+3, nl), s_dist(1:ni, 1:nj, 1:nk))
  call c_f_pointer(fld(1), f, [ni + 6, nj + 6, nk + 6]); call load(s_r, f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(2), f, [ni + 6, nj + 6, nk + 6]); call load(s_t, f, ni, nj, nk, -2, 3)
  call c_f_pointer(fld(3), f, [ni + 6, nj + 6, nk + 6]); call load(s_qke(:,:,:,1), f, ni, nj, nk, -2, 3)
  if (nl == 2) then
    call c_f_pointer(fld(4), f, [ni + 6, nj + 6, nk + 6]); call load(s_qke(:,:,:,2), f, ni, nj, nk, -2, 3)
  end if
  call c_f_pointer(fld(5), f, [ni + 6, nj + 6, nk + 6]); call load(s_dist, f, ni, nj, nk, 1, 0)
  if (associated(mb_qke)) deallocate(mb_qke)
  if (associated(mb_distance)) deallocate(mb_distance)
  allocate(mb_qke(1), mb_distance(1))
  mb_qke(1)%a4d => s_qke
  mb_distance(1)%a3d => s_dist
  r => s_r; t => s_t
  method = 1; nlamtur = nl
  if (nl == 2) then
    nameturb = 'SST'
  else
    nameturb = 'SA'
  end if
  reynolds = scal(1); visc = scal(2)
  do q = 1, npts
    call BC_wall_turbulence(1, idx(1,q), idx(2,q), idx(3,q), idx(4,q), idx(5,q), idx(6,q), &
                            idx(7,q), idx(8,q), idx(9,q))
    call BC_face_dif_turbulence(1, 1, idx(1,q), idx(2,q), idx(3,q), idx(4,q), idx(5,q), idx(6,q), &
                                idx(7,q), idx(8,q), idx(9,q))
  end do
  qke_out = s_qke(:,:,:,1)
  qkw_out = 0.0
  if (nl == 2) qkw_out = s_qke(:,:,:,2)
  m = 0
  ref_wall_turb = m
end function ref_wall_turb
