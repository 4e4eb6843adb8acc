! ref_step_bridge.f90 -- TEST INFRASTRUCTURE ONLY, never part of the library.
!
! Loads a whole multi-block case handed over from C into the reference's
! module variables and runs the reference's own serial step routines on it:
!   boundary_all        sub_tgh.f90:123-186 (-> boundary_sequence, BC.f90:169-210)
!   initial_q_dq_c_t    sub_tgh.f90:1-32
!   timestep_tgh        sub_tgh.f90:1270-1348
!   r_h_s_vis           RHS.f90:2-30 (turbulence(nb), WCNSE5_VIS_virtual)
!   r_h_s_invis         RHS.f90:35-72
!   l_h_s_tgh           LHS.f90:1-34 (GS_PR_Single, update_limited)
!   residual_curve      PostProcess.f90:1195-1347
! oracle/Makefile links it with those reference files compiled from source in
! the reference's serial configuration (PARALLEL undefined) into
! _ref/libref_step.so.  PreProcess.f90 and Main.f90 do not build serially
! (they name mod_parallels symbols that exist only under PARALLEL), so the
! set-up they would do is done here instead:
!   * the allocations of allocate_other_variable (PreProcess.f90:682-818),
!     with the reference's bounds;
!   * the per-region analysis of read_bc (PreProcess.f90:413-482), after which
!     the reference's own read_bc_connect (Sub2.f90:2-114) reads the target
!     line from unit 1 and forms the image maps;
!   * the region order of set_bc_index (PreProcess.f90:2543-2572);
!   * the scalars of read_parameter / init_inflow (PreProcess.f90:80-159,
!     968-1113), handed over by the caller.
! Every field crosses in the C oracle's frame: (ni+6)*(nj+6)*(nk+6) doubles,
! i fastest, index 0 = node -2.
module ref_step_state
  implicit none
  integer, save :: nb_loaded = 0
contains
  ! copy nodes lo..n+hi of a C-frame field into dst(lo:ni+hi, ...)
  subroutine from_c(dst, src, ni, nj, nk, lo, hi)
    integer, intent(in) :: ni, nj, nk, lo, hi
    real, intent(in) :: src(-2:ni+3, -2:nj+3, -2:nk+3)
    real, intent(inout) :: dst(lo:ni+hi, lo:nj+hi, lo:nk+hi)
    dst = src(lo:ni+hi, lo:nj+hi, lo:nk+hi)
  end subroutine from_c

  ! copy dst(lo:ni+hi, ...) into a C-frame field (other nodes left as they are)
  subroutine to_c(dst, src, ni, nj, nk, lo, hi)
    integer, intent(in) :: ni, nj, nk, lo, hi
    real, intent(inout) :: dst(-2:ni+3, -2:nj+3, -2:nk+3)
    real, intent(in) :: src(lo:ni+hi, lo:nj+hi, lo:nk+hi)
    dst(lo:ni+hi, lo:nj+hi, lo:nk+hi) = src
  end subroutine to_c
end module ref_step_state

! scal = (moo, reynolds (after init_inflow's rescaling), attack, sideslip
!         (radians), tref, twall, cfl, timedt_rate, efix, csrv, gama, prl, prt,
!         roo, uoo, voo, woo, poo, coo, too, visc, lfref, lref, kmaxlim,
!         timedt_turb, xk, xb)                                   (27 doubles)
! iscal = (nblocks, nvis, nlamtur, nscheme, nflux, ntmst)       (6 ints)
! With a turbulence model the freestream values come from the reference's
! inflow_turbulence (turblence.f90:3166-3181).
integer(c_int) function ref_step_init(scal, iscal) bind(C, name="ref_step_init")
  use iso_c_binding
  use global_variables
  use resdual_module
  use ref_step_state
  implicit none
  real(c_double), intent(in) :: scal(27)
  integer(c_int), intent(in) :: iscal(6)
  integer :: n
  external :: inflow_turbulence

  nblocks = iscal(1)
  nvis = iscal(2); nlamtur = iscal(3); nscheme = iscal(4); nflux = iscal(5); ntmst = iscal(6)
  moo = scal(1); reynolds = scal(2); attack = scal(3); sideslip = scal(4)
  tref = scal(5); twall = scal(6); cfl = scal(7); timedt_rate = scal(8)
  efix = scal(9); csrv = scal(10); gama = scal(11); prl = scal(12); prt = scal(13)
  roo = scal(14); uoo = scal(15); voo = scal(16); woo = scal(17); poo = scal(18)
  coo = scal(19); too = scal(20); visc = scal(21); lfref = scal(22); lref = scal(23)
  kmaxlim = scal(24); timedt_turb = scal(25); xk = scal(26); xb = scal(27)
  re = 1.0 / reynolds
  eoo = poo / (gama - 1.0) + 0.5 * roo * (uoo * uoo + voo * voo + woo * woo)
  ! options the hot path of the reference is run with (param.dat defaults,
  ! Main.f90:20-21)
  method = 1; nijk2d = 4; nijk2nd = 4; ndim = 3
  gcl = 1; cic1 = 0; cic2 = 0; cic3 = 0; cic4 = 0; cic5 = 0
  ndualtst = 0; nsubstep = 1; nchem = 0; nchem_source = 0
  ncmpcor = 0; ndes = 0; ntrans = 0; nwtmax = 0; nwallfun = 0; connect_order = 0
  if (nlamtur == 2) then
    nameturb = 'SST'
  else
    nameturb = 'SA'
  end if
  nstart = 0; nreset = 0
  if (nlamtur > 0) call inflow_turbulence
  nmax = 0

  n = nblocks
  allocate(mb_dim(n, 3), mb_bc(n))
  allocate(mb_x(n), mb_y(n), mb_z(n), mb_vol(n))
  allocate(mb_kcx(n), mb_kcy(n), mb_kcz(n), mb_kct(n))
  allocate(mb_etx(n), mb_ety(n), mb_etz(n), mb_ett(n))
  allocate(mb_ctx(n), mb_cty(n), mb_ctz(n), mb_ctt(n))
  allocate(mb_r(n), mb_u(n), mb_v(n), mb_w(n), mb_p(n), mb_t(n), mb_c(n))
  allocate(mb_sra(n), mb_srb(n), mb_src(n), mb_srva(n), mb_srvb(n), mb_srvc(n))
  allocate(mb_dtdt(n), mb_visl(n), mb_vist(n), mb_q(n), mb_dq(n), mb_qke(n), mb_distance(n))
  allocate(res(n), resmax(n), nres(n, 4))
  nb_loaded = n
  ref_step_init = 0
end function ref_step_init

! Allocates block nb (1-based) with the bounds of allocate_other_variable and
! loads x, y, z (PLOT3D order, 1..n).
integer(c_int) function ref_step_block(nb, dims, xc, yc, zc) bind(C, name="ref_step_block")
  use iso_c_binding
  use global_variables, only: nl, nlamtur, nvis, nmax, mb_dim, mb_bc, mb_x, mb_y, mb_z, mb_vol, &
       mb_kcx, mb_kcy, mb_kcz, mb_kct, mb_etx, mb_ety, mb_etz, mb_ett, mb_ctx, mb_cty, mb_ctz, mb_ctt, &
       mb_r, mb_u, mb_v, mb_w, mb_p, mb_t, mb_c, mb_sra, mb_srb, mb_src, mb_srva, mb_srvb, mb_srvc, &
       mb_visl, mb_dtdt, mb_q, mb_dq, mb_vist, mb_qke, mb_distance
  implicit none
  integer(c_int), value, intent(in) :: nb
  integer(c_int), intent(in) :: dims(3)
  real(c_double), intent(in) :: xc(dims(1), dims(2), dims(3)), yc(dims(1), dims(2), dims(3))
  real(c_double), intent(in) :: zc(dims(1), dims(2), dims(3))
  integer :: ni, nj, nk

  ni = dims(1); nj = dims(2); nk = dims(3)
  mb_dim(nb, 1) = ni; mb_dim(nb, 2) = nj; mb_dim(nb, 3) = nk
  nmax = max(nmax, ni, nj, nk)
  allocate(mb_x(nb)%a3d(ni, nj, nk), mb_y(nb)%a3d(ni, nj, nk), mb_z(nb)%a3d(ni, nj, nk))
  mb_x(nb)%a3d = xc; mb_y(nb)%a3d = yc; mb_z(nb)%a3d = zc
  allocate(mb_vol(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_kcx(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_kcy(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_kcz(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_kct(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_etx(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_ety(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_etz(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_ett(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_ctx(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_cty(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_ctz(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_ctt(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  mb_kct(nb)%a3d = 0.0; mb_ett(nb)%a3d = 0.0; mb_ctt(nb)%a3d = 0.0
  mb_vol(nb)%a3d = 0.0
  mb_kcx(nb)%a3d = 0.0; mb_kcy(nb)%a3d = 0.0; mb_kcz(nb)%a3d = 0.0
  mb_etx(nb)%a3d = 0.0; mb_ety(nb)%a3d = 0.0; mb_etz(nb)%a3d = 0.0
  mb_ctx(nb)%a3d = 0.0; mb_cty(nb)%a3d = 0.0; mb_ctz(nb)%a3d = 0.0
  allocate(mb_r(nb)%a3d(-2:ni+3, -2:nj+3, -2:nk+3), mb_u(nb)%a3d(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(mb_v(nb)%a3d(-2:ni+3, -2:nj+3, -2:nk+3), mb_w(nb)%a3d(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(mb_p(nb)%a3d(-2:ni+3, -2:nj+3, -2:nk+3), mb_t(nb)%a3d(-2:ni+3, -2:nj+3, -2:nk+3))
  allocate(mb_c(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_sra(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_srb(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_src(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_srva(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_srvb(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1), mb_srvc(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_visl(nb)%a3d(-1:ni+1, -1:nj+1, -1:nk+1))
  allocate(mb_dtdt(nb)%a3d(1:ni, 1:nj, 1:nk))
  allocate(mb_q(nb)%a4d(1:nl, -2:ni+3, -2:nj+3, -2:nk+3), mb_dq(nb)%a4d(1:nl, -2:ni+3, -2:nj+3, -2:nk+3))
  allocate(mb_vist(nb)%a3d(-2:ni+3, -2:nj+3, -2:nk+3))
  mb_t(nb)%a3d = 0.0; mb_c(nb)%a3d = 0.0; mb_visl(nb)%a3d = 0.0; mb_vist(nb)%a3d = 0.0
  mb_sra(nb)%a3d = 0.0; mb_srb(nb)%a3d = 0.0; mb_src(nb)%a3d = 0.0
  mb_srva(nb)%a3d = 0.0; mb_srvb(nb)%a3d = 0.0; mb_srvc(nb)%a3d = 0.0
  mb_q(nb)%a4d = 0.0; mb_dq(nb)%a4d = 0.0; mb_dtdt(nb)%a3d = 0.0
  if (nlamtur >= 0) then
    allocate(mb_qke(nb)%a4d(-2:ni+3, -2:nj+3, -2:nk+3, nlamtur))
    mb_qke(nb)%a4d = 0.0
  end if
  if (nvis > 0 .and. nlamtur >= 0) then
    allocate(mb_distance(nb)%a3d(1:ni, 1:nj, 1:nk))
    mb_distance(nb)%a3d = 0.0
  end if
  mb_bc(nb)%nregions = 0
  ref_step_block = 0
end function ref_step_block

! Region nr (1-based) of block nb out of nregions, one line of the .inp file:
! s_raw = (s_st1, s_ed1, s_st2, s_ed2, s_st3, s_ed3) with the file's signs;
! for bctype < 0 the target line t_raw and the target block nbt (1-based).
! The analysis is read_bc's (PreProcess.f90:448-480); the image maps come from
! the reference's read_bc_connect (Sub2.f90:2-114), which reads the target
! line from unit 1.
integer(c_int) function ref_step_region(nb, nr, nregions, s_raw, bctype, t_raw, nbt, scratch_file) &
    bind(C, name="ref_step_region")
  use iso_c_binding
  use global_variables
  implicit none
  integer(c_int), value, intent(in) :: nb, nr, nregions, bctype, nbt
  integer(c_int), intent(in) :: s_raw(6), t_raw(6)
  character(kind=c_char), intent(in) :: scratch_file(*)
  integer :: m, s_st(3), s_ed(3), s_nd, s_lr, s_fix, l
  character(len=512) :: fname
  external :: read_bc_connect

  if (nr == 1) then
    mb_bc(nb)%nregions = nregions
    allocate(mb_bc(nb)%bc(nregions))
  end if
  s_st(1) = s_raw(1); s_ed(1) = s_raw(2); s_st(2) = s_raw(3); s_ed(2) = s_raw(4)
  s_st(3) = s_raw(5); s_ed(3) = s_raw(6)
  mb_bc(nb)%bc(nr)%bctype = bctype
  mb_bc(nb)%bc(nr)%nbs = nb
  s_nd = 0; s_lr = 0; s_fix = 0
  do m = 1, 3
    mb_bc(nb)%bc(nr)%s_st(m) = min(abs(s_st(m)), abs(s_ed(m)))
    mb_bc(nb)%bc(nr)%s_ed(m) = max(abs(s_st(m)), abs(s_ed(m)))
    mb_bc(nb)%bc(nr)%s_lr3d(m) = 0
    if (s_st(m) == s_ed(m)) then
      s_nd = m
      if (s_st(m) == 1) then
        s_lr = -1
        mb_bc(nb)%bc(nr)%s_lr3d(m) = -1
      else
        s_lr = 1
        mb_bc(nb)%bc(nr)%s_lr3d(m) = 1
      end if
      s_fix = s_st(m)
    else
      mb_bc(nb)%bc(nr)%s_ed(m) = mb_bc(nb)%bc(nr)%s_ed(m) + method - 1
    end if
  end do
  mb_bc(nb)%bc(nr)%s_lr = s_lr
  mb_bc(nb)%bc(nr)%s_nd = s_nd
  mb_bc(nb)%bc(nr)%s_fix = s_fix
  if (bctype < 0) then
    fname = ' '
    l = 0
    do while (scratch_file(l + 1) /= c_null_char .and. l < 511)
      l = l + 1
      fname(l:l) = scratch_file(l)
    end do
    open(1, file=fname(1:l), status='replace')
    write(1, *) t_raw(1), t_raw(2), t_raw(3), t_raw(4), t_raw(5), t_raw(6), nbt
    rewind(1)
    call read_bc_connect(1, nb, nr, s_nd, s_st, s_ed)
    close(1, status='delete')
  end if
  ref_step_region = 0
end function ref_step_region

! The conservative metrics of every block by the reference's own
! GRID_DERIVATIVE_gcl (WCNSE5.f90:1478-1728), as set_grid_derivative /
! grid_derivative_nb (PreProcess.f90:1618-1660) call it.
integer(c_int) function ref_step_metrics() bind(C, name="ref_step_metrics")
  use iso_c_binding
  use global_variables, only: nblocks, ni, nj, nk, kct, ett, ctt
  implicit none
  integer :: nb
  external :: recast_grid, GRID_DERIVATIVE_gcl
  do nb = 1, nblocks
    call recast_grid(nb)
    call GRID_DERIVATIVE_gcl(nb)
    kct(1:ni, 1:nj, 1:nk) = 0.0
    ett(1:ni, 1:nj, 1:nk) = 0.0
    ctt(1:ni, 1:nj, 1:nk) = 0.0
  end do
  ref_step_metrics = 0
end function ref_step_metrics

! The uniform initial turbulence field of the reference, init_SA / init_SST
! (turblence.f90:1772-1782, 2701-2733, 3090-3117).
integer(c_int) function ref_step_turb_init() bind(C, name="ref_step_turb_init")
  use iso_c_binding
  implicit none
  external :: init_uniform_turbulence
  call init_uniform_turbulence
  ref_step_turb_init = 0
end function ref_step_turb_init

! initial_t (sub_tgh.f90:35-58): the initial temperature that pre_processing
! forms before the first step when nvis = 1 (PreProcess.f90:62).
integer(c_int) function ref_step_initial_t() bind(C, name="ref_step_initial_t")
  use iso_c_binding
  implicit none
  external :: initial_t
  call initial_t
  ref_step_initial_t = 0
end function ref_step_initial_t

! force + force_dif (PostProcess.f90:1415-1680) with the namelist force values
! fref = (sref, xref, yref, zref) and nwhole; force appends its line to the
! file named by fname (forcename).  out = cfx, cfy, cfz, cmx, cmy, cmz, cd, cl, xcp.
integer(c_int) function ref_step_force(fref, nwhole, fname, out) bind(C, name="ref_step_force")
  use iso_c_binding
  use global_variables, only: sref, xref, yref, zref, nwholefield, forcename, cfx, cfy, cfz, cmx, cmy, &
       cmz, cd, cl, xcp, nout
  implicit none
  real(c_double), intent(in) :: fref(4)
  integer(c_int), value, intent(in) :: nwhole
  character(kind=c_char), intent(in) :: fname(*)
  real(c_double), intent(out) :: out(9)
  integer :: l
  external :: force
  sref = fref(1); xref = fref(2); yref = fref(3); zref = fref(4)
  nwholefield = nwhole
  forcename = ' '
  l = 0
  do while (fname(l + 1) /= c_null_char .and. l < 119)
    l = l + 1
    forcename(l:l) = fname(l)
  end do
  nout = 0
  call force
  out = (/ cfx, cfy, cfz, cmx, cmy, cmz, cd, cl, xcp /)
  ref_step_force = 0
end function ref_step_force

! The reference's Tecplot writers, on the loaded state: kind 0
! tecout_bin_surface (PostProcess.f90:97-336, the file tecname), 1
! tecout_bin_wall (:338-582, tecname's stem + "_wall.plt"), 2
! tecout_bin_volume (:584-772, tecname's stem + "_vol.plt").
integer(c_int) function ref_step_tecout(kind, fname) bind(C, name="ref_step_tecout")
  use iso_c_binding
  use global_variables, only: tecname
  implicit none
  integer(c_int), value, intent(in) :: kind
  character(kind=c_char), intent(in) :: fname(*)
  integer :: l
  external :: tecout_bin_surface, tecout_bin_wall, tecout_bin_volume
  tecname = ' '
  l = 0
  do while (fname(l + 1) /= c_null_char .and. l < len(tecname))
    l = l + 1
    tecname(l:l) = fname(l)
  end do
  ref_step_tecout = 0
  select case (kind)
  case (0)
    call tecout_bin_surface
  case (1)
    call tecout_bin_wall
  case (2)
    call tecout_bin_volume
  case default
    ref_step_tecout = -1
  end select
end function ref_step_tecout

! The turbulence freestream values: muoo, koo, omgoo, vtoo.
integer(c_int) function ref_step_turb_inflow(out) bind(C, name="ref_step_turb_inflow")
  use iso_c_binding
  use global_variables, only: muoo, koo, omgoo, vtoo
  implicit none
  real(c_double), intent(out) :: out(4)
  out(1) = muoo; out(2) = koo; out(3) = omgoo; out(4) = vtoo
  ref_step_turb_inflow = 0
end function ref_step_turb_inflow

! The region order of set_bc_index (PreProcess.f90:2543-2572): -1, 4, 5, 6, 2, 3.
integer(c_int) function ref_step_bcindex() bind(C, name="ref_step_bcindex")
  use iso_c_binding
  use global_variables
  implicit none
  integer, parameter :: bclists(6) = (/-1, 4, 5, 6, 2, 3/)
  integer :: nb, nr, i, ibl, nrmax, n1, n2, n3, m, j, k
  ref_step_bcindex = 0
  do nb = 1, nblocks
    nrmax = mb_bc(nb)%nregions
    allocate(mb_bc(nb)%bcindexs(nrmax))
    i = 0
    do ibl = 1, 6
      do nr = 1, nrmax
        if (mb_bc(nb)%bc(nr)%bctype == bclists(ibl)) then
          i = i + 1
          mb_bc(nb)%bcindexs(i) = nr
        end if
      end do
    end do
    if (i /= nrmax) ref_step_bcindex = -1
  end do
  ! the face-node region flags of allocate_other_variable and fill_bc_flag
  ! (PreProcess.f90:695, 761-766, 820-857): per face, the last region in
  ! the order above whose box holds the node
  allocate(mb_flg(nblocks, 6))
  do nb = 1, nblocks
    n1 = mb_dim(nb, 1); n2 = mb_dim(nb, 2); n3 = mb_dim(nb, 3)
    allocate(mb_flg(nb, 1)%a3d(1:1, 1:n2, 1:n3), mb_flg(nb, 2)%a3d(n1:n1, 1:n2, 1:n3))
    allocate(mb_flg(nb, 3)%a3d(1:n1, 1:1, 1:n3), mb_flg(nb, 4)%a3d(1:n1, n2:n2, 1:n3))
    allocate(mb_flg(nb, 5)%a3d(1:n1, 1:n2, 1:1), mb_flg(nb, 6)%a3d(1:n1, 1:n2, n3:n3))
    do ibl = 1, mb_bc(nb)%nregions
      nr = mb_bc(nb)%bcindexs(ibl)
      m = 2 * mb_bc(nb)%bc(nr)%s_nd + (mb_bc(nb)%bc(nr)%s_lr - 1) / 2
      do k = mb_bc(nb)%bc(nr)%s_st(3), mb_bc(nb)%bc(nr)%s_ed(3)
        do j = mb_bc(nb)%bc(nr)%s_st(2), mb_bc(nb)%bc(nr)%s_ed(2)
          do i = mb_bc(nb)%bc(nr)%s_st(1), mb_bc(nb)%bc(nr)%s_ed(1)
            mb_flg(nb, m)%a3d(i, j, k) = nr
          end do
        end do
      end do
    end do
  end do
end function ref_step_bcindex

! image/jmage/kmage of region nr of block nb over its face box, i fastest
integer(c_int) function ref_step_images(nb, nr, img, jmg, kmg) bind(C, name="ref_step_images")
  use iso_c_binding
  use global_variables
  implicit none
  integer(c_int), value, intent(in) :: nb, nr
  integer(c_int), intent(out) :: img(*), jmg(*), kmg(*)
  integer :: i, j, k, o
  ref_step_images = -1
  if (mb_bc(nb)%bc(nr)%bctype >= 0) return
  o = 0
  do k = mb_bc(nb)%bc(nr)%s_st(3), mb_bc(nb)%bc(nr)%s_ed(3)
    do j = mb_bc(nb)%bc(nr)%s_st(2), mb_bc(nb)%bc(nr)%s_ed(2)
      do i = mb_bc(nb)%bc(nr)%s_st(1), mb_bc(nb)%bc(nr)%s_ed(1)
        o = o + 1
        img(o) = mb_bc(nb)%bc(nr)%image(i, j, k)
        jmg(o) = mb_bc(nb)%bc(nr)%jmage(i, j, k)
        kmg(o) = mb_bc(nb)%bc(nr)%kmage(i, j, k)
      end do
    end do
  end do
  ref_step_images = 0
end function ref_step_images

! which (the oracle's numbering): 0 r, 1 u, 2 v, 3 w, 4 p, 5 t, 6 c, 7 dtdt,
! 8 sra, 9 srb, 10 src, 11 srva, 12 srvb, 13 srvc, 14 visl, 15 vist, 16 vol,
! 17..25 kcx kcy kcz etx ety etz ctx cty ctz, 26 qke(1), 27 distance,
! 28 qke(2); 100 dq, 101 q (5 doubles per node, variable fastest).
! set = 1 copies C -> module, 0 module -> C (nodes outside the reference's
! bounds keep the C array's values).
integer(c_int) function ref_step_field(nb, which, set, a) bind(C, name="ref_step_field")
  use iso_c_binding
  use global_variables, only: mb_dim, mb_vol, mb_kcx, mb_kcy, mb_kcz, mb_etx, mb_ety, mb_etz, mb_ctx, mb_cty, &
       mb_ctz, mb_r, mb_u, mb_v, mb_w, mb_p, mb_t, mb_c, mb_sra, mb_srb, mb_src, mb_srva, mb_srvb, mb_srvc, &
       mb_visl, mb_dtdt, mb_q, mb_dq, mb_vist, mb_qke, mb_distance
  use ref_step_state
  implicit none
  integer(c_int), value, intent(in) :: nb, which, set
  real(c_double), intent(inout) :: a(*)
  integer :: ni, nj, nk
  real, pointer :: f3(:,:,:)
  integer :: lo, hi
  ni = mb_dim(nb, 1); nj = mb_dim(nb, 2); nk = mb_dim(nb, 3)
  ref_step_field = 0
  if (which == 100 .or. which == 101) then
    call q5(a)
    return
  end if
  lo = -1; hi = 1
  select case (which)
  case (0); f3 => mb_r(nb)%a3d; lo = -2; hi = 3
  case (1); f3 => mb_u(nb)%a3d; lo = -2; hi = 3
  case (2); f3 => mb_v(nb)%a3d; lo = -2; hi = 3
  case (3); f3 => mb_w(nb)%a3d; lo = -2; hi = 3
  case (4); f3 => mb_p(nb)%a3d; lo = -2; hi = 3
  case (5); f3 => mb_t(nb)%a3d; lo = -2; hi = 3
  case (6); f3 => mb_c(nb)%a3d
  case (7); f3 => mb_dtdt(nb)%a3d; lo = 1; hi = 0
  case (8); f3 => mb_sra(nb)%a3d
  case (9); f3 => mb_srb(nb)%a3d
  case (10); f3 => mb_src(nb)%a3d
  case (11); f3 => mb_srva(nb)%a3d
  case (12); f3 => mb_srvb(nb)%a3d
  case (13); f3 => mb_srvc(nb)%a3d
  case (14); f3 => mb_visl(nb)%a3d
  case (15); f3 => mb_vist(nb)%a3d; lo = -2; hi = 3
  case (16); f3 => mb_vol(nb)%a3d
  case (17); f3 => mb_kcx(nb)%a3d
  case (18); f3 => mb_kcy(nb)%a3d
  case (19); f3 => mb_kcz(nb)%a3d
  case (20); f3 => mb_etx(nb)%a3d
  case (21); f3 => mb_ety(nb)%a3d
  case (22); f3 => mb_etz(nb)%a3d
  case (23); f3 => mb_ctx(nb)%a3d
  case (24); f3 => mb_cty(nb)%a3d
  case (25); f3 => mb_ctz(nb)%a3d
  case (26); f3 => mb_qke(nb)%a4d(:,:,:,1); lo = -2; hi = 3
  case (27); f3 => mb_distance(nb)%a3d; lo = 1; hi = 0
  case (28); f3 => mb_qke(nb)%a4d(:,:,:,2); lo = -2; hi = 3
  case default
    ref_step_field = -1
    return
  end select
  if (set == 1) then
    call from_c(f3, a, ni, nj, nk, lo, hi)
  else
    call to_c(a, f3, ni, nj, nk, lo, hi)
  end if
contains
  subroutine q5(c)
    real(c_double), intent(inout) :: c(5, -2:ni+3, -2:nj+3, -2:nk+3)
    if (which == 100) then
      if (set == 1) then
        mb_dq(nb)%a4d = c
      else
        c = mb_dq(nb)%a4d
      end if
    else
      if (set == 1) then
        mb_q(nb)%a4d = c
      else
        c = mb_q(nb)%a4d
      end if
    end if
  end subroutine q5
end function ref_step_field

! One of the reference's step routines on every block: 0 boundary_all,
! 1 initial_q_dq_c_t, 2 timestep_tgh, 3 r_h_s_vis, 4 r_h_s_invis, 5 l_h_s_tgh.
integer(c_int) function ref_step_stage(which) bind(C, name="ref_step_stage")
  use iso_c_binding
  implicit none
  integer(c_int), value, intent(in) :: which
  external :: boundary_all, initial_q_dq_c_t, timestep_tgh, r_h_s_vis, r_h_s_invis, l_h_s_tgh
  ref_step_stage = 0
  select case (which)
  case (0); call boundary_all
  case (1); call initial_q_dq_c_t
  case (2); call timestep_tgh
  case (3); call r_h_s_vis
  case (4); call r_h_s_invis
  case (5); call l_h_s_tgh
  case default; ref_step_stage = -1
  end select
end function ref_step_stage

! residual_curve (PostProcess.f90:1195-1347): rms (dres_add), max
! (resmax_total) and its location n_error(1:5) = i, j, k, m, nb.
integer(c_int) function ref_step_residual(out, loc) bind(C, name="ref_step_residual")
  use iso_c_binding
  use resdual_module
  implicit none
  real(c_double), intent(out) :: out(2)
  integer(c_int), intent(out) :: loc(5)
  external :: residual_curve
  call residual_curve(1, 1)
  out(1) = dres_add
  out(2) = resmax_total
  loc = n_error
  ref_step_residual = 0
end function ref_step_residual

! Allocates, fills with v and frees an array of n reals, twice, so that the
! heap's next allocation of that size (inviscd3d's rpres, Convect.f90:79)
! starts from those values: the entries of rpres its loops never write
! (Convect.f90:81-132) but its rpmin reads (Convect.f90:169-177, 335-343,
! 499-507) then hold v.  Test infrastructure for pinning those reads.
integer(c_int) function ref_step_poison(n, v) bind(C, name="ref_step_poison")
  use iso_c_binding
  implicit none
  integer(c_long), value, intent(in) :: n
  real(c_double), value, intent(in) :: v
  real, pointer :: a(:)
  integer :: rep
  do rep = 1, 2
    allocate(a(n))
    a = v
    deallocate(a)
  end do
  ref_step_poison = 0
end function ref_step_poison
