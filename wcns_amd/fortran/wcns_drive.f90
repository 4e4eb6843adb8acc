! wcns_drive.f90 -- a Fortran host that runs the reference's pseudo-time loop
! (src/Main.f90:85-121) with the hot path on the GPU library through the shims
! of wcns_shims.f90 and the bind(C) module wcns_amd_c.
!
! It reads the reference's own inputs: param.dat (namelists of read_parameter,
! PreProcess.f90:80-159), the PLOT3D grid (read_plot3d_grid,
! PreProcess.f90:365-412), the "1976" boundary file (read_bc_parallel,
! parallel.f90:852-929) and, with nstart >= 1, the restart file flowname
! (read_flow_file, PreProcess.f90:1542-1616; with a turbulence model also
! flowname.turb).  It writes what the reference writes: a res.dat line every
! nwerror steps in saveres2file's format (PostProcess.f90:1349-1378), a
! force.dat line every nforce steps in force's format (PostProcess.f90:1479-1481)
! and flowname every ndisk steps and at the end (post_processing,
! PostProcess.f90:11-34).  It also writes errhis//'.full': step, rms and max of
! the monitor in full precision, for exact comparisons.
!
! usage: wcns_drive [param.dat] [stages|graph]
!   stages  every routine of the loop is its own call (the default)
!   graph   the steps between two res.dat lines run as one wcns_step_graph call
program wcns_drive
  use iso_c_binding
  use wcns_amd_c
  use wcns_monitor
  implicit none
  ! namelist values (read_parameter); gama, prl and prt are the perfect-gas
  ! model's (read_perfect_gasmodel, GasModel.f90:184-217)
  real(8) :: moo = 0.8d0, reynolds = 1.0d5, attack = 0.0d0, sideslip = 0.0d0, tref = 288.15d0
  real(8) :: twall = -1.0d0, pref = 0.0d0, rref = 0.0d0, vref = 0.0d0, height = -1.0d0
  integer :: ndim = 3
  integer :: nstart = 0, nmethod = 1, nbgmax = 1, ndisk = 1000, newton = 0, nomax = 10, nforce = 100
  integer :: nwerror = 1, method = 1, nplot = 0
  integer :: ntmst = 0, ndualtst = 0, nsubstmx = 1
  real(8) :: cfl = 5.0d0, timedt = 0.0d0, timedt_rate = 1.1d0, timedt_turb = 10.0d0, dtdts = 0.0d0
  real(8) :: tolsub = 0.0d0
  integer :: nvis = 0, nchem = 0, ntmodel = 1
  integer :: nlamtur = -1, nreset = 0, nwallfun = 0, ntrans = 0, nwtmax = 0, ncmpcor = 0, ndes = 0
  character(len=120) :: nameturb = 'SA'
  real(8) :: xtrans = 0.0d0, kmaxlim = 30000.0d0
  integer :: nlhs = 21, nscheme = 41, nlimiter = 1, nsmooth = 0, nflux = 1
  real(8) :: efix = 0.1d0, csrv = 1.0d0
  real(8) :: xk = -1.0d0, xb = 1.0d0, c_k2 = 0.8d0, c_k4 = 0.05d0
  character(len=120) :: flowname = 'flow.dat', gridname = 'grid.x', bcname = 'grid.inp'
  character(len=120) :: forcename = 'force.dat', errhis = 'res.dat', tecname = 'tec.plt'
  integer :: nwholefield = 1
  real(8) :: sref = 1.0d0, lfref = 1.0d0, lref = 1.0d0, xref = 0.0d0, yref = 0.0d0, zref = 0.0d0
  namelist /inflow/ moo, reynolds, attack, sideslip, tref, twall, pref, rref, vref, ndim, height
  namelist /control/ nstart, nmethod, nbgmax, ndisk, newton, nomax, nforce, nwerror, method, nplot
  namelist /step/ ntmst, cfl, timedt, timedt_rate, timedt_turb, ndualtst, dtdts, nsubstmx, tolsub
  namelist /flowtype/ nvis, nchem, ntmodel
  namelist /turbulence/ nlamtur, nameturb, nreset, nwallfun, ntrans, xtrans, nwtmax, ncmpcor, kmaxlim, ndes
  namelist /technic/ nlhs, nscheme, nlimiter, efix, csrv, nsmooth, nflux
  namelist /interplate/ xk, xb, c_k2, c_k4
  namelist /filename/ flowname, gridname, bcname, forcename, errhis, tecname
  namelist /force/ nwholefield, sref, lfref, lref, xref, yref, zref

  type(wcns_params) :: prm
  character(len=256) :: pfile, mode, line
  integer :: nblocks, nb, nr, nrmax, i, ngs, nout, m, ios, nsteps_graph
  integer, allocatable :: dims(:, :)
  integer :: s_raw(6), t_raw(6), bctype, nbt, ibcwin, marker, nprocs, nbf, pid, idum(3)
  character(len=120) :: blockname
  real(8), allocatable :: xyz(:)
  type xyz_block
    real(8), allocatable :: x(:, :, :), y(:, :, :), z(:, :, :)
  end type
  type(xyz_block), allocatable :: g(:)
  real(8), allocatable :: r(:), u(:), v(:), w(:), p(:)
  integer(c_int) :: rc, nbc, nout_c
  real(8) :: wholetime, pai, att, ssl, cpu0, cpu1, fc(9)
  real(8) :: hrms(8192), hmax(8192)
  integer(c_int) :: hloc(5, 8192)
  integer(c_long_long) :: hcount
  integer :: n3

  pfile = 'param.dat'
  mode = 'stages'
  if (command_argument_count() >= 1) call get_command_argument(1, pfile)
  if (command_argument_count() >= 2) call get_command_argument(2, mode)

  ! ---- param.dat: the reference writes its groups as $name ... /
  call read_namelists(trim(pfile))

  ! ---- grid (read_plot3d_grid)
  open(101, file=trim(gridname), form='unformatted', status='old', action='read')
  read(101) nblocks
  allocate(dims(3, nblocks), g(nblocks))
  do nb = 1, nblocks
    read(101) dims(1, nb), dims(2, nb), dims(3, nb)
  end do
  do nb = 1, nblocks
    allocate(g(nb)%x(dims(1, nb), dims(2, nb), dims(3, nb)), g(nb)%y(dims(1, nb), dims(2, nb), dims(3, nb)))
    allocate(g(nb)%z(dims(1, nb), dims(2, nb), dims(3, nb)))
    read(101) g(nb)%x, g(nb)%y, g(nb)%z
  end do
  close(101)

  ! ---- the library context (wcns_amd_c)
  prm%moo = moo; prm%reynolds = reynolds; prm%attack = attack; prm%sideslip = sideslip
  prm%tref = tref; prm%twall = twall; prm%cfl = cfl; prm%timedt_rate = timedt_rate
  prm%efix = efix; prm%csrv = csrv; prm%gama = 1.4d0; prm%prl = 0.72d0; prm%prt = 0.90d0
  prm%nvis = nvis; prm%nlamtur = nlamtur; prm%nscheme = nscheme; prm%nflux = nflux; prm%ntmst = ntmst
  prm%sref = sref; prm%lfref = lfref; prm%xref = xref; prm%yref = yref; prm%zref = zref
  prm%nwholefield = nwholefield
  prm%timedt_turb = timedt_turb; prm%kmaxlim = kmaxlim; prm%xk = xk; prm%xb = xb; prm%lref = lref
  rc = wcns_create(wcns_ctx_h, 0_c_int, prm)
  if (rc /= 0) then
    write(*, '(a,i4)') ' wcns_amd: wcns_create failed: ', rc
    stop 2
  end if
  do nb = 1, nblocks
    rc = wcns_add_block(wcns_ctx_h, dims(1, nb), dims(2, nb), dims(3, nb), g(nb)%x, g(nb)%y, g(nb)%z, nbc)
    if (rc /= 0) call wcns_abort('wcns_add_block', rc)
  end do

  ! ---- boundary file (read_bc_parallel, "1976" format)
  open(102, file=trim(bcname), status='old', action='read')
  read(102, *) marker
  if (marker /= 1976) then
    write(*, *) 'wcns_drive: ', trim(bcname), ' is not a 1976-format boundary file'
    stop 2
  end if
  read(102, *) nprocs
  read(102, *) nbf
  if (nbf /= nblocks) then
    write(*, *) 'wcns_drive: block counts of grid and boundary file differ'
    stop 2
  end if
  do nb = 1, nblocks
    read(102, *) pid
    read(102, *) idum(1), idum(2), idum(3)
    if (any(idum /= dims(:, nb))) then
      write(*, *) 'wcns_drive: block', nb, ' dimensions differ between grid and boundary file'
      stop 2
    end if
    read(102, *) blockname
    read(102, *) nrmax
    do nr = 1, nrmax
      read(102, *) s_raw, bctype
      t_raw = 0; nbt = 0; ibcwin = 0
      if (bctype < 0) read(102, *) t_raw, nbt, ibcwin
      rc = wcns_add_region(wcns_ctx_h, nb - 1, s_raw, bctype, t_raw, nbt, ibcwin)
      if (rc /= 0) call wcns_abort('wcns_add_region', rc)
    end do
  end do
  close(102)
  rc = wcns_setup(wcns_ctx_h)
  if (rc /= 0) call wcns_abort('wcns_setup', rc)

  ! ---- initial field: the freestream of init_inflow everywhere (init_nstart0,
  ! PreProcess.f90:1250-1315), then with nstart >= 1 the restart file
  pai = 3.14159265d0
  att = attack * pai / 180.0d0
  ssl = sideslip * pai / 180.0d0
  do nb = 1, nblocks
    n3 = (dims(1, nb) + 6) * (dims(2, nb) + 6) * (dims(3, nb) + 6)
    allocate(r(n3), u(n3), v(n3), w(n3), p(n3))
    r = 1.0d0
    u = cos(att) * cos(ssl)
    v = sin(att) * cos(ssl)
    w = sin(ssl)
    p = 1.0d0 / (1.4d0 * moo * moo)
    rc = wcns_set_prim(wcns_ctx_h, nb - 1, r, u, v, w, p)
    if (rc /= 0) call wcns_abort('wcns_set_prim', rc)
    deallocate(r, u, v, w, p)
  end do
  nout = 0
  wholetime = 0.0d0
  if (nstart >= 1) then
    rc = wcns_load_flow(wcns_ctx_h, trim(flowname) // c_null_char, 0_c_int, nout_c, wholetime)
    if (rc /= 0) call wcns_abort('wcns_load_flow', rc)
    if (nstart /= 1) nout = nout_c
  end if

  open(201, file=trim(errhis), access='append', status='unknown')
  open(202, file=trim(errhis) // '.full', status='replace')
  open(203, file=trim(forcename), access='append', status='unknown')
  call cpu_time(cpu0)

  ! ---- the pseudo-time loop (Main.f90:85-121)
  if (trim(mode) == 'graph') then
    ngs = 0
    do while (ngs < nomax)
      nsteps_graph = min(nwerror - mod(nout, nwerror), nomax - ngs)
      rc = wcns_step_graph(wcns_ctx_h, nsteps_graph)
      if (rc /= 0) call wcns_abort('wcns_step_graph', rc)
      rc = wcns_residual_history(wcns_ctx_h, nsteps_graph, hrms, hmax, hloc, hcount)
      if (rc /= 0) call wcns_abort('wcns_residual_history', rc)
      do m = 1, nsteps_graph
        nout = nout + 1
        ngs = ngs + 1
        dres_add = hrms(m); resmax_total = hmax(m); n_error = hloc(:, m)
        call cpu_time(cpu1)
        call write_monitor(nout, cpu1 - cpu0)
      end do
      call post(nout, ngs)
    end do
  else
    do ngs = 1, nomax
      nout = nout + 1
      call boundary_all
      call initial_q_dq_c_t
      call timestep_tgh
      if (nvis == 1) then
        call r_h_s_vis
        call communicate_dq_npp
      end if
      call r_h_s_invis
      call communicate_dq_npp
      call l_h_s_tgh
      call communicate_pv_npp
      call residual_curve(1, 1)
      call cpu_time(cpu1)
      call write_monitor(nout, cpu1 - cpu0)
      call post(nout, ngs)
    end do
  end if
  close(201); close(202); close(203)
  rc = wcns_destroy(wcns_ctx_h)

contains

  subroutine read_namelists(fn)
    character(len=*), intent(in) :: fn
    integer :: u1, u2, ios2
    character(len=512) :: ln
    character(len=*), parameter :: tmpf = '.wcns_drive_param.nml'
    ! the reference's files mark groups with '$'; standard namelist input uses '&'
    open(newunit=u1, file=fn, status='old', action='read')
    open(newunit=u2, file=tmpf, status='replace')
    do
      read(u1, '(a)', iostat=ios2) ln
      if (ios2 /= 0) exit
      ln = adjustl(ln)
      if (ln(1:1) == '$') ln(1:1) = '&'

      write(u2, '(a)') trim(ln)
    end do
    close(u1)
    close(u2)
    open(newunit=u1, file=tmpf, status='old', action='read')
    rewind(u1); read(u1, nml=inflow, iostat=ios2)
    rewind(u1); read(u1, nml=force, iostat=ios2)
    rewind(u1); read(u1, nml=filename, iostat=ios2)
    rewind(u1); read(u1, nml=control, iostat=ios2)
    rewind(u1); read(u1, nml=step, iostat=ios2)
    rewind(u1); read(u1, nml=flowtype, iostat=ios2)
    rewind(u1); read(u1, nml=turbulence, iostat=ios2)
    rewind(u1); read(u1, nml=technic, iostat=ios2)
    rewind(u1); read(u1, nml=interplate, iostat=ios2)
    close(u1, status='delete')
  end subroutine read_namelists

  ! saveres2file (PostProcess.f90:1349-1378) every nwerror steps, and the full-precision line
  subroutine write_monitor(it, cpu)
    integer, intent(in) :: it
    real(8), intent(in) :: cpu
    write(202, '(i8,2es26.17e3,5i6)') it, dres_add, resmax_total, n_error(5), n_error(1:4)
    if ((it / nwerror) * nwerror == it) then
      write(201, 10) it, cpu, cfl, dres_add, resmax_total, n_error(5), n_error(1:4)
      write(*, 10) it, cpu / 3600., cfl, dres_add, resmax_total, n_error(5), n_error(1:4)
    end if
10  format(i7, 2x, e10.3, f9.3, 2x, e11.4, 1x, e11.4, 1x, 5(1x, i4))
  end subroutine write_monitor

  ! post_processing (PostProcess.f90:1-96): flow file every ndisk steps and at
  ! the end, force every nforce steps
  subroutine post(it, ng)
    integer, intent(in) :: it, ng
    integer(c_int) :: rc2
    real(8) :: cpu
    if ((it / ndisk) * ndisk == it .or. ng == nomax) then
      rc2 = wcns_save_flow(wcns_ctx_h, trim(flowname) // c_null_char, int(it, c_int), wholetime)
      if (rc2 /= 0) call wcns_abort('wcns_save_flow', rc2)
      ! the Tecplot files of nplot (PostProcess.f90:37-55)
      select case (nplot)
      case (0, 3)
        call tecplot(WCNS_TEC_SURFACE)
        call tecplot(WCNS_TEC_WALL)
      case (1)
        call tecplot(WCNS_TEC_WALL)
      case (2)
        call tecplot(WCNS_TEC_VOLUME)
      case (4, 5)
        write(*, *) 'nplot = ', nplot, ': the two-dimensional Tecplot output is not supported'
      case default
        write(*, *) 'nplot = ', nplot, ': no such output option'
      end select
    end if
    if ((it / nforce) * nforce == it) then
      rc2 = wcns_force(wcns_ctx_h, fc)
      if (rc2 /= 0) call wcns_abort('wcns_force', rc2)
      call cpu_time(cpu)
      write(203, 20) it, cpu - cpu0, fc(1:9)
20    format(1x, i6, 2x, e10.4, 9(f12.5, 1x))
    end if
  end subroutine post

  ! one Tecplot file, named as the reference names it: tecname for the
  ! surfaces, its stem + "_wall.plt" / "_vol.plt" (PostProcess.f90:375-380, 626-631)
  subroutine tecplot(kind)
    integer(c_int), intent(in) :: kind
    integer(c_int) :: rc2
    character(len=130) :: fname
    integer :: m
    m = index(tecname, '.', back=.true.)
    if (kind == WCNS_TEC_SURFACE) then
      fname = tecname
    else if (m == 0) then
      fname = trim(adjustl(tecname))
    else
      fname = trim(adjustl(tecname(1:m-1)))
    end if
    if (kind == WCNS_TEC_WALL) fname = trim(fname) // '_wall.plt'
    if (kind == WCNS_TEC_VOLUME) fname = trim(fname) // '_vol.plt'
    rc2 = wcns_tecplot(wcns_ctx_h, trim(fname) // c_null_char, kind)
    if (rc2 /= 0) call wcns_abort('wcns_tecplot', rc2)
  end subroutine tecplot
end program wcns_drive
