! mod_substep.f90 -- host-side modules of the MI355X build of the InterPartS
! substep.  They keep the module and procedure names of the reference's
! Fortran (ajayhudge/InterPartS_MPICH, src/*.f90) so that the time loop of
! program interpt reads as the reference's does; the bodies forward to the
! C-ABI of libipg.so (module mod_ipg_capi).  The fields, the particle state
! and the markers stay resident in HBM inside one context: the reference's
! array arguments (unew.., dudt.., p) become selectors, and host copies of a
! field exist only for input and output (mod_init, mod_loadflds).
!
! Compiled, as the reference is, with -fdefault-real-8: every real and every
! real literal is FP64.

! ---------------------------------------------------------------------------
! src/param.f90:19-86 (&PARAMETERS) and src/gridsize.f90:6-26, which are
! compile-time constants in the reference and namelist input here.
module mod_param
  use mod_ipg_capi, only: c_int
  implicit none
  real(8) :: lx = 1., ly = 1., lz = 1.
  character(len=9) :: iniu = 'zer'
  character(len=9) :: inic = 'ZERC'
  character(len=15) :: bc_uvw_top_type = 'BC_NOSLIP', bc_uvw_bot_type = 'BC_NOSLIP'
  character(len=15) :: bc_c_top_type = 'BC_DIRICHLET', bc_c_bot_type = 'BC_DIRICHLET'
  real(8) :: bc_c_top_val = 0., bc_c_bot_val = 0.
  real(8) :: visc = 3.21e-3, Pran = 1.0, beta = 0., radius = 0.5, ratiorho = 1.5
  logical :: particle_rotation = .false.
  integer :: runmode = 0
  integer :: ioutchk = 5, iout1d = 5, iout2d = 3000, ioutfld = 3000
  namelist /parameters/ lx, ly, lz, iniu, inic, bc_uvw_top_type, bc_uvw_bot_type, bc_c_top_type, &
       bc_c_bot_type, bc_c_top_val, bc_c_bot_val, visc, Pran, beta, radius, ratiorho, particle_rotation, &
       ioutchk, iout1d, iout2d, ioutfld, runmode
  ! src/gridsize.f90 (runtime here)
  integer :: itot = 64, jtot = 64, ktot = 64, np = 0, nlmax = 0
  real(8) :: t_end = 1.0d30
  namelist /gridsize/ itot, jtot, ktot, np, nlmax, t_end
  ! run control: the reference hard-codes begin = 0 and nstep (src/main.f90:137-138)
  ! and seeds with random_number; here a counter-based generator is keyed by seed
  integer :: begin = 0, nstep = 100
  integer(8) :: seed = 1
  real(8) :: perturb = 0.1
  logical :: use_graphs = .true., fused_substep = .true.
  ! the per-pass lines of updtintermediatevel (src/forcing.f90:430-468)
  logical :: ibm_log = .true.
  ! arithmetic (no reference counterpart): precision 0 FP64, 1 mixed (FP32 RK
  ! history; Poisson solve mixed_solve = 32, FP32-stored spectra with nref
  ! FP64-residual refinement sweeps, or 64, the FP64 solve)
  integer :: precision = 0, mixed_solve = 32, nref = 1
  namelist /runcontrol/ begin, nstep, seed, perturb, use_graphs, fused_substep, ibm_log, precision, mixed_solve, nref
  integer :: nl = 0
  integer :: imax, jmax, kmax, i1, j1, k1
  real(8), parameter :: pi = acos(-1.d0)
contains
  ! src/init.f90:10-72 (input_parameters): parameters.inp and run.nml
  subroutine input_parameters
    integer :: u, ios
    open(newunit=u, file='parameters.inp', status='old', action='read', iostat=ios)
    if (ios /= 0) error stop 'parameters.inp not found'
    read(u, nml=parameters)
    close(u)
    open(newunit=u, file='run.nml', status='old', action='read', iostat=ios)
    if (ios /= 0) error stop 'run.nml not found'
    read(u, nml=gridsize)
    rewind(u)
    read(u, nml=runcontrol, iostat=ios)
    close(u)
    imax = itot; jmax = jtot; kmax = ktot
    i1 = imax + 1; j1 = jmax + 1; k1 = kmax + 1
  end subroutine input_parameters
end module mod_param

! ---------------------------------------------------------------------------
! src/common.f90: the state that stays on the host (time, dt, the stage) and
! the context that owns everything else.
module mod_common
  use mod_ipg_capi
  implicit none
  type(c_ptr) :: ctx = c_null_ptr
  real(8) :: time = 0., dt = 0.
  integer :: rkiter = 1, istep = 0
  integer, parameter :: FLD_NEW = 0, FLD_DT = 1     ! bounduvw / boundp selectors
  real(8), allocatable :: hbuf(:,:,:)               ! (0:i1,0:j1,0:k1) host staging for I/O
end module mod_common

! src/initmpi.f90 + src/init.f90:74-243 for one GPU holding the whole domain
module mod_initmpi
  use mod_param
  use mod_common
  implicit none
contains
  subroutine initmpi
    type(ipg_params) :: prm
    character(len=1) :: s
    prm%itot = itot; prm%jtot = jtot; prm%ktot = ktot; prm%np = np
    nl = 0
    if (np > 0) nl = ipg_nl_formula(int(itot, c_int), lx, radius)
    ! src/main.f90:117-133: the run stops when nl /= nlmax
    if (np > 0 .and. nlmax > 0 .and. nl /= nlmax) then
      write(6,*) 'ERROR: Number of Lagrangian Force Points (nl =', nl, ') does not match expected value (nlmax =', &
           nlmax, ')'
      error stop 1
    end if
    prm%nl = nl
    prm%lx = lx; prm%ly = ly; prm%lz = lz
    prm%visc = visc; prm%pran = Pran; prm%beta = beta; prm%radius = radius; prm%ratiorho = ratiorho
    prm%particle_rotation = merge(1, 0, particle_rotation)
    prm%bc_uvw_top = merge(1, 0, trim(bc_uvw_top_type) == 'BC_FREESLIP')
    prm%bc_uvw_bot = merge(1, 0, trim(bc_uvw_bot_type) == 'BC_FREESLIP')
    prm%bc_c_top = merge(1, 0, trim(bc_c_top_type) == 'BC_NEUMANN')
    prm%bc_c_bot = merge(1, 0, trim(bc_c_bot_type) == 'BC_NEUMANN')
    prm%bc_c_top_val = bc_c_top_val; prm%bc_c_bot_val = bc_c_bot_val
    prm%precision = precision
    call ipg_ok(ipg_init(prm, ctx), 'ipg_init')
    if (precision == 1) then
      call ipg_ok(ipg_set_mixed_solve(ctx, int(mixed_solve, c_int)), 'set_mixed_solve')
      if (mixed_solve == 32) call ipg_ok(ipg_set_refinement(ctx, int(nref, c_int)), 'set_refinement')
    end if
    if (use_graphs) call ipg_ok(ipg_use_graphs(ctx, 1), 'ipg_use_graphs')
    allocate(hbuf(0:i1, 0:j1, 0:k1))
    s = ' '
  end subroutine initmpi
end module mod_initmpi

! src/init.f90:270-505 (init_ics): 'poi' / 'zer' profiles; the random
! perturbation of src/init.f90:287-305 is drawn from the counter-based
! generator, U[0,1) per (seed, 100 + component, interior cell index)
module mod_init
  use mod_param
  use mod_common
  implicit none
contains
  subroutine init_ics
    real(8), allocatable :: r(:)
    real(8) :: dz, coorz, prof
    integer :: i, j, k, comp
    integer(c_int) :: sel(3)
    integer(8) :: n
    dz = 1. / (kmax / lz)
    sel = [IPG_U, IPG_V, IPG_W]
    allocate(r(imax * jmax * kmax))
    do comp = 1, 3
      hbuf = 0.
      if (comp == 2 .and. trim(iniu) == 'poi') then
        do k = 0, k1
          coorz = (k - 0.5) * dz / lz
          prof = 6. * 1. * coorz * (1. - coorz)
          hbuf(:, :, k) = prof
        end do
      end if
      if (perturb /= 0.) then
        n = int(imax, 8) * jmax * kmax
        call ipg_ok(ipg_uniform(seed, int(99 + comp, 8), 0_8, n, r), 'ipg_uniform')
        do k = 1, kmax
          do j = 1, jmax
            do i = 1, imax
              n = (int(k - 1, 8) * jmax + (j - 1)) * imax + i
              hbuf(i, j, k) = hbuf(i, j, k) + perturb * (2. * r(n) - 1.)
            end do
          end do
        end do
      end if
      call ipg_ok(ipg_field_put(ctx, sel(comp), hbuf), 'field_put')
    end do
    hbuf = 0.
    call ipg_ok(ipg_field_put(ctx, IPG_PNEW, hbuf), 'field_put')
    call ipg_ok(ipg_field_put(ctx, IPG_C, hbuf), 'field_put')
  end subroutine init_ics
end module mod_init

! src/initparticles.f90: seeded placement by rejection sampling (centre gap
! >= D + h with the minimum image in x and y, wall gap >= R + h), draws
! (seed, 200, 3t .. 3t+2) for trial t; velocities and forces start at zero.
module mod_initparticles
  use mod_param
  use mod_common
  implicit none
contains
  subroutine initparticles
    real(8), allocatable :: xs(:), ys(:), zs(:)
    real(8) :: h, gap, zlo, zhi, r3(3), x, y, z, dx, dy, dzz
    integer(8) :: t
    integer :: n, q
    logical :: ok
    allocate(xs(np), ys(np), zs(np))
    h = lx / itot
    gap = 2. * radius + h
    zlo = radius + h
    zhi = lz - radius - h
    n = 0
    t = 0
    do while (n < np)
      if (t >= 2000000_8) error stop 'initparticles: could not place the spheres'
      call ipg_ok(ipg_uniform(seed, 200_8, 3 * t, 3_8, r3), 'ipg_uniform')
      t = t + 1
      x = lx * r3(1)
      y = ly * r3(2)
      z = zlo + (zhi - zlo) * r3(3)
      ok = .true.
      do q = 1, n
        dx = x - xs(q)
        dx = dx - lx * anint(dx / lx)
        dy = y - ys(q)
        dy = dy - ly * anint(dy / ly)
        dzz = z - zs(q)
        if (dx * dx + dy * dy + dzz * dzz < gap * gap) then
          ok = .false.
          exit
        end if
      end do
      if (ok) then
        n = n + 1
        xs(n) = x; ys(n) = y; zs(n) = z
      end if
    end do
    call ipg_ok(ipg_initparticles(ctx, xs, ys, zs), 'ipg_initparticles')
  end subroutine initparticles
end module mod_initparticles

! src/coordsfp.f90:10-106: read lagrangianforcepoints2 (3 header lines, nl
! records of 6E16.8: x, y, z, theta, phi, radfp) and set up the markers
module mod_coordsfp
  use mod_param
  use mod_common
  implicit none
contains
  subroutine coordsfp
    real(8), allocatable :: thetarc(:), phirc(:)
    real(8) :: xd, yd, zd, radfp
    integer :: u, l, ios
    allocate(thetarc(nl), phirc(nl))
    open(newunit=u, file='lagrangianforcepoints2', status='old', action='read', iostat=ios)
    if (ios /= 0) error stop 'lagrangianforcepoints2 not found'
    read(u, *)
    read(u, *)
    read(u, *)
    do l = 1, nl
      read(u, '(6E16.8)') xd, yd, zd, thetarc(l), phirc(l), radfp
    end do
    close(u)
    call ipg_ok(ipg_coordsfp(ctx, thetarc, phirc, radfp), 'ipg_coordsfp')
  end subroutine coordsfp
end module mod_coordsfp

module mod_initsolver
  use mod_common
  implicit none
contains
  subroutine initsolver                 ! src/initsolver.f90:10-67
    call ipg_ok(ipg_initsolver(ctx), 'initsolver')
  end subroutine
end module mod_initsolver

module mod_bound
  use mod_common
  implicit none
contains
  subroutine bounduvw(which)            ! src/bound.f90:11-80; FLD_NEW: unew.., FLD_DT: dudt..
    integer, intent(in) :: which
    call ipg_ok(ipg_bounduvw(ctx, int(which, c_int)), 'bounduvw')
  end subroutine
  subroutine boundp(which)              ! src/bound.f90:82-107; FLD_NEW: pnew, FLD_DT: p
    integer, intent(in) :: which
    call ipg_ok(ipg_boundp(ctx, int(which, c_int)), 'boundp')
  end subroutine
  subroutine boundscal                  ! src/bound.f90:109-160 on cnew
    call ipg_ok(ipg_boundscal(ctx), 'boundscal')
  end subroutine
end module mod_bound

module mod_rk
  use mod_common
  implicit none
contains
  subroutine rk1                        ! src/rk3.f90:13-106
    call ipg_ok(ipg_rk(ctx, 1_c_int), 'rk1')
  end subroutine
  subroutine rk2                        ! src/rk3.f90:108-222
    call ipg_ok(ipg_rk(ctx, 2_c_int), 'rk2')
  end subroutine
  subroutine rk3                        ! src/rk3.f90:224-329
    call ipg_ok(ipg_rk(ctx, 3_c_int), 'rk3')
  end subroutine
end module mod_rk

! src/interp_spread.f90 and src/forcing.f90.  ibmiter of the current pass is
! kept here, as lagr2eulr of the reference takes no argument.
module mod_interp_spread
  use mod_common
  implicit none
  integer :: curiter = 0
contains
  subroutine eulr2lagr(ibmiter)         ! src/interp_spread.f90:12-228
    integer, intent(in) :: ibmiter
    curiter = ibmiter
    call ipg_ok(ipg_eulr2lagr(ctx, int(ibmiter, c_int)), 'eulr2lagr')
  end subroutine
  subroutine lagr2eulr                  ! src/interp_spread.f90:230-474
    call ipg_ok(ipg_lagr2eulr(ctx, int(curiter, c_int)), 'lagr2eulr')
  end subroutine
end module mod_interp_spread

module mod_forcing
  use mod_param, only: itot, jtot, kmax
  use mod_common
  use mod_interp_spread, only: curiter
  implicit none
contains
  subroutine ibm_begin                  ! dudtf = dudt, src/main.f90:301-303
    call ipg_ok(ipg_ibm_begin(ctx), 'ibm_begin')
  end subroutine
  subroutine complagrforces             ! src/forcing.f90:9-187
    call ipg_ok(ipg_lagrforces(ctx, int(curiter, c_int)), 'complagrforces')
  end subroutine
  subroutine updtlagrforces(maxerror, avererror)  ! src/forcing.f90:189-391 (errors are always 0 there)
    real(8), intent(out) :: maxerror, avererror
    call ipg_ok(ipg_lagrforces(ctx, int(curiter, c_int)), 'updtlagrforces')
    maxerror = 0.
    avererror = 0.
  end subroutine
  ! src/forcing.f90:393-513, after IBM pass curiter: one line appended to
  ! meanforces (istep, time/tref and the mean of forcex, forcey, forcez over
  ! the interior, arrays the reference never writes: zero) and one to
  ! IBM_forcing.out (istep, time and the sums over the spheres of fxltot,
  ! fyltot, fzltot after the pass).  The velocity update the routine once
  ! made is commented out in the reference.
  subroutine updtintermediatevel(istep)
    integer, intent(in) :: istep
    real(8), parameter :: tref = 1.                   ! src/init.f90:102
    real(8) :: tot(3, 0:2), sumfx, sumfy, sumfz
    integer :: u
    call ipg_ok(ipg_ibm_forces(ctx, tot), 'updtintermediatevel')
    sumfx = 0. / (1. * itot * jtot * kmax)
    sumfy = sumfx
    sumfz = sumfx
    open(newunit=u, file='meanforces', position='append')
    write(u, '(I8,4E16.8)') istep, time / tref, sumfx, sumfy, sumfz
    close(u)
    open(newunit=u, file='IBM_forcing.out', position='append')
    write(u, '(I8,4E16.8)') istep, time, tot(1, curiter), tot(2, curiter), tot(3, curiter)
    close(u)
  end subroutine
end module mod_forcing

module mod_fillps
  use mod_common
  implicit none
contains
  subroutine fillps(rk)                 ! src/fillps.f90:8-53, p out
    integer, intent(in) :: rk
    call ipg_ok(ipg_fillps(ctx, int(rk, c_int)), 'fillps')
  end subroutine
end module mod_fillps

module mod_solver
  use mod_common
  implicit none
contains
  subroutine solver2d                   ! src/solver.f90:43-172 on p
    call ipg_ok(ipg_solver2d(ctx), 'solver2d')
  end subroutine
end module mod_solver

module mod_correc
  use mod_common
  implicit none
contains
  subroutine correc(rk)                 ! src/correc.f90:11-42
    integer, intent(in) :: rk
    call ipg_ok(ipg_correc(ctx, int(rk, c_int)), 'correc')
  end subroutine
  subroutine pnew_update                ! src/main.f90:370-376: pnew += p, boundp, cnew = dcdt, boundscal
    call ipg_ok(ipg_pnew_update(ctx), 'pnew_update')
  end subroutine
end module mod_correc

module mod_intgr_over_sphere
  use mod_common
  implicit none
contains
  subroutine intgr_over_sphere(typ)     ! src/intgr_over_sphere.f90:12-625
    integer, intent(in) :: typ
    call ipg_ok(ipg_intgr_over_sphere(ctx, int(typ, c_int)), 'intgr_over_sphere')
  end subroutine
end module mod_intgr_over_sphere

module mod_intgr_nwtn_eulr
  use mod_common
  implicit none
contains
  subroutine intgr_nwtn_eulr            ! src/intgr_nwtn_eulr.f90:12-714 at stage rkiter
    call ipg_ok(ipg_intgr_nwtn_eulr(ctx, int(rkiter, c_int)), 'intgr_nwtn_eulr')
  end subroutine
end module mod_intgr_nwtn_eulr

module mod_chkdt
  use mod_common
  implicit none
contains
  subroutine chkdt(dtmax)               ! src/chkdt.f90:6-271
    real(8), intent(out) :: dtmax
    call ipg_ok(ipg_chkdt(ctx, dtmax), 'chkdt')
  end subroutine
end module mod_chkdt

module mod_chkdiv
  use mod_common
  implicit none
contains
  subroutine chkdiv                     ! src/chkdiv.f90:7-61
    real(8) :: divtot, divmax
    call ipg_ok(ipg_chkdiv(ctx, divtot, divmax), 'chkdiv')
    write(6, *) 'Total divergence = ', divtot, '| Maximum divergence = ', divmax
    if (abs(divtot) > 1.e-3) then
      write(6, *) 'Fatal error: total divergence > 1.e-3. Program aborted...'
      error stop 1
    end if
  end subroutine
end module mod_chkdiv

! src/loadd.f90:12-68 (loadflds): fld%07d holds unew, vnew, wnew, pnew, cnew
! over the interior (itot x jtot x ktot, i fastest, FP64, no record markers)
! followed by (time, istep, dt)
module mod_loadflds
  use mod_param
  use mod_common
  implicit none
contains
  subroutine loadflds(in, nr)
    integer, intent(in) :: in
    integer, intent(inout) :: nr
    character(len=7) :: istepchar
    integer(c_int) :: sel(5)
    real(8) :: fldinfo(3)
    integer :: u, q
    sel = [IPG_U, IPG_V, IPG_W, IPG_PNEW, IPG_C]
    write(istepchar, '(i7.7)') nr
    if (in == 0) then
      open(newunit=u, file='fld'//istepchar, access='stream', form='unformatted', status='old', action='read')
      do q = 1, 5
        hbuf = 0.
        read(u) hbuf(1:imax, 1:jmax, 1:kmax)
        call ipg_ok(ipg_field_put(ctx, sel(q), hbuf), 'field_put')
      end do
      read(u) fldinfo
      close(u)
      time = fldinfo(1)
      nr = int(fldinfo(2))
      dt = fldinfo(3)
    else
      open(newunit=u, file='fld'//istepchar, access='stream', form='unformatted', status='replace', action='write')
      do q = 1, 5
        call ipg_ok(ipg_field_get(ctx, sel(q), hbuf), 'field_get')
        write(u) hbuf(1:imax, 1:jmax, 1:kmax)
      end do
      fldinfo(1) = time
      fldinfo(2) = 1. * nr
      fldinfo(3) = dt
      write(u) fldinfo
      close(u)
    end if
  end subroutine loadflds
end module mod_loadflds

! Particle state dump and restart: partstate%07d holds the IPG_NPART
! quantities of type particle kept on the device (x y z theta phi u v w omx
! omy omz omtheta intu intv intw intomx intomy intomz colfx colfy colfz coltx
! colty coltz fxltot fyltot fzltot torqx torqy torqz torqtheta), each as np
! FP64 values, quantity after quantity, then (time, istep, dt): the complete
! state, from which a restart continues exactly.  The reference's own file,
! allpartdata%07d of src/loadd.f90:121-692, is written beside it
! (loadpart_ref) and read when no partstate%07d is present.
module mod_loadpart
  use mod_param
  use mod_common
  implicit none
contains
  subroutine loadpart(in, nr)
    integer, intent(in) :: in
    integer, intent(inout) :: nr
    character(len=7) :: istepchar
    real(8), allocatable :: q(:, :)
    real(8) :: info(3)
    integer :: u, m
    allocate(q(np, 0:IPG_NPART - 1))
    write(istepchar, '(i7.7)') nr
    if (in == 0) then
      open(newunit=u, file='partstate'//istepchar, access='stream', form='unformatted', status='old', action='read')
      read(u) q
      read(u) info
      close(u)
      call ipg_ok(ipg_initparticles(ctx, q(:, IPG_PX), q(:, IPG_PY), q(:, IPG_PZ)), 'ipg_initparticles')
      do m = 0, IPG_NPART - 1
        call ipg_ok(ipg_particles_put(ctx, int(m, c_int), q(:, m)), 'particles_put')
      end do
    else
      do m = 0, IPG_NPART - 1
        call ipg_ok(ipg_particles_get(ctx, int(m, c_int), q(:, m)), 'particles_get')
      end do
      info(1) = time
      info(2) = 1. * nr
      info(3) = dt
      open(newunit=u, file='partstate'//istepchar, access='stream', form='unformatted', status='replace', &
           action='write')
      write(u) q
      write(u) info
      close(u)
    end if
  end subroutine loadpart

  ! allpartdata%07d of src/loadd.f90:121-692 on one rank.  A particle is a
  ! record of type particle_restart (lines 123-138), nrec = 31 + 11 nqmax
  ! reals, nqmax = min(np + 2, 15) (src/param.f90:100): x .. intomz (18),
  ! flag1, flag2, colfx .. coltz (6), fxltot, fyltot, fzltot, then the contact
  ! history dx, dy, dz, dxt, dyt, dzt, dut, dvt, dwt, psi (nqmax each), qmax,
  ! idp and firstc (nqmax).  The reference moves skip = 29 + 11 nqmax reals
  ! per particle over that record array (lines 140-142, 206, 690), so its file
  ! holds the first (29 + 11 nqmax) np reals of the array and the last 2 np are
  ! never written; this writes the same bytes, the particles in id order (the
  ! order of the masters on one rank).  Slots the reference leaves unassigned
  ! (flag1, flag2, dut, dvt, dwt, psi) are 0, and so are the contact history,
  ! qmax and firstc, which this build does not keep (the tangential force they
  ! feed never reaches the state, src/collisions.f90:199-205).
  ! Reading follows lines 190-339: each particle takes the first
  ! 24 + 10 nqmax reals of its record in the layout of type particle_old
  ! (src/common.f90:74-88), which has no flag1, flag2: colfx, colfy, colfz,
  ! coltx, colty, coltz are read from the record's flag1, flag2, colfx, colfy,
  ! colfz, coltx slots, as the reference reads them; fxltot.. are not read.
  ! Where the missing tail has taken a record's idp, the record's position
  ! gives it (the reference would abort there); a particle whose state lies
  ! in the missing tail cannot be read, and the run stops, as the reference's
  ! does.
  subroutine loadpart_ref(in, nr)
    integer, intent(in) :: in
    integer, intent(inout) :: nr
    character(len=7) :: istepchar
    real(8), allocatable :: rec(:), q(:, :)
    integer :: nq, nrec, nfile, u, p, m, idp, base, ios
    integer(8) :: fsize
    logical, allocatable :: seen(:)
    nq = min(np + 2, 15)
    nrec = 31 + 11 * nq
    nfile = 29 + 11 * nq
    allocate(rec(nrec * np), q(np, 0:IPG_NPART - 1), seen(np))
    rec = 0.
    q = 0.
    write(istepchar, '(i7.7)') nr
    if (in == 0) then
      open(newunit=u, file='allpartdata'//istepchar, access='stream', form='unformatted', status='old', &
           action='read', iostat=ios)
      if (ios /= 0) error stop 'allpartdata: file not found'
      inquire(unit=u, size=fsize)
      if (fsize /= 8_8 * nfile * np) then
        write(6, *) 'allpartdata: ', fsize, ' bytes, expected ', 8_8 * nfile * np
        error stop 'allpartdata: wrong size for np'
      end if
      read(u) rec(1:nfile * np)
      close(u)
      seen = .false.
      do p = 1, np
        base = (p - 1) * nrec
        ! the state a particle takes: reals 1 .. 24 + 10 nqmax of its record
        if (base + 24 + 10 * nq > nfile * np) then
          write(6, *) 'Fatal error in initialisation of particle positions: allpartdata holds no state for', &
               ' the particle of record ', p
          error stop 1
        end if
        idp = p
        if (base + 29 + 10 * nq + 2 <= nfile * np) idp = nint(rec(base + 29 + 10 * nq + 2))
        if (idp < 1 .or. idp > np) error stop 'allpartdata: particle id out of range'
        if (seen(idp)) error stop 'allpartdata: particle id repeated'
        seen(idp) = .true.
        ! src/loadd.f90:194-205
        if (rec(base + 1) < 0. .or. rec(base + 1) > lx .or. rec(base + 2) < 0. .or. rec(base + 2) > ly .or. &
            rec(base + 3) < 0. .or. rec(base + 3) > lz) then
          write(6, *) 'Fatal error in initialisation of particle positions - '
          write(6, *) 'particle outside the domain!'
          error stop 1
        end if
        do m = 0, 17
          q(idp, m) = rec(base + 1 + m)
        end do
        do m = 0, 5
          q(idp, IPG_PCOLFX + m) = rec(base + 19 + m)
        end do
      end do
      call ipg_ok(ipg_initparticles(ctx, q(:, IPG_PX), q(:, IPG_PY), q(:, IPG_PZ)), 'ipg_initparticles')
      do m = 0, IPG_NPART - 1
        call ipg_ok(ipg_particles_put(ctx, int(m, c_int), q(:, m)), 'particles_put')
      end do
    else
      do m = 0, IPG_NPART - 1
        call ipg_ok(ipg_particles_get(ctx, int(m, c_int), q(:, m)), 'particles_get')
      end do
      do p = 1, np
        base = (p - 1) * nrec
        do m = 0, 17
          rec(base + 1 + m) = q(p, m)
        end do
        do m = 0, 5
          rec(base + 21 + m) = q(p, IPG_PCOLFX + m)
        end do
        do m = 0, 2
          rec(base + 27 + m) = q(p, IPG_PFXLTOT + m)
        end do
        rec(base + 29 + 10 * nq + 2) = 1. * p
      end do
      open(newunit=u, file='allpartdata'//istepchar, access='stream', form='unformatted', status='replace', &
           action='write')
      write(u) rec(1:nfile * np)
      close(u)
    end if
  end subroutine loadpart_ref
end module mod_loadpart

! src/output.f90: post1d (13-593), whose files are info_1dstat.out and, through
! outputpart, partvisu%07d.bin (its profile statistics are computed and not
! written), and post3d (997-1074):
! pre, vez, con, vox, voy and the phase indicator sol
! (src/phase_indicator.f90:11-559, itype = 1) as name%07d_3d.bin, the
! interior in Fortran order (decomp_2d_write_every with no skipping),
! computed on the device.
module mod_output
  use mod_param
  use mod_common
  implicit none
contains
  subroutine post1d(tim, nr)
    real(8), intent(in) :: tim
    integer, intent(in) :: nr
    integer :: u
    real(8), parameter :: uref = 1., lref = 1., tref = 1.   ! src/init.f90:100-102
    open(newunit=u, file='info_1dstat.out')
    write(u, '(8E15.7)') 1. * nr, tim, visc, 1. * ktot, lx / itot, uref, lref, tref
    close(u)
    call outputpart(nr)
  end subroutine post1d

  ! src/output.f90:1127-1273: partvisu%07d.bin holds fltot(3,np) and then
  ! accel(3,np), the substep sums rkp%fxltot.. and rkp%dudt.. of the last step,
  ! spheres in id order, native real(8).
  subroutine outputpart(nr)
    integer, intent(in) :: nr
    character(len=7) :: istepchar
    real(8), allocatable :: fltot(:, :), accel(:, :), a(:)
    integer :: u, m
    integer(c_int), parameter :: sf(3) = [IPG_PRKFX, IPG_PRKFY, IPG_PRKFZ], sa(3) = [IPG_PRKDUDT, IPG_PRKDVDT, IPG_PRKDWDT]
    allocate(fltot(3, np), accel(3, np), a(max(np, 1)))
    do m = 1, 3
      if (np > 0) then
        call ipg_ok(ipg_particles_get(ctx, sf(m), a), 'outputpart')
        fltot(m, :) = a(1:np)
        call ipg_ok(ipg_particles_get(ctx, sa(m), a), 'outputpart')
        accel(m, :) = a(1:np)
      end if
    end do
    write(istepchar, '(i7.7)') nr
    open(newunit=u, file='partvisu'//istepchar//'.bin', access='stream', form='unformatted', status='replace', &
         action='write')
    write(u) fltot, accel
    close(u)
  end subroutine outputpart

  subroutine post3d(nr)
    integer, intent(in) :: nr
    character(len=7) :: istepchar
    character(len=3), parameter :: names(6) = ['pre', 'vez', 'con', 'vox', 'voy', 'sol']
    integer(c_int), parameter :: sel(6) = [IPG_POST_PRE, IPG_POST_VEZ, IPG_POST_CON, IPG_POST_VOX, IPG_POST_VOY, &
         IPG_POST_SOL]
    real(8), allocatable :: var(:, :, :)
    integer :: u, q
    allocate(var(imax, jmax, kmax))
    write(istepchar, '(i7.7)') nr
    do q = 1, 6
      call ipg_ok(ipg_post3d(ctx, sel(q), var), 'post3d')
      open(newunit=u, file=names(q)//istepchar//'_3d.bin', access='stream', form='unformatted', status='replace', &
           action='write')
      write(u) var
      close(u)
    end do
  end subroutine post3d
end module mod_output
