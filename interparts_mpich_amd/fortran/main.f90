! main.f90 -- program interpt on one MI355X: the time loop of the reference's
! src/main.f90 (start-up 98-227, substep 268-414, dt policy 425-441) over the
! MI355X library.  Inputs: parameters.inp (&PARAMETERS of src/param.f90),
! run.nml (&GRIDSIZE: the compile-time sizes of src/gridsize.f90, and
! &RUNCONTROL), lagrangianforcepoints2 when np > 0, fld%07d when begin > 0.
! Outputs: timestep.txt (src/main.f90:425-431); meanforces and IBM_forcing.out,
! a line each per IBM pass (updtintermediatevel, src/main.f90:315; run.nml
! ibm_log = .false. leaves them out); info_1dstat.out and partvisu%07d.bin every iout1d
! steps, partstate%07d and allpartdata%07d (src/loadd.f90:121-692) every ioutchk steps with particles, the post3d fields
! name%07d_3d.bin every iout2d steps (src/main.f90:443-472); fld%07d
! (src/loadd.f90 format), partstate%07d and allpartdata%07d when the run stops (step nstep,
! t_end or tmax hours, src/main.f90:498-512).  nstep is the last step index, as
! in `do istep = begin+1,nstep` (src/main.f90:252).
program interpt
  use mod_param
  use mod_common
  use mod_initmpi
  use mod_init
  use mod_initparticles
  use mod_coordsfp
  use mod_initsolver
  use mod_bound
  use mod_rk
  use mod_interp_spread
  use mod_forcing
  use mod_fillps
  use mod_solver
  use mod_correc
  use mod_intgr_over_sphere
  use mod_intgr_nwtn_eulr
  use mod_chkdt
  use mod_chkdiv
  use mod_loadflds
  use mod_loadpart
  use mod_output
  implicit none
  real :: dtmax, maxerror, avererror, t1, sumk
  integer :: ibmiter, u, last
  integer(8) :: cs, c0, c1, crate
  integer(c_int) :: kbtune
  character(len=7) :: istepchar
  logical :: have_partstate
  real, parameter :: tmax = 23.5   ! hours, src/main.f90:140

  call system_clock(cs, crate)
  call input_parameters
  call initmpi
  if (np > 0) then
    write(6, *) 'Number of Lagrangian Force Points of first shell = ', nl
  end if
  if (begin == 0) then
    call init_ics
    if (np > 0) then
      ! src/main.f90:142-145
      call ipg_ok(ipg_kerneltest(sumk), 'kerneltest')
      write(6, *) 'Integral over kernel = ', sumk
      call initparticles
    end if
    time = 0.
    call bounduvw(FLD_NEW)
    call boundscal
    call chkdt(dtmax)
    dt = 0.5 * dtmax
  else
    if (np > 0) then
      write(istepchar, '(i7.7)') begin
      inquire(file='partstate'//istepchar, exist=have_partstate)
      if (have_partstate) then
        call loadpart(0, begin)
      else
        call loadpart_ref(0, begin)
      end if
    end if
    call loadflds(0, begin)
    call chkdt(dtmax)
    write(6, *) 'nr steps at beginning simulation = ', begin
  end if
  call ipg_ok(ipg_set_dt(ctx, dt), 'set_dt')
  call initsolver
  call bounduvw(FLD_NEW)
  call boundscal
  call boundp(FLD_NEW)
  call chkdiv
  if (np > 0) then
    call coordsfp
    if (begin == 0) then
      call intgr_over_sphere(1)
      call intgr_over_sphere(2)
      call intgr_over_sphere(3)
    end if
  end if
  write(6, *) 'dtmax = ', dtmax, ' dt = ', dt
  ! launch shape of the RK stage for this grid (overwrites dudt.. only)
  call ipg_ok(ipg_autotune(ctx, kbtune), 'autotune')

  call system_clock(c0, crate)
  last = begin
  do istep = begin + 1, nstep
    if (time + dt > t_end) then
      dt = t_end - time
      call ipg_ok(ipg_set_dt(ctx, dt), 'set_dt')
    end if
    time = time + dt
    do rkiter = 1, 3
      if (fused_substep) then
        ! the whole substep below as one launch sequence (one hipGraph per stage)
        call ipg_ok(ipg_substep(ctx, int(rkiter, c_int)), 'substep')
        ! updtintermediatevel's lines of the substep's three IBM passes
        if (np > 0 .and. ibm_log) then
          do ibmiter = 0, 2
            curiter = ibmiter
            call updtintermediatevel(istep)
          end do
        end if
        cycle
      end if
      if (rkiter == 1) call rk1
      if (rkiter == 2) call rk2
      if (rkiter == 3) call rk3
      if (np > 0) then
        call ibm_begin
        do ibmiter = 0, 2
          call eulr2lagr(ibmiter)
          if (ibmiter == 0) then
            call complagrforces
          else
            call updtlagrforces(maxerror, avererror)
          end if
          call lagr2eulr
          if (ibm_log) call updtintermediatevel(istep)
        end do
      end if
      call bounduvw(FLD_DT)
      call fillps(rkiter)
      call solver2d
      call boundp(FLD_DT)
      call correc(rkiter)
      call bounduvw(FLD_NEW)
      call pnew_update
      if (np > 0) call intgr_nwtn_eulr
    end do
    if (mod(istep, ioutchk) == 0) then
      call chkdiv
      call chkdt(dtmax)
      write(6, *) 'dtmax = ', dtmax, ' dt = ', dt
      open(newunit=u, file='timestep.txt', position='append')
      write(u, '(3E16.8)') 1. * istep, time, dt
      close(u)
      dt = 0.75 * dtmax
      call ipg_ok(ipg_set_dt(ctx, dt), 'set_dt')
    end if
    last = istep
    if (mod(istep, iout1d) == 0) call post1d(time, istep)
    if (np > 0 .and. mod(istep, ioutchk) == 0) then
      call loadpart(1, last)
      call loadpart_ref(1, last)
    end if
    if (mod(istep, iout2d) == 0) call post3d(istep)
    ! src/main.f90:498-512: after tmax hours, at step nstep or at t_end, write
    ! the particle and field state and leave the loop
    call system_clock(c1)
    if (real(c1 - cs) / real(crate) > tmax * 3600. .or. istep == nstep .or. time >= t_end) then
      call ipg_ok(ipg_sync(ctx), 'sync')
      call system_clock(c1)
      call dump(last)
      exit
    end if
  end do
  t1 = real(c1 - c0) / real(crate)
  if (last > begin) write(6, '(A,I0,A,F10.3,A,F12.3,A)') 'time loop: ', 3 * (last - begin), ' substeps in ', t1, &
       ' s (', 3 * (last - begin) / max(t1, 1.e-9), ' substeps/s)'
  ! no step ran: write the start-up state (the reference writes nothing then)
  if (last == begin) call dump(last)
  call ipg_ok(ipg_finalize(ctx), 'finalize')
contains
  subroutine dump(nr)
    integer, intent(inout) :: nr
    call loadflds(1, nr)
    if (np > 0) then
      call loadpart(1, nr)
      call loadpart_ref(1, nr)
    end if
  end subroutine dump
end program interpt
