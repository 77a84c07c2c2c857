! mod_ipg_capi.f90 -- iso_c_binding interface to libipg.so (include/ipg.h).
!
! Every routine of the C-ABI returns 0 on success; ipg_ok turns a non-zero
! code into `error stop` with the library's message, the way the reference
! aborts inside the failing routine (src/chkdiv.f90:51-58).
module mod_ipg_capi
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: ipg_params, ipg_ok, c_ptr, c_null_ptr, c_int, c_double, c_loc, c_long_long
  public :: ipg_init, ipg_finalize, ipg_nl_formula, ipg_field_put, ipg_field_get, ipg_particles_put, &
            ipg_particles_get, ipg_set_dt, ipg_get_dt, ipg_initsolver, ipg_initparticles, ipg_coordsfp, ipg_rk, &
            ipg_ibm, ipg_ibm_begin, ipg_eulr2lagr, ipg_lagrforces, ipg_lagr2eulr, ipg_bounduvw, ipg_boundp, &
            ipg_boundscal, ipg_fillps, ipg_solver2d, ipg_correc, ipg_pnew_update, ipg_intgr_over_sphere, &
            ipg_intgr_nwtn_eulr, ipg_substep, ipg_chkdt, ipg_chkdiv, ipg_uniform, ipg_sync, ipg_use_graphs, &
            ipg_use_side_stream, ipg_autotune, ipg_tuning, ipg_post3d, ipg_kerneltest, ipg_ibm_forces, &
            ipg_set_mixed_solve, ipg_set_refinement
  ! the batched ensemble (no reference counterpart: many cases through each
  ! substep together, include/ipg.h)
  public :: ipg_batch_create, ipg_batch_destroy, ipg_batch_substep, ipg_batch_sync, ipg_batch_autotune

  ! field selectors (enum of include/ipg.h)
  integer(c_int), parameter, public :: IPG_U = 0, IPG_V = 1, IPG_W = 2, IPG_PNEW = 3, IPG_C = 4, &
       IPG_DUDT = 5, IPG_DVDT = 6, IPG_DWDT = 7, IPG_DCDT = 8, IPG_DURK = 9, IPG_DVRK = 10, IPG_DWRK = 11, &
       IPG_DCRK = 12, IPG_P = 13
  ! particle quantity selectors
  integer(c_int), parameter, public :: IPG_PX = 0, IPG_PY = 1, IPG_PZ = 2, IPG_PTHETA = 3, IPG_PPHI = 4, &
       IPG_PU = 5, IPG_PV = 6, IPG_PW = 7, IPG_POMX = 8, IPG_POMY = 9, IPG_POMZ = 10, IPG_POMTHETA = 11, &
       IPG_PINTU = 12, IPG_PINTV = 13, IPG_PINTW = 14, IPG_PINTOMX = 15, IPG_PINTOMY = 16, IPG_PINTOMZ = 17, &
       IPG_PCOLFX = 18, IPG_PCOLFY = 19, IPG_PCOLFZ = 20, IPG_PCOLTX = 21, IPG_PCOLTY = 22, IPG_PCOLTZ = 23, &
       IPG_PFXLTOT = 24, IPG_PFYLTOT = 25, IPG_PFZLTOT = 26, IPG_PTORQX = 27, IPG_PTORQY = 28, IPG_PTORQZ = 29, &
       IPG_PTORQTHETA = 30, IPG_PRKFX = 31, IPG_PRKFY = 32, IPG_PRKFZ = 33, IPG_PRKDUDT = 34, IPG_PRKDVDT = 35, &
       IPG_PRKDWDT = 36, IPG_NPART = 37
  ! post3d quantities
  integer(c_int), parameter, public :: IPG_POST_PRE = 0, IPG_POST_VEZ = 1, IPG_POST_CON = 2, IPG_POST_VOX = 3, &
       IPG_POST_VOY = 4, IPG_POST_VOZ = 5, IPG_POST_SOL = 6

  type, bind(C) :: ipg_params
    integer(c_int) :: itot, jtot, ktot, np, nl
    real(c_double) :: lx, ly, lz, visc, pran, beta, radius, ratiorho
    integer(c_int) :: particle_rotation, bc_uvw_top, bc_uvw_bot, bc_c_top, bc_c_bot
    real(c_double) :: bc_c_top_val, bc_c_bot_val
    integer(c_int) :: precision
  end type ipg_params

  interface
    integer(c_int) function ipg_init(prm, ctx) bind(C, name="ipg_init")
      import :: c_int, c_ptr, ipg_params
      type(ipg_params), intent(in) :: prm
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function ipg_finalize(ctx) bind(C, name="ipg_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function ipg_last_error() bind(C, name="ipg_last_error")
      import :: c_ptr
    end function
    integer(c_int) function ipg_nl_formula(itot, lx, radius) bind(C, name="ipg_nl_formula")
      import :: c_int, c_double
      integer(c_int), value :: itot
      real(c_double), value :: lx, radius
    end function
    integer(c_int) function ipg_field_put(ctx, field, host) bind(C, name="ipg_field_put")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function ipg_field_get(ctx, field, host) bind(C, name="ipg_field_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function ipg_particles_put(ctx, q, host) bind(C, name="ipg_particles_put")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: q
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function ipg_particles_get(ctx, q, host) bind(C, name="ipg_particles_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: q
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function ipg_set_dt(ctx, dt) bind(C, name="ipg_set_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function ipg_get_dt(ctx, dt) bind(C, name="ipg_get_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function ipg_initsolver(ctx) bind(C, name="ipg_initsolver")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ipg_initparticles(ctx, x, y, z) bind(C, name="ipg_initparticles")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*), y(*), z(*)
    end function
    integer(c_int) function ipg_coordsfp(ctx, thetarc, phirc, radfp) bind(C, name="ipg_coordsfp")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: thetarc(*), phirc(*)
      real(c_double), value :: radfp
    end function
    integer(c_int) function ipg_rk(ctx, rkiter) bind(C, name="ipg_rk")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: rkiter
    end function
    integer(c_int) function ipg_ibm(ctx) bind(C, name="ipg_ibm")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ipg_ibm_begin(ctx) bind(C, name="ipg_ibm_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ipg_eulr2lagr(ctx, ibmiter) bind(C, name="ipg_eulr2lagr")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ibmiter
    end function
    integer(c_int) function ipg_lagrforces(ctx, ibmiter) bind(C, name="ipg_lagrforces")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ibmiter
    end function
    integer(c_int) function ipg_lagr2eulr(ctx, ibmiter) bind(C, name="ipg_lagr2eulr")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ibmiter
    end function
    integer(c_int) function ipg_ibm_forces(ctx, tot) bind(C, name="ipg_ibm_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: tot(3, 0:2)
    end function
    integer(c_int) function ipg_kerneltest(sum) bind(C, name="ipg_kerneltest")
      import :: c_int, c_double
      real(c_double), intent(out) :: sum
    end function
    integer(c_int) function ipg_bounduvw(ctx, which) bind(C, name="ipg_bounduvw")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
    end function
    integer(c_int) function ipg_boundp(ctx, which) bind(C, name="ipg_boundp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
    end function
    integer(c_int) function ipg_boundscal(ctx) bind(C, name="ipg_boundscal")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ipg_fillps(ctx, rkiter) bind(C, name="ipg_fillps")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: rkiter
    end function
    integer(c_int) function ipg_solver2d(ctx) bind(C, name="ipg_solver2d")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ipg_correc(ctx, rkiter) bind(C, name="ipg_correc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: rkiter
    end function
    integer(c_int) function ipg_pnew_update(ctx) bind(C, name="ipg_pnew_update")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ipg_intgr_over_sphere(ctx, typ) bind(C, name="ipg_intgr_over_sphere")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: typ
    end function
    integer(c_int) function ipg_intgr_nwtn_eulr(ctx, rkiter) bind(C, name="ipg_intgr_nwtn_eulr")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: rkiter
    end function
    integer(c_int) function ipg_substep(ctx, rkiter) bind(C, name="ipg_substep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: rkiter
    end function
    integer(c_int) function ipg_chkdt(ctx, dtmax) bind(C, name="ipg_chkdt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dtmax
    end function
    integer(c_int) function ipg_chkdiv(ctx, divtot, divmax) bind(C, name="ipg_chkdiv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: divtot, divmax
    end function
    integer(c_int) function ipg_post3d(ctx, quantity, host) bind(C, name="ipg_post3d")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: quantity
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function ipg_uniform(seed, stream, first, count, out) bind(C, name="ipg_uniform")
      import :: c_int, c_long_long, c_double
      integer(c_long_long), value :: seed, stream, first, count
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function ipg_sync(ctx) bind(C, name="ipg_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ipg_use_graphs(ctx, on) bind(C, name="ipg_use_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function ipg_use_side_stream(ctx, on) bind(C, name="ipg_use_side_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    ! mixed precision: the Poisson solve (32 FP32-stored with refinement, 64 FP64)
    integer(c_int) function ipg_set_mixed_solve(ctx, bits) bind(C, name="ipg_set_mixed_solve")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: bits
    end function
    integer(c_int) function ipg_set_refinement(ctx, nref) bind(C, name="ipg_set_refinement")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nref
    end function
    integer(c_int) function ipg_autotune(ctx, kb) bind(C, name="ipg_autotune")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: kb
    end function
    integer(c_int) function ipg_tuning(ctx, rk_kb, sphere_blocks) bind(C, name="ipg_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: rk_kb, sphere_blocks
    end function
    integer(c_int) function ipg_batch_create(cases, ncases, batch) bind(C, name="ipg_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: cases(*)
      integer(c_int), value :: ncases
      type(c_ptr), intent(out) :: batch
    end function
    integer(c_int) function ipg_batch_destroy(batch) bind(C, name="ipg_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function
    integer(c_int) function ipg_batch_substep(batch, rkiter) bind(C, name="ipg_batch_substep")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: rkiter
    end function
    integer(c_int) function ipg_batch_sync(batch) bind(C, name="ipg_batch_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function
    integer(c_int) function ipg_batch_autotune(batch, substeps) bind(C, name="ipg_batch_autotune")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), intent(out) :: substeps
    end function
  end interface

contains

  ! Stop the run with the library's message when a call failed.
  subroutine ipg_ok(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n
    if (rc == 0) return
    text = ''
    call c_f_pointer(ipg_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
      text(n:n) = msg(n)
    end do
    write(6, '(A,A,A,I0,A,A)') 'ERROR in ', what, ' (code ', rc, '): ', trim(text)
    error stop 1
  end subroutine ipg_ok

end module mod_ipg_capi
