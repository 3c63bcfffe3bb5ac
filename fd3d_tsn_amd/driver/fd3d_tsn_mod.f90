! iso_c_binding interface of the MI355X time-stepping core
! (include/fd3d_tsn.h).  Every entry point takes plain pointers and
! integers, so each C function maps to one bind(C) interface here.
module fd3d_tsn_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FD3D_SW = 0, FD3D_FVW = 1
  integer(c_int), parameter :: FD3D_NOUT = 5, FD3D_NFINAL = 7

  ! mirrors struct fd3d_config
  type, bind(C) :: fd3d_config
    integer(c_int) :: nxt, nyt, nzt, nabc, fspace, friction, nucleation, nstations
    type(c_ptr) :: stations
    real(c_float) :: dh, dt, damp_s, pml_vp, pml_fact
    real(c_double) :: Sn, uini, wini, f0, v0, fw
    real(c_double) :: hx0, hz0, perturb, RR2, TT2, strinixI
    integer(c_int) :: newton_max, ring_steps
  end type

  interface
    integer(c_int) function fd3d_create(cfg, h) bind(C, name='fd3d_create')
      import :: c_int, c_ptr, fd3d_config
      type(fd3d_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function fd3d_upload_medium(h, lam1, mu1, d1) bind(C, name='fd3d_upload_medium')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(in) :: lam1(*), mu1(*), d1(*)
    end function
    integer(c_int) function fd3d_upload_fault_sw(h, striniX, striniZ, peak_xz, dyn_xz, Dc, coh) &
        bind(C, name='fd3d_upload_fault_sw')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(in) :: striniX(*), striniZ(*), peak_xz(*), dyn_xz(*), Dc(*), coh(*)
    end function
    integer(c_int) function fd3d_upload_fault_fvw(h, Dc, striniX, striniZ, a, b, psi, vw, aX, bX, psiX, vwX) &
        bind(C, name='fd3d_upload_fault_fvw')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: Dc(*), striniX(*), striniZ(*), a(*), b(*), psi(*), vw(*)
      real(c_double), intent(in) :: aX(*), bX(*), psiX(*), vwX(*)
    end function
    integer(c_int) function fd3d_init_state(h) bind(C, name='fd3d_init_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fd3d_step(h, nsteps, out, seis) bind(C, name='fd3d_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
      type(c_ptr), value :: out, seis
    end function
    integer(c_int) function fd3d_step_begin(h, nsteps) bind(C, name='fd3d_step_begin')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function fd3d_step_wait(h, nsteps, out, seis) bind(C, name='fd3d_step_wait')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nsteps
      type(c_ptr), intent(out) :: out, seis
    end function
    integer(c_int) function fd3d_download_final(h, out7) bind(C, name='fd3d_download_final')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(out) :: out7(*)
    end function
    integer(c_int) function fd3d_set_tuning(h, mode, tz_vel, ny_vel, tz_str, ny_str) &
        bind(C, name='fd3d_set_tuning')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode, tz_vel, ny_vel, tz_str, ny_str
    end function
    function fd3d_last_error(h) bind(C, name='fd3d_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
      type(c_ptr) :: fd3d_last_error
    end function
    subroutine fd3d_destroy(h) bind(C, name='fd3d_destroy')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
  end interface

contains

  ! stop with the library's message when a call failed
  subroutine fd3d_check(rc, h, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n
    if (rc == 0) return
    text = ''
    if (c_associated(h)) then
      call c_f_pointer(fd3d_last_error(h), msg, [512])
      do n = 1, 512
        if (msg(n) == c_null_char) exit
        text(n:n) = msg(n)
      end do
    end if
    print '(a,a,a,i0,a,a)', 'fd3d_tsn: ', what, ' failed, code ', rc, ': ', trim(text)
    stop 2
  end subroutine

end module fd3d_tsn_mod
