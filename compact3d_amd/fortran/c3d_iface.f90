! c3d_iface.f90 -- ISO_C_BINDING interface to libc3d.so (include/c3d.h).
!
! A maintainer of compact3d adds this module to the build and replaces the
! bodies of GETUP, DIVCHECK, CFL and CHKMF by the wrappers in c3d_wrappers.f90;
! MAIN (compact3d.f:11-189) is unchanged apart from one C3D_BEGIN call after
! INIUP/READUP and C3D_END before the output routines.
module c3d_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: C3D_FP64 = 0, C3D_MIXED = 1

  type, bind(C) :: c3d_config
    integer(c_int) :: n1, n2, n3
    real(c_double) :: re, alx, aly, alz, dt, files, wmodel, cks
    type(c_ptr) :: y
    integer(c_int) :: precision, device
    type(c_ptr) :: stream
    real(c_double) :: cflmax
  end type c3d_config

  interface
    integer(c_int) function c3d_create(cfg, h) bind(C, name='c3d_create')
      import :: c3d_config, c_ptr, c_int
      type(c3d_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function c3d_destroy(h) bind(C, name='c3d_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    type(c_ptr) function c3d_last_error(h) bind(C, name='c3d_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function c3d_upload_state(h, u, p, hh, sgsvis, presg, presg3) &
        bind(C, name='c3d_upload_state')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u(*), p(*), hh(*), sgsvis(*)
      real(c_double), value :: presg, presg3
    end function
    integer(c_int) function c3d_download_state(h, u, p, hh, sgsvis, presg, presg3) &
        bind(C, name='c3d_download_state')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: u(*), p(*), hh(*), sgsvis(*)
      real(c_double), intent(out) :: presg, presg3
    end function
    integer(c_int) function c3d_getup(h, ntime) bind(C, name='c3d_getup')
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: ntime
    end function
    integer(c_int) function c3d_divcheck(h, divmax) bind(C, name='c3d_divcheck')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: divmax
    end function
    integer(c_int) function c3d_cfl(h, cflm) bind(C, name='c3d_cfl')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: cflm
    end function
    integer(c_int) function c3d_chkmf(h, flow1, flow3) bind(C, name='c3d_chkmf')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: flow1, flow3
    end function
    integer(c_int) function c3d_wallss(h) bind(C, name='c3d_wallss')
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function c3d_write_flow(h, path, ntime) bind(C, name='c3d_write_flow')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: ntime
    end function
    integer(c_int) function c3d_write_strain(h, path, ntime, utau) bind(C, name='c3d_write_strain')
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: ntime
      real(c_double), value :: utau
    end function
    integer(c_int) function c3d_get_scalar(h, which, v) bind(C, name='c3d_get_scalar')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: which
      real(c_double), intent(out) :: v
    end function
    integer(c_int) function c3d_set_scalar(h, which, v) bind(C, name='c3d_set_scalar')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: which
      real(c_double), value :: v
    end function
    integer(c_int) function c3d_set_dt(h, dt) bind(C, name='c3d_set_dt')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
    end function
  end interface

contains

  ! stops with the library's message on a nonzero status (the reference has
  ! no status codes; a failed step ends the run)
  subroutine c3d_check(rc, h, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: s
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c3d_last_error(h), msg, [512])
    s = ' '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      s(i:i) = msg(i)
    end do
    write(*,'(A,": ",A)') what, trim(s)
    stop 1
  end subroutine c3d_check

end module c3d_iface
