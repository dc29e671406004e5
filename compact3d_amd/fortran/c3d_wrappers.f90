! c3d_wrappers.f90 -- the reference's per-step entry points, redirected to the
! device.  Argument lists are those of compact3d.f:
!   GETUP(U,P,HH,SS,NTIME,PRESG,PRESG3,SGSVIS)   :879
!   DIVCHECK(U,TIME,DIVMAX,NTIME)                :9049
!   CFL(U,CFLM)                                  :9138
!   CHKMF(U,TIME)                                :9463
!   WALLSS(U,P,TIME)                             :11077
!   WRITEUP(U,P,PRESG,PRESG3,NTIME)              :757
! The state is resident in HBM between calls: C3D_BEGIN uploads it once and
! C3D_END (or C3D_SYNC) copies it back into the caller's arrays.  The array
! arguments of the per-step routines are not read; they are kept so that
! MAIN's calls stay unchanged.  GETUP writes the state back into its U, P,
! HH and SGSVIS arguments every C3D_SET_HOSTSYNC(NEVERY) steps (NTIME a
! multiple of NEVERY; default 0 = never), for MAIN's host routines that read
! U between steps (ENERGY, WALLSS_WRITE, compact3d.f:114-132).
module c3d_state
  use iso_c_binding
  use c3d_iface
  implicit none
  type(c_ptr), save :: h3d = c_null_ptr
  integer, save :: m1 = 0, m2 = 0, m3 = 0
  integer, save :: hostsync = 0
  real(c_double), save :: flow1 = 0.0d0, flow3 = 0.0d0
end module c3d_state

! CFLMAX is /TSTEP/'s (compact3d.f:247); the integral wall model (WMODEL = 2)
! takes its first time-filter weight from it (:3700) and needs it > 0
subroutine C3D_BEGIN(N1, N2, N3, RE, ALX, ALY, ALZ, DT, FILES, WMODEL, CKS, CFLMAX, PRECISION, &
                     U, P, HH, SGSVIS, PRESG, PRESG3)
  use c3d_state
  implicit none
  integer, intent(in) :: N1, N2, N3, PRECISION
  real(8), intent(in) :: RE, ALX, ALY, ALZ, DT, FILES, WMODEL, CKS, CFLMAX, PRESG, PRESG3
  real(8), intent(in) :: U(0:N1, 0:N2, 0:N3, 3), P(N1, N2, N3), HH(0:N1, 0:N2, 0:N3, 3)
  real(8), intent(in) :: SGSVIS(0:N1, 0:N2, 0:N3)
  type(c3d_config) :: cfg
  integer(c_int) :: rc
  cfg%n1 = N1; cfg%n2 = N2; cfg%n3 = N3
  cfg%re = RE; cfg%alx = ALX; cfg%aly = ALY; cfg%alz = ALZ; cfg%dt = DT
  cfg%files = FILES; cfg%wmodel = WMODEL; cfg%cks = CKS
  cfg%y = c_null_ptr; cfg%precision = PRECISION; cfg%device = 0; cfg%stream = c_null_ptr
  cfg%cflmax = CFLMAX
  rc = c3d_create(cfg, h3d)
  call c3d_check(rc, c_null_ptr, 'c3d_create')
  m1 = N1; m2 = N2; m3 = N3
  rc = c3d_upload_state(h3d, U, P, HH, SGSVIS, PRESG, PRESG3)
  call c3d_check(rc, h3d, 'c3d_upload_state')
end subroutine C3D_BEGIN

! GETUP copies the state back into its arguments every NEVERY steps (0: never)
subroutine C3D_SET_HOSTSYNC(NEVERY)
  use c3d_state
  implicit none
  integer, intent(in) :: NEVERY
  hostsync = NEVERY
end subroutine C3D_SET_HOSTSYNC

! changes CFLMAX of /TSTEP/ after C3D_BEGIN (an adaptive-DT run changes it)
subroutine C3D_SET_CFLMAX(CFLMAX)
  use c3d_state
  implicit none
  real(8), intent(in) :: CFLMAX
  integer(c_int) :: rc
  rc = c3d_set_scalar(h3d, 12_c_int, CFLMAX)
  call c3d_check(rc, h3d, 'c3d_set_scalar')
end subroutine C3D_SET_CFLMAX

subroutine C3D_SYNC(U, P, HH, SGSVIS, PRESG, PRESG3)
  use c3d_state
  implicit none
  real(8) :: U(*), P(*), HH(*), SGSVIS(*), PRESG, PRESG3
  integer(c_int) :: rc
  rc = c3d_download_state(h3d, U, P, HH, SGSVIS, PRESG, PRESG3)
  call c3d_check(rc, h3d, 'c3d_download_state')
end subroutine C3D_SYNC

subroutine C3D_END(U, P, HH, SGSVIS, PRESG, PRESG3)
  use c3d_state
  implicit none
  real(8) :: U(*), P(*), HH(*), SGSVIS(*), PRESG, PRESG3
  integer(c_int) :: rc
  call C3D_SYNC(U, P, HH, SGSVIS, PRESG, PRESG3)
  rc = c3d_destroy(h3d)
  h3d = c_null_ptr
end subroutine C3D_END

subroutine GETUP(U, P, HH, SS, NTIME, PRESG, PRESG3, SGSVIS)
  use c3d_state
  implicit none
  real(8) :: U(*), P(*), HH(*), SS(*), SGSVIS(*), PRESG, PRESG3
  integer :: NTIME
  integer(c_int) :: rc
  rc = c3d_getup(h3d, int(NTIME, c_int))
  call c3d_check(rc, h3d, 'GETUP')
  if (hostsync > 0) then
    if (mod(NTIME, hostsync) == 0) then
      rc = c3d_download_state(h3d, U, P, HH, SGSVIS, PRESG, PRESG3)
      call c3d_check(rc, h3d, 'GETUP (host sync)')
      return
    end if
  end if
  rc = c3d_get_scalar(h3d, 0_c_int, PRESG)
  call c3d_check(rc, h3d, 'GETUP (PRESG)')
  rc = c3d_get_scalar(h3d, 1_c_int, PRESG3)
  call c3d_check(rc, h3d, 'GETUP (PRESG3)')
end subroutine GETUP

subroutine DIVCHECK(U, TIME, DIVMAX, NTIME)
  use c3d_state
  implicit none
  real(8) :: U(*), TIME, DIVMAX
  integer :: NTIME
  integer(c_int) :: rc
  rc = c3d_divcheck(h3d, DIVMAX)
  call c3d_check(rc, h3d, 'DIVCHECK')
end subroutine DIVCHECK

subroutine CFL(U, CFLM)
  use c3d_state
  implicit none
  real(8) :: U(*), CFLM
  integer(c_int) :: rc
  rc = c3d_cfl(h3d, CFLM)
  call c3d_check(rc, h3d, 'CFL')
end subroutine CFL

subroutine CHKMF(U, TIME)
  use c3d_state
  implicit none
  real(8) :: U(*), TIME
  integer(c_int) :: rc
  rc = c3d_chkmf(h3d, flow1, flow3)
  call c3d_check(rc, h3d, 'CHKMF')
end subroutine CHKMF

subroutine WALLSS(U, P, TIME)
  use c3d_state
  implicit none
  real(8) :: U(*), P(*), TIME
  integer(c_int) :: rc
  rc = c3d_wallss(h3d)
  call c3d_check(rc, h3d, 'WALLSS')
end subroutine WALLSS

! RESULT/FLOW-*.plt and RESULT/STRAIN-*.plt of the resident field.  UTAU is
! formed as WRITEUP forms it (:785-789), from the caller's COMMON blocks
! WSMEM (WSMO) and WALLTAOAVE (TAOWO); the RESULT directory must exist, as
! for the reference's OPEN.
subroutine WRITEUP(U, P, PRESG, PRESG3, NTIME)
  use c3d_state
  implicit none
  real(8) :: U(*), P(*), PRESG, PRESG3
  integer :: NTIME
  real(8) :: WSM, WSMO, TAOWO, WMODEL, TAUW1, TAUB1, UTAU
  integer :: IFWALL, IS
  ! the reference's member lists and implicit types (compact3d.f:46, COMMON1.FI:31)
  common /WSMEM/ WSM(3), WSMO(3)
  common /WALLTAOAVE/ TAOWO
  common /WALLMODEL/ WMODEL, IFWALL, TAUW1, TAUB1, IS
  character(len=6) :: tag
  integer(c_int) :: rc
  write(tag, '(I6)') 100000 + NTIME
  if (WMODEL == 0.0d0) then
    UTAU = sqrt(WSMO(1))
  else
    UTAU = sqrt(TAOWO)
  end if
  rc = c3d_write_flow(h3d, 'RESULT/FLOW-'//tag//'.plt'//c_null_char, int(NTIME, c_int))
  call c3d_check(rc, h3d, 'WRITEUP')
  rc = c3d_write_strain(h3d, 'RESULT/STRAIN-'//tag//'.plt'//c_null_char, int(NTIME, c_int), UTAU)
  call c3d_check(rc, h3d, 'WRITEUP')
end subroutine WRITEUP
