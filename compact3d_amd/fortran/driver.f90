! driver.f90 -- the reference's time loop (MAIN, compact3d.f:83-172) on the
! device through the redirected subroutines.
!
! Usage: c3d_driver N1 N2 N3 RE ALX ALY ALZ DT FILES WMODEL CKS NTST NPRN CONTINUE [CFLMAX [NSYNC]]
! CONTINUE is a READUP file (compact3d.f:741-746: unformatted records
! PRESG,PRESG3 / U(K fastest, J=0..N2, I slowest) / P).  Fixed DT (IDTOPT=0).
! CFLMAX is /TSTEP/'s (default 0.5, parame.ter's usual value).  Every NPRN
! steps it prints MAIN's FORMAT 100 fields that this path computes.  With
! NSYNC > 0 GETUP copies the state into the host arrays every NSYNC steps and
! the driver prints U(5,5,5,1) of the host array then (as MAIN's host
! routines would read it).
program c3d_driver
  implicit none
  integer :: N1, N2, N3, NTST, NPRN, NTIME, I, J, K
  real(8) :: RE, ALX, ALY, ALZ, DT, FILES, WMODEL, CKS, PRESG, PRESG3, TIME, DIVMAX, CFLM, CFLMAX
  integer :: NSYNC
  real(8), allocatable :: U(:,:,:,:), P(:,:,:), HH(:,:,:,:), SS(:,:,:,:), SGSVIS(:,:,:)
  character(len=256) :: arg, fname
  integer(8) :: c0, c1, cr
  call getarg(1, arg); read(arg, *) N1
  call getarg(2, arg); read(arg, *) N2
  call getarg(3, arg); read(arg, *) N3
  call getarg(4, arg); read(arg, *) RE
  call getarg(5, arg); read(arg, *) ALX
  call getarg(6, arg); read(arg, *) ALY
  call getarg(7, arg); read(arg, *) ALZ
  call getarg(8, arg); read(arg, *) DT
  call getarg(9, arg); read(arg, *) FILES
  call getarg(10, arg); read(arg, *) WMODEL
  call getarg(11, arg); read(arg, *) CKS
  call getarg(12, arg); read(arg, *) NTST
  call getarg(13, arg); read(arg, *) NPRN
  call getarg(14, fname)
  CFLMAX = 0.5d0
  NSYNC = 0
  if (command_argument_count() >= 15) then
    call getarg(15, arg); read(arg, *) CFLMAX
  end if
  if (command_argument_count() >= 16) then
    call getarg(16, arg); read(arg, *) NSYNC
  end if
  allocate(U(0:N1,0:N2,0:N3,3), P(N1,N2,N3), HH(0:N1,0:N2,0:N3,3), SS(0:N1,0:N2,0:N3,3))
  allocate(SGSVIS(0:N1,0:N2,0:N3))
  U = 0; P = 0; HH = 0; SS = 0; SGSVIS = 0
  ! READUP (compact3d.f:741-746)
  open(3, file=trim(fname), form='UNFORMATTED', status='OLD')
  read(3) PRESG, PRESG3
  read(3) (((U(I,J,K,1), U(I,J,K,2), U(I,J,K,3), K=1,N3-1), J=0,N2), I=1,N1-1)
  read(3) (((P(I,J,K), K=1,N3-1), J=1,N2-1), I=1,N1-1)
  close(3)
  call C3D_BEGIN(N1, N2, N3, RE, ALX, ALY, ALZ, DT, FILES, WMODEL, CKS, CFLMAX, 0, U, P, HH, SGSVIS, PRESG, PRESG3)
  call C3D_SET_HOSTSYNC(NSYNC)
  TIME = 0.0d0
  NTIME = 1
  call DIVCHECK(U, TIME, DIVMAX, NTIME)
  call CHKMF(U, TIME)
  call system_clock(c0, cr)
  do while (NTIME <= NTST)
    call CFL(U, CFLM)
    TIME = TIME + DT
    if (WMODEL == 0.0d0) call WALLSS(U, P, TIME)
    call GETUP(U, P, HH, SS, NTIME, PRESG, PRESG3, SGSVIS)
    call DIVCHECK(U, TIME, DIVMAX, NTIME)
    call CHKMF(U, TIME)
    call CFL(U, CFLM)
    if (NSYNC > 0) then
      if (mod(NTIME, NSYNC) == 0) write(*, '(A,I6,2X,A,ES22.14)') 'HOSTSYNC=', NTIME, 'U(5,5,5,1)=', U(5,5,5,1)
    end if
    if (mod(NTIME, NPRN) == 0) then
      write(*, '(A,I6,2X,A,ES22.14,2X,A,ES12.5,2X,A,ES10.3,2X,A,ES12.5)') 'STEP=', NTIME, 'DP=', PRESG, &
          'DT=', DT, 'DIVMAX=', DIVMAX, 'CFL=', CFLM * DT
    end if
    NTIME = NTIME + 1
  end do
  call system_clock(c1)
  write(*, '(A,F12.6)') 'ms_per_step=', 1000.0d0 * dble(c1 - c0) / dble(cr) / dble(NTST)
  ! the final WRITEUP of MAIN (:177), into ./RESULT
  call WRITEUP(U, P, PRESG, PRESG3, NTIME)
  call C3D_END(U, P, HH, SGSVIS, PRESG, PRESG3)
  write(*, '(A,ES22.14)') 'U(5,5,5,1)=', U(5,5,5,1)
end program c3d_driver
