! Fortran driver of the drop-in test: calls the reference's routine names
! through MODULE relax_hip (relax_amd/fortran/relax_hip.f90) exactly as
! relax.f90 does, and writes inputs and outputs to a stream file.
PROGRAM drop_in
  USE relax_hip
  IMPLICIT NONE
  INTEGER, PARAMETER :: sx1=32, sx2=32, sx3=32
  REAL*8, PARAMETER :: dx=1.0d0, lam=3.d4, mu=3.d4, gam=8.33d-4
  REAL*8, PARAMETER :: pi2=6.28318530717958623199592693708837032318d0
  REAL*4, ALLOCATABLE :: c1(:,:,:),c2(:,:,:),c3(:,:,:),t1(:,:),t2(:,:),t3(:,:)
  TYPE(TENSOR), ALLOCATABLE :: sig(:,:,:)
  CHARACTER(512) :: fname
  INTEGER :: i1,i2,i3
  REAL*8 :: g

  CALL GET_COMMAND_ARGUMENT(1,fname)
  ALLOCATE(c1(sx1+2,sx2,sx3),c2(sx1+2,sx2,sx3),c3(sx1+2,sx2,sx3), &
           t1(sx1+2,sx2),t2(sx1+2,sx2),t3(sx1+2,sx2),sig(sx1,sx2,sx3/2))
  c1=0; c2=0; c3=0; t1=0; t2=0; t3=0
  DO i3=1,sx3/2
     g=EXP(-((i3-8.d0)/3.d0)**2)
     DO i2=1,sx2
        DO i1=1,sx1
           c1(i1,i2,i3)=REAL(SIN(pi2*i1/sx1)*COS(pi2*2*i2/sx2)*g,4)
           c2(i1,i2,i3)=REAL(COS(pi2*3*i1/sx1)*SIN(pi2*i2/sx2)*g,4)
           c3(i1,i2,i3)=REAL(SIN(pi2*(i1+i2)/sx1)*g,4)
        END DO
     END DO
  END DO
  DO i2=1,sx2
     DO i1=1,sx1
        t3(i1,i2)=REAL(0.1d0*COS(pi2*i1/sx1)*COS(pi2*i2/sx2),4)
     END DO
  END DO
  OPEN(10,FILE=TRIM(fname),ACCESS='STREAM',FORM='UNFORMATTED',STATUS='REPLACE')
  WRITE(10) c1,c2,c3,t1,t2,t3
  CALL greenfunctioncowling(c1,c2,c3,t1,t2,t3,dx,dx,dx,lam,mu,gam)
  IF (relax_hip_status .NE. 0) STOP 1
  sig=TENSOR(0,0,0,0,0,0)
  CALL stressupdate(c1,c2,c3,lam,mu,dx,dx,dx,sx1,sx2,sx3/2,sig)
  IF (relax_hip_status .NE. 0) STOP 2
  WRITE(10) c1,c2,c3,sig
  CLOSE(10)
  PRINT '("drop_in: ok")'
END PROGRAM drop_in
