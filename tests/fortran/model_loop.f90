! Fortran driver of the model-loop test: sets up a model through the
! device-resident part of MODULE relax_hip as relax.f90 would (structure,
! prestress, the first event: one shear patch and one surface load), runs RK2
! steps, and writes the step scalars and the fields to a stream file.
PROGRAM model_loop
  USE relax_hip
  IMPLICIT NONE
  INTEGER, PARAMETER :: sx1=32, sx2=32, sx3=32, h3=sx3/2, nsteps=3
  REAL*8, PARAMETER :: dx=1.d0, lam=3.d4, mu=3.d4, gam=8.33d-4
  REAL*8, PARAMETER :: pi=3.14159265358979323846d0
  TYPE(relax_model_desc) :: d
  TYPE(relax_model) :: m
  TYPE(LAYER_STRUCT), TARGET :: lin(h3)
  TYPE(TENSOR_LAYER_STRUCT), TARGET :: pre(h3)
  TYPE(SOURCE_STRUCT), TARGET :: pat(1), load(1)
  TYPE(relax_step_info) :: info(nsteps)
  REAL*4, ALLOCATABLE :: u1(:,:,:),u2(:,:,:),u3(:,:,:)
  TYPE(TENSOR), ALLOCATABLE :: tau(:,:,:)
  CHARACTER(512) :: fname
  INTEGER :: k

  CALL GET_COMMAND_ARGUMENT(1,fname)
  d=relax_model_desc(sx1,sx2,sx3,dx,dx,dx,lam,mu,gam,0.2d0,1.d9,1.d0,1.d0,1,0,C_NULL_PTR)
  CALL relax_hip_model_create(d,m)
  DO k=1,h3
     ! a Maxwell layer below 6 km (gammadot0 = 0.2 /yr, n = 1)
     lin(k)=LAYER_STRUCT(DBLE(k-1)*dx,MERGE(0.2d0,0.d0,k>6),1.d0,0.d0,0.6d0,0.d0)
     pre(k)=TENSOR_LAYER_STRUCT(REAL(k-1,4),0.,TENSOR(0.,0.,0.,0.,0.,0.))
  END DO
  CALL relax_hip_model_set_prestress(m,pre)
  CALL relax_hip_model_set_structure(m,RELAX_LINEAR,lin)
  ! one vertical strike-slip patch, 10 km long and 6 km wide, top at 1 km
  ! (internal convention: dip measured from the vertical, radians), and a
  ! 4 x 3 km surface load of 0.5 MPa
  pat(1)=SOURCE_STRUCT(1.d-3,0.d0,0.d0,1.d0,6.d0,10.d0,0.d0,0.d0,0.d0,0.d0,0.d0,0.2d0,0.d0)
  load(1)=SOURCE_STRUCT(0.5d0,3.d0,-2.d0,0.d0,3.d0,4.d0,0.d0,0.d0,0.d0,0.d0,0.d0,0.2d0,0.d0)
  CALL relax_hip_model_event(m,pat,load,1.d0)
  DO k=1,nsteps
     CALL relax_hip_model_step(m,info(k))
     IF (relax_hip_status .NE. 0) STOP 1
  END DO
  ALLOCATE(u1(sx1+2,sx2,h3),u2(sx1+2,sx2,h3),u3(sx1+2,sx2,h3),tau(sx1,sx2,h3))
  CALL relax_hip_model_download(m,u1,u2,u3,tau=tau)
  OPEN(10,FILE=TRIM(fname),ACCESS='STREAM',FORM='UNFORMATTED',STATUS='REPLACE')
  WRITE(10) (info(k)%Dt,info(k)%t,info(k)%power,info(k)%cei,k=1,nsteps)
  WRITE(10) u1,u2,u3,tau
  CLOSE(10)
  CALL relax_hip_model_destroy(m)
  PRINT '("model_loop: ok")'
END PROGRAM model_loop
