! Fortran driver of the batch test: three models set up as model_loop.f90
! sets up one (structure, prestress, the first event), two on a 32^3 grid with
! different viscosities and one on 32x32x64, stepped together as one batch
! through MODULE relax_hip; writes the step scalars and u1 of each model.
PROGRAM batch_loop
  USE relax_hip
  IMPLICIT NONE
  INTEGER, PARAMETER :: sx1=32, sx2=32, nsteps=3, nm=3
  REAL*8, PARAMETER :: dx=1.d0, lam=3.d4, mu=3.d4, gam=8.33d-4
  INTEGER, PARAMETER :: depth(nm)=(/32,32,64/)
  REAL*8, PARAMETER :: g0(nm)=(/0.2d0,0.05d0,0.1d0/)
  TYPE(relax_model) :: m(nm)
  TYPE(relax_batch) :: b
  TYPE(relax_step_info) :: info(nm,nsteps)
  REAL*4, ALLOCATABLE :: u1(:,:,:),u2(:,:,:),u3(:,:,:)
  CHARACTER(512) :: fname
  INTEGER :: j,k

  CALL GET_COMMAND_ARGUMENT(1,fname)
  DO j=1,nm
     CALL setup(m(j),depth(j),g0(j))
  END DO
  CALL relax_hip_batch_create(b)
  DO j=1,nm
     CALL relax_hip_batch_add(b,m(j))
  END DO
  DO k=1,nsteps
     CALL relax_hip_batch_step(b,info(:,k))
     IF (relax_hip_status .NE. 0) STOP 1
  END DO
  CALL relax_hip_batch_destroy(b)
  OPEN(10,FILE=TRIM(fname),ACCESS='STREAM',FORM='UNFORMATTED',STATUS='REPLACE')
  WRITE(10) ((info(j,k)%Dt,info(j,k)%t,info(j,k)%power,info(j,k)%cei,j=1,nm),k=1,nsteps)
  DO j=1,nm
     ALLOCATE(u1(sx1+2,sx2,depth(j)/2),u2(sx1+2,sx2,depth(j)/2),u3(sx1+2,sx2,depth(j)/2))
     CALL relax_hip_model_download(m(j),u1,u2,u3)
     WRITE(10) u1
     DEALLOCATE(u1,u2,u3)
     CALL relax_hip_model_destroy(m(j))
  END DO
  CLOSE(10)
  PRINT '("batch_loop: ok")'
CONTAINS
  SUBROUTINE setup(mo,sx3,gd)
    TYPE(relax_model), INTENT(OUT) :: mo
    INTEGER, INTENT(IN) :: sx3
    REAL*8, INTENT(IN) :: gd
    TYPE(relax_model_desc) :: d
    TYPE(LAYER_STRUCT), ALLOCATABLE, TARGET :: lin(:)
    TYPE(TENSOR_LAYER_STRUCT), ALLOCATABLE, TARGET :: pre(:)
    TYPE(SOURCE_STRUCT), TARGET :: pat(1), load(1)
    INTEGER :: i
    d=relax_model_desc(sx1,sx2,sx3,dx,dx,dx,lam,mu,gam,0.2d0,1.d9,1.d0,1.d0,1,0,C_NULL_PTR)
    CALL relax_hip_model_create(d,mo)
    ALLOCATE(lin(sx3/2),pre(sx3/2))
    DO i=1,sx3/2
       ! a Maxwell layer below 6 km (gammadot0 = gd /yr, n = 1)
       lin(i)=LAYER_STRUCT(DBLE(i-1)*dx,MERGE(gd,0.d0,i>6),1.d0,0.d0,0.6d0,0.d0)
       pre(i)=TENSOR_LAYER_STRUCT(REAL(i-1,4),0.,TENSOR(0.,0.,0.,0.,0.,0.))
    END DO
    CALL relax_hip_model_set_prestress(mo,pre)
    CALL relax_hip_model_set_structure(mo,RELAX_LINEAR,lin)
    pat(1)=SOURCE_STRUCT(1.d-3,0.d0,0.d0,1.d0,6.d0,10.d0,0.d0,0.d0,0.d0,0.d0,0.d0,0.2d0,0.d0)
    load(1)=SOURCE_STRUCT(0.5d0,3.d0,-2.d0,0.d0,3.d0,4.d0,0.d0,0.d0,0.d0,0.d0,0.d0,0.2d0,0.d0)
    CALL relax_hip_model_event(mo,pat,load,1.d0)
  END SUBROUTINE setup
END PROGRAM batch_loop
