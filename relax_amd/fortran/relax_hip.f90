!-----------------------------------------------------------------------
! relax_hip.f90 -- Fortran drop-in for RELAX's per-time-step routines on an
! MI355X through the C ABI of librelax_hip.so (include/relax_hip.h).
!
! The subroutines keep the names, argument order, shapes and optional
! arguments of the reference's routines (paths relative to its src/):
!   fft3, fft2                         fourier.f90:177-355, :372-535
!   elasticresponse                    green.f90:51-101
!   surfacetractioncowling             green.f90:327-377
!   cerrutimodified                    green.f90:520-612
!   greenfunctioncowling               green.f90:1008-1129
!   stressupdate, equivalentbodyforce  elastic3d.f90:956-1506
!   fieldadd, tensorfieldadd           elastic3d.f90:3774-3959
!   tensoramplitude                    elastic3d.f90:446-466
!   viscouseigenstress                 viscoelastic3d.f90:217-314
!   frictioneigenstress                friction3d.f90:82-241
!   monitorfriction, frictionadd       friction3d.f90:263-453
! A STOP of the reference becomes a non-zero relax_hip_status and a message
! on unit 0; the routines never stop the program themselves.
!
! The derived types are SEQUENCE types with the reference's names and
! components (types.f90), so they are the same types as the driver's.  Built
! inside the reference tree, define RELAX_REFERENCE_TYPES to take them from
! its module types instead of relax_hip_types.
!-----------------------------------------------------------------------
MODULE relax_hip_types
  IMPLICIT NONE
  TYPE LAYER_STRUCT
     SEQUENCE
     REAL*8 :: z,gammadot0,stressexponent,cohesion,friction,Gk
  END TYPE LAYER_STRUCT
  TYPE TENSOR
     SEQUENCE
     REAL*4 :: s11,s12,s13,s22,s23,s33
  END TYPE TENSOR
  TYPE TENSOR_LAYER_STRUCT
     SEQUENCE
     REAL*4 :: z,dum
     TYPE(TENSOR) :: t
  END TYPE TENSOR_LAYER_STRUCT
  TYPE SLIPPATCH_STRUCT
     SEQUENCE
     REAL*8 :: x1,x2,x3
     REAL*8 :: lx,lz
     REAL*8 :: slip,ss,ds
     REAL*8 :: v,vss,vds
     REAL*8 :: taus
     TYPE(TENSOR) :: sig
  END TYPE SLIPPATCH_STRUCT
  TYPE PLANE_STRUCT
     SEQUENCE
     REAL*8 :: x,y,z,width,length,strike,dip,rake
     INTEGER :: px2,px3
     TYPE(SLIPPATCH_STRUCT), DIMENSION(:,:), ALLOCATABLE :: patch
  END TYPE PLANE_STRUCT
END MODULE relax_hip_types

MODULE relax_hip
  USE, INTRINSIC :: iso_c_binding
#ifdef RELAX_REFERENCE_TYPES
  USE types
#else
  USE relax_hip_types
#endif
  IMPLICIT NONE
  PRIVATE :: check, c_interfaces_dummy

  !> space-domain layout of the real fields: 1 = padded rows of sx1+2 (FFTW,
  !! MKL builds), 0 = packed rows of sx1 (ctfft build)
  INTEGER, PUBLIC :: relax_hip_align = 1
  !> status of the last call (0 = success, RELAX_ERR_* of relax_hip.h)
  INTEGER, PUBLIC :: relax_hip_status = 0
  INTEGER, PARAMETER, PUBLIC :: FFT_FORWARD=-1, FFT_INVERSE=1

  INTERFACE
     INTEGER(c_int) FUNCTION c_fft3(d,sx1,sx2,sx3,dx1,dx2,dx3,dir,al) BIND(C,name="relax_fft3")
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: d
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,dir,al
       REAL(c_double), VALUE :: dx1,dx2,dx3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_fft2(d,sx1,sx2,dx1,dx2,dir,al) BIND(C,name="relax_fft2")
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: d
       INTEGER(c_int), VALUE :: sx1,sx2,dir,al
       REAL(c_double), VALUE :: dx1,dx2
     END FUNCTION
     INTEGER(c_int) FUNCTION c_elasticresponse(l,m,f1,f2,f3,sx1,sx2,sx3,dx1,dx2,dx3) &
          BIND(C,name="relax_elasticresponse")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: l,m,dx1,dx2,dx3
       TYPE(c_ptr), VALUE :: f1,f2,f3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_surfacetraction(l,m,g,u1,u2,u3,sx1,sx2,sx3,dx1,dx2,dx3,p1,p2,p3) &
          BIND(C,name="relax_surfacetractioncowling")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: l,m,g,dx1,dx2,dx3
       TYPE(c_ptr), VALUE :: u1,u2,u3,p1,p2,p3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_cerruti(p1,p2,p3,l,m,g,u1,u2,u3,sx1,sx2,sx3,dx1,dx2,dx3) &
          BIND(C,name="relax_cerrutimodified")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: l,m,g,dx1,dx2,dx3
       TYPE(c_ptr), VALUE :: u1,u2,u3,p1,p2,p3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_green(c1,c2,c3,t1,t2,t3,sx1,sx2,sx3,dx1,dx2,dx3,l,m,g,al) &
          BIND(C,name="relax_greenfunctioncowling")
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: c1,c2,c3,t1,t2,t3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,al
       REAL(c_double), VALUE :: dx1,dx2,dx3,l,m,g
     END FUNCTION
     INTEGER(c_int) FUNCTION c_stressupdate(v1,v2,v3,l,m,dx1,dx2,dx3,sx1,sx2,sx3,sig,al) &
          BIND(C,name="relax_stressupdate")
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: v1,v2,v3,sig
       REAL(c_double), VALUE :: l,m,dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,al
     END FUNCTION
     INTEGER(c_int) FUNCTION c_bodyforce(sig,dx1,dx2,dx3,sx1,sx2,sx3,c1,c2,c3,t1,t2,t3,mask,al) &
          BIND(C,name="relax_equivalentbodyforce")
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: sig,c1,c2,c3,t1,t2,t3,mask
       REAL(c_double), VALUE :: dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,al
     END FUNCTION
     INTEGER(c_int) FUNCTION c_fieldadd(u,v,n,c1,c2) BIND(C,name="relax_fieldadd")
       IMPORT :: c_int, c_long, c_ptr
       TYPE(c_ptr), VALUE :: u,v,c1,c2
       INTEGER(c_long), VALUE :: n
     END FUNCTION
     INTEGER(c_int) FUNCTION c_tensorfieldadd(t1,t2,n,c1,c2) BIND(C,name="relax_tensorfieldadd")
       IMPORT :: c_int, c_long, c_ptr
       TYPE(c_ptr), VALUE :: t1,t2,c1,c2
       INTEGER(c_long), VALUE :: n
     END FUNCTION
     INTEGER(c_int) FUNCTION c_tensoramplitude(t,n,dx1,dx2,dx3,out) BIND(C,name="relax_tensoramplitude")
       IMPORT :: c_int, c_long, c_double, c_ptr
       TYPE(c_ptr), VALUE :: t,out
       INTEGER(c_long), VALUE :: n
       REAL(c_double), VALUE :: dx1,dx2,dx3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_viscous(mu,st,nst,pre,npre,sig,sx1,sx2,sx3,dx1,dx2,dx3,mom,mw,dg,gam,al) &
          BIND(C,name="relax_viscouseigenstress")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: mu,dx1,dx2,dx3
       TYPE(c_ptr), VALUE :: st,pre,sig,mom,mw,dg,gam
       INTEGER(c_int), VALUE :: nst,npre,sx1,sx2,sx3,al
     END FUNCTION
     INTEGER(c_int) FUNCTION c_friction(x,y,z,L,W,strike,dip,rake,beta,sig,pre,mu,st,sx1,sx2,sx3, &
          dx1,dx2,dx3,mom,mw,vel,al) BIND(C,name="relax_frictioneigenstress")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: x,y,z,L,W,strike,dip,rake,beta,mu,dx1,dx2,dx3
       TYPE(c_ptr), VALUE :: sig,pre,st,mom,mw,vel
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,al
     END FUNCTION
     INTEGER(c_int) FUNCTION c_monitor(x,y,z,L,W,strike,dip,rake,beta,sx1,sx2,sx3,sx3a,dx1,dx2,dx3, &
          sig,pre,st,patch,px2,px3) BIND(C,name="relax_monitorfriction")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: x,y,z,L,W,strike,dip,rake,beta,dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,sx3a,px2,px3
       TYPE(c_ptr), VALUE :: sig,pre,st,patch
     END FUNCTION
     INTEGER(c_int) FUNCTION c_frictionadd(patch,n,dt) BIND(C,name="relax_frictionadd")
       IMPORT :: c_int, c_long, c_double, c_ptr
       TYPE(c_ptr), VALUE :: patch
       INTEGER(c_long), VALUE :: n
       REAL(c_double), VALUE :: dt
     END FUNCTION
     FUNCTION c_status_string(s) BIND(C,name="relax_hip_status_string")
       IMPORT :: c_int, c_ptr
       INTEGER(c_int), VALUE :: s
       TYPE(c_ptr) :: c_status_string
     END FUNCTION
  END INTERFACE

CONTAINS

  SUBROUTINE c_interfaces_dummy()
  END SUBROUTINE c_interfaces_dummy

  SUBROUTINE check(rc,what)
    INTEGER(c_int), INTENT(IN) :: rc
    CHARACTER(*), INTENT(IN) :: what
    relax_hip_status=rc
    IF (0 .NE. rc) THEN
       WRITE (0,'("relax_hip: ",a," failed with status ",I3)') what, rc
    END IF
  END SUBROUTINE check

  SUBROUTINE fft3(data,sx1,sx2,sx3,dx1,dx2,dx3,direction)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3,direction
    REAL*4, DIMENSION(:,:,:), INTENT(INOUT), TARGET, CONTIGUOUS :: data
    REAL*8, INTENT(IN) :: dx1,dx2,dx3
    CALL check(c_fft3(C_LOC(data),sx1,sx2,sx3,dx1,dx2,dx3,direction,relax_hip_align),"fft3")
  END SUBROUTINE fft3

  SUBROUTINE fft2(data,sx1,sx2,dx1,dx2,direction)
    INTEGER, INTENT(IN) :: sx1,sx2,direction
    REAL*4, DIMENSION(:,:), INTENT(INOUT), TARGET, CONTIGUOUS :: data
    REAL*8, INTENT(IN) :: dx1,dx2
    CALL check(c_fft2(C_LOC(data),sx1,sx2,dx1,dx2,direction,relax_hip_align),"fft2")
  END SUBROUTINE fft2

  SUBROUTINE elasticresponse(lambda,mu,f1,f2,f3,dx1,dx2,dx3)
    REAL*8, INTENT(IN) :: lambda,mu,dx1,dx2,dx3
    REAL*4, DIMENSION(:,:,:), INTENT(INOUT), TARGET, CONTIGUOUS :: f1,f2,f3
    CALL check(c_elasticresponse(lambda,mu,C_LOC(f1),C_LOC(f2),C_LOC(f3), &
         SIZE(f1,1)-2,SIZE(f1,2),SIZE(f1,3),dx1,dx2,dx3),"elasticresponse")
  END SUBROUTINE elasticresponse

  SUBROUTINE surfacetractioncowling(lambda,mu,gamma,u1,u2,u3,dx1,dx2,dx3,p1,p2,p3)
    REAL*4, INTENT(IN), DIMENSION(:,:,:), TARGET, CONTIGUOUS :: u1,u2,u3
    REAL*8, INTENT(IN) :: lambda,mu,gamma,dx1,dx2,dx3
    REAL*4, INTENT(OUT), DIMENSION(:,:), TARGET, CONTIGUOUS :: p1,p2,p3
    CALL check(c_surfacetraction(lambda,mu,gamma,C_LOC(u1),C_LOC(u2),C_LOC(u3), &
         SIZE(u1,1)-2,SIZE(u1,2),SIZE(u1,3),dx1,dx2,dx3,C_LOC(p1),C_LOC(p2),C_LOC(p3)),"surfacetractioncowling")
  END SUBROUTINE surfacetractioncowling

  SUBROUTINE cerrutimodified(p1,p2,p3,lambda,mu,gamma,u1,u2,u3,dx1,dx2,dx3)
    REAL*4, DIMENSION(:,:), INTENT(IN), TARGET, CONTIGUOUS :: p1,p2,p3
    REAL*4, DIMENSION(:,:,:), INTENT(INOUT), TARGET, CONTIGUOUS :: u1,u2,u3
    REAL*8, INTENT(IN) :: lambda,mu,gamma,dx1,dx2,dx3
    CALL check(c_cerruti(C_LOC(p1),C_LOC(p2),C_LOC(p3),lambda,mu,gamma,C_LOC(u1),C_LOC(u2),C_LOC(u3), &
         SIZE(u1,1)-2,SIZE(u1,2),SIZE(u1,3),dx1,dx2,dx3),"cerrutimodified")
  END SUBROUTINE cerrutimodified

  SUBROUTINE greenfunctioncowling(c1,c2,c3,t1,t2,t3,dx1,dx2,dx3,lambda,mu,gamma)
    REAL*4, INTENT(INOUT), DIMENSION(:,:,:), TARGET, CONTIGUOUS :: c1,c2,c3
    REAL*4, INTENT(INOUT), DIMENSION(:,:), TARGET, CONTIGUOUS :: t1,t2,t3
    REAL*8, INTENT(IN) :: dx1,dx2,dx3,lambda,mu,gamma
    CALL check(c_green(C_LOC(c1),C_LOC(c2),C_LOC(c3),C_LOC(t1),C_LOC(t2),C_LOC(t3), &
         SIZE(c1,1)-2,SIZE(c1,2),SIZE(c1,3),dx1,dx2,dx3,lambda,mu,gamma,relax_hip_align),"greenfunctioncowling")
  END SUBROUTINE greenfunctioncowling

  SUBROUTINE stressupdate(v1,v2,v3,lambda,mu,dx1,dx2,dx3,sx1,sx2,sx3,sig)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: dx1,dx2,dx3,lambda,mu
    REAL*4, INTENT(IN), DIMENSION(*), TARGET :: v1,v2,v3
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: sig
    CALL check(c_stressupdate(C_LOC(v1),C_LOC(v2),C_LOC(v3),lambda,mu,dx1,dx2,dx3,sx1,sx2,sx3, &
         C_LOC(sig),relax_hip_align),"stressupdate")
  END SUBROUTINE stressupdate

  SUBROUTINE equivalentbodyforce(sig,dx1,dx2,dx3,sx1,sx2,sx3,c1,c2,c3,t1,t2,t3,mask)
    REAL*8, INTENT(IN) :: dx1,dx2,dx3
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*4, INTENT(INOUT), DIMENSION(*), TARGET :: c1,c2,c3,t1,t2,t3
    TYPE(TENSOR), INTENT(IN), DIMENSION(sx1,sx2,sx3), TARGET :: sig
    REAL*4, INTENT(IN), DIMENSION(sx3), OPTIONAL, TARGET :: mask
    TYPE(c_ptr) :: pm
    pm=C_NULL_PTR
    IF (PRESENT(mask)) pm=C_LOC(mask)
    CALL check(c_bodyforce(C_LOC(sig),dx1,dx2,dx3,sx1,sx2,sx3,C_LOC(c1),C_LOC(c2),C_LOC(c3), &
         C_LOC(t1),C_LOC(t2),C_LOC(t3),pm,relax_hip_align),"equivalentbodyforce")
  END SUBROUTINE equivalentbodyforce

  SUBROUTINE fieldadd(u,v,sx1,sx2,sx3,c1,c2)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*4, INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: u
    REAL*4, INTENT(IN), DIMENSION(sx1,sx2,sx3), TARGET :: v
    REAL*4, INTENT(IN), OPTIONAL, TARGET :: c1,c2
    TYPE(c_ptr) :: p1,p2
    p1=C_NULL_PTR; p2=C_NULL_PTR
    IF (PRESENT(c1)) p1=C_LOC(c1)
    IF (PRESENT(c2)) p2=C_LOC(c2)
    CALL check(c_fieldadd(C_LOC(u),C_LOC(v),INT(sx1,c_long)*sx2*sx3,p1,p2),"fieldadd")
  END SUBROUTINE fieldadd

  SUBROUTINE tensorfieldadd(t1,t2,sx1,sx2,sx3,c1,c2)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: t1
    TYPE(TENSOR), INTENT(IN), DIMENSION(sx1,sx2,sx3), TARGET :: t2
    REAL*4, INTENT(IN), OPTIONAL, TARGET :: c1,c2
    TYPE(c_ptr) :: p1,p2
    p1=C_NULL_PTR; p2=C_NULL_PTR
    IF (PRESENT(c1)) p1=C_LOC(c1)
    IF (PRESENT(c2)) p2=C_LOC(c2)
    CALL check(c_tensorfieldadd(C_LOC(t1),C_LOC(t2),INT(sx1,c_long)*sx2*sx3,p1,p2),"tensorfieldadd")
  END SUBROUTINE tensorfieldadd

  REAL*8 FUNCTION tensoramplitude(t,dx1,dx2,dx3)
    TYPE(TENSOR), INTENT(IN), DIMENSION(:,:,:), TARGET, CONTIGUOUS :: t
    REAL*8, INTENT(IN) :: dx1,dx2,dx3
    REAL(c_double), TARGET :: out
    out=0
    CALL check(c_tensoramplitude(C_LOC(t),INT(SIZE(t),c_long),dx1,dx2,dx3,C_LOC(out)),"tensoramplitude")
    tensoramplitude=out
  END FUNCTION tensoramplitude

  SUBROUTINE viscouseigenstress(mu,structure,sig,prestress,sx1,sx2,sx3, &
       dx1,dx2,dx3,moment,maxwelltime,dgammadot0,gamma)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*4, DIMENSION(sx1,sx2,sx3), INTENT(IN), OPTIONAL, TARGET :: dgammadot0
    REAL*8, INTENT(IN) :: mu,dx1,dx2,dx3
    TYPE(LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: structure
    TYPE(TENSOR_LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: prestress
    TYPE(TENSOR), INTENT(IN), DIMENSION(sx1,sx2,sx3), TARGET :: sig
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: moment
    REAL*8, OPTIONAL, INTENT(INOUT), TARGET :: maxwelltime
    REAL*4, DIMENSION(*), INTENT(INOUT), OPTIONAL, TARGET :: gamma
    TYPE(c_ptr) :: pm,pd,pg
    pm=C_NULL_PTR; pd=C_NULL_PTR; pg=C_NULL_PTR
    IF (PRESENT(maxwelltime)) pm=C_LOC(maxwelltime)
    IF (PRESENT(dgammadot0)) pd=C_LOC(dgammadot0)
    IF (PRESENT(gamma)) pg=C_LOC(gamma)
    CALL check(c_viscous(mu,C_LOC(structure),SIZE(structure),C_LOC(prestress),SIZE(prestress),C_LOC(sig), &
         sx1,sx2,sx3,dx1,dx2,dx3,C_LOC(moment),pm,pd,pg,relax_hip_align),"viscouseigenstress")
  END SUBROUTINE viscouseigenstress

  SUBROUTINE frictioneigenstress(x,y,z,L,W,strike,dip,rake,beta, &
       sig,prestress,mu,structure,sx1,sx2,sx3,dx1,dx2,dx3,moment,maxwelltime,vel)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: mu,dx1,dx2,dx3,x,y,z,L,W,strike,dip,rake,beta
    TYPE(LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: structure
    TYPE(TENSOR_LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: prestress
    TYPE(TENSOR), INTENT(IN), DIMENSION(sx1,sx2,sx3), TARGET :: sig
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: moment
    REAL*8, OPTIONAL, INTENT(INOUT), TARGET :: maxwelltime
    REAL*4, INTENT(INOUT), DIMENSION(*), OPTIONAL, TARGET :: vel
    TYPE(c_ptr) :: pm,pv
    pm=C_NULL_PTR; pv=C_NULL_PTR
    IF (PRESENT(maxwelltime)) pm=C_LOC(maxwelltime)
    IF (PRESENT(vel)) pv=C_LOC(vel)
    CALL check(c_friction(x,y,z,L,W,strike,dip,rake,beta,C_LOC(sig),C_LOC(prestress),mu,C_LOC(structure), &
         sx1,sx2,sx3,dx1,dx2,dx3,C_LOC(moment),pm,pv,relax_hip_align),"frictioneigenstress")
  END SUBROUTINE frictioneigenstress

  SUBROUTINE monitorfriction(x,y,z,L,W,strike,dip,rake,beta, &
       sx1,sx2,sx3,dx1,dx2,dx3,sig,prestress,structure,patch)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: x,y,z,L,W,strike,rake,dip,beta,dx1,dx2,dx3
    TYPE(TENSOR), DIMENSION(:,:,:), INTENT(IN), TARGET, CONTIGUOUS :: sig
    TYPE(SLIPPATCH_STRUCT), DIMENSION(:,:), INTENT(INOUT), TARGET, CONTIGUOUS :: patch
    TYPE(LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: structure
    TYPE(TENSOR_LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: prestress
    ! the driver passes the full grid depth (relax.f90:758); the stress field
    ! holds SIZE(sig,3) planes
    CALL check(c_monitor(x,y,z,L,W,strike,dip,rake,beta,sx1,sx2,sx3,SIZE(sig,3),dx1,dx2,dx3, &
         C_LOC(sig),C_LOC(prestress),C_LOC(structure),C_LOC(patch),SIZE(patch,1),SIZE(patch,2)), &
         "monitorfriction")
  END SUBROUTINE monitorfriction

  SUBROUTINE frictionadd(np,n,dt)
    INTEGER, INTENT(IN) :: np
    TYPE(PLANE_STRUCT), INTENT(INOUT), DIMENSION(np), TARGET :: n
    REAL*8, INTENT(IN) :: dt
    INTEGER :: k
    DO k=1,np
       CALL check(c_frictionadd(C_LOC(n(k)%patch),INT(SIZE(n(k)%patch),c_long),dt),"frictionadd")
    END DO
  END SUBROUTINE frictionadd

END MODULE relax_hip
