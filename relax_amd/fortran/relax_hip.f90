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
!   transienteigenstress               viscoelastic3d.f90:123-202
!   builddgammadot0                    viscoelastic3d.f90:320-435
!   source, traction                   elastic3d.f90:2064-2252, :2517-2590
!   tensilesource, mogisource          elastic3d.f90:2315-2411, :2458-2509
!   momentdensitytensile/-mogi         elastic3d.f90:3137-3215, :3240-3286
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
  TYPE SOURCE_STRUCT
     SEQUENCE
     REAL*8 :: slip,x,y,z,width,length,strike,dip,rake,period,phase,beta,opening
  END TYPE SOURCE_STRUCT
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
  TYPE WEAK_STRUCT
     SEQUENCE
     REAL*8 :: dgammadot0,x,y,z,width,length,thickness,strike,dip
     TYPE(TENSOR) :: e
  END TYPE WEAK_STRUCT
  TYPE TETRAHEDRON_STRUCT
     SEQUENCE
     TYPE(TENSOR) :: e
     INTEGER :: i1,i2,i3,i4
  END TYPE TETRAHEDRON_STRUCT
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

  !> device-resident model: the body of the time loop of relax.f90:555-890
  !! (relax_model_* of relax_hip.h); the layouts are those of the C structs
  TYPE, BIND(C), PUBLIC :: relax_model_desc
     INTEGER(c_int) :: sx1,sx2,sx3
     REAL(c_double) :: dx1,dx2,dx3,lambda,mu,gamma,beta,interval,odt,tscale
     INTEGER(c_int) :: skip,ne
     TYPE(c_ptr) :: event_times=C_NULL_PTR
  END TYPE relax_model_desc
  TYPE, BIND(C), PUBLIC :: relax_step_info
     INTEGER(c_int) :: i
     REAL(c_double) :: Dt,tm,maxwell(5),power,cei,t
     INTEGER(c_int) :: oi,status,e
  END TYPE relax_step_info
  !> a fault plane or slip patch in the internal convention (angles in
  !! radians): PLANE_STRUCT without its patch array
  TYPE, BIND(C), PUBLIC :: relax_plane_c
     REAL(c_double) :: x,y,z,width,length,strike,dip,rake
     INTEGER(c_int) :: px2,px3
  END TYPE relax_plane_c
  TYPE, PUBLIC :: relax_model
     TYPE(c_ptr) :: h=C_NULL_PTR
  END TYPE relax_model
  !> models stepped together (relax_batch_t): a parameter search's models,
  !! of one grid or of several
  TYPE, PUBLIC :: relax_batch
     TYPE(c_ptr) :: h=C_NULL_PTR
  END TYPE relax_batch
  !> the source lists of one event (relax_event_t)
  TYPE, BIND(C), PUBLIC :: relax_event_c
     TYPE(c_ptr) :: s=C_NULL_PTR
     INTEGER(c_int) :: ns=0
     TYPE(c_ptr) :: ts=C_NULL_PTR
     INTEGER(c_int) :: nt=0
     TYPE(c_ptr) :: m=C_NULL_PTR
     INTEGER(c_int) :: nm=0
     TYPE(c_ptr) :: l=C_NULL_PTR
     INTEGER(c_int) :: nl=0
     TYPE(c_ptr) :: c=C_NULL_PTR
     INTEGER(c_int) :: nc=0
     TYPE(c_ptr) :: th=C_NULL_PTR
     INTEGER(c_int) :: nth=0
     TYPE(c_ptr) :: v=C_NULL_PTR
     INTEGER(c_int) :: nv=0
  END TYPE relax_event_c
  INTEGER, PARAMETER, PUBLIC :: RELAX_LINEAR=0, RELAX_POWERLAW=1, RELAX_FAULTCREEP=2, &
       RELAX_LTRANSIENT=3, RELAX_NLTRANSIENT=4

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
     INTEGER(c_int) FUNCTION c_transient(mu,st,nst,sig,eik,sx1,sx2,sx3,dx1,dx2,dx3,mom,eikdot,mw,dg) &
          BIND(C,name="relax_transienteigenstress")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: mu,dx1,dx2,dx3
       TYPE(c_ptr), VALUE :: st,sig,eik,mom,eikdot,mw,dg
       INTEGER(c_int), VALUE :: nst,sx1,sx2,sx3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_source(mu,s,x,y,z,L,W,strike,dip,rake,beta,sx1,sx2,sx3,dx1,dx2,dx3, &
          f1,f2,f3,t1,t2,t3,al) BIND(C,name="relax_source")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: mu,s,x,y,z,L,W,strike,dip,rake,beta,dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,al
       TYPE(c_ptr), VALUE :: f1,f2,f3,t1,t2,t3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_traction(mu,l,nl,sx1,sx2,dx1,dx2,t,Dt,t3,rate,al) BIND(C,name="relax_traction")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: mu,dx1,dx2,t,Dt
       INTEGER(c_int), VALUE :: nl,sx1,sx2,rate,al
       TYPE(c_ptr), VALUE :: l,t3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_builddg(sx1,sx2,sx3,dx1,dx2,dx3,beta,nz,zones,dg) &
          BIND(C,name="relax_builddgammadot0")
       IMPORT :: c_int, c_double, c_ptr
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,nz
       REAL(c_double), VALUE :: dx1,dx2,dx3,beta
       TYPE(c_ptr), VALUE :: zones,dg
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_set_weakzones(h,which,zones,nz) BIND(C,name="relax_model_set_weakzones")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h,zones
       INTEGER(c_int), VALUE :: which,nz
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_transient_strain(h,eik) BIND(C,name="relax_model_get_transient_strain")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h,eik
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
     ! device-resident model (relax_model_* of relax_hip.h)
     INTEGER(c_int) FUNCTION c_model_create(d,h) BIND(C,name="relax_model_create")
       IMPORT :: c_int, c_ptr, relax_model_desc
       TYPE(relax_model_desc), INTENT(IN) :: d
       TYPE(c_ptr), INTENT(OUT) :: h
     END FUNCTION
     INTEGER(c_int) FUNCTION c_batch_create(h) BIND(C,name="relax_batch_create")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), INTENT(OUT) :: h
     END FUNCTION
     INTEGER(c_int) FUNCTION c_batch_destroy(h) BIND(C,name="relax_batch_destroy")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
     END FUNCTION
     INTEGER(c_int) FUNCTION c_batch_add(h,m) BIND(C,name="relax_batch_add")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, m
     END FUNCTION
     INTEGER(c_int) FUNCTION c_batch_remove(h,m) BIND(C,name="relax_batch_remove")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, m
     END FUNCTION
     INTEGER(c_int) FUNCTION c_batch_step(h,infos) BIND(C,name="relax_batch_step")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, infos
     END FUNCTION
     INTEGER(c_int) FUNCTION c_batch_autotune(h,choice,ms) BIND(C,name="relax_batch_autotune")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, choice, ms
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_destroy(h) BIND(C,name="relax_model_destroy")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_set_structure(h,which,layers,n) BIND(C,name="relax_model_set_structure")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, layers
       INTEGER(c_int), VALUE :: which, n
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_set_prestress(h,p,n) BIND(C,name="relax_model_set_prestress")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, p
       INTEGER(c_int), VALUE :: n
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_set_planes(h,planes,np,patches) BIND(C,name="relax_model_set_planes")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, planes, patches
       INTEGER(c_int), VALUE :: np
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_upload(h,u1,u2,u3,tau,sig,gam,t,oi,i,al) BIND(C,name="relax_model_upload")
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h, u1, u2, u3, tau, sig, gam
       REAL(c_double), VALUE :: t
       INTEGER(c_int), VALUE :: oi, i, al
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_download(h,u1,u2,u3,tau,sig,gam,t,al) BIND(C,name="relax_model_download")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, u1, u2, u3, tau, sig, gam, t
       INTEGER(c_int), VALUE :: al
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_step(h,dt,info) BIND(C,name="relax_model_step")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, dt, info
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_add_shear(h,patches,slip,n) BIND(C,name="relax_model_add_shear")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, patches, slip
       INTEGER(c_int), VALUE :: n
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_event(h,ev,nyquist) BIND(C,name="relax_model_event")
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: h,ev
       REAL(c_double), VALUE :: nyquist
     END FUNCTION
     INTEGER(c_int) FUNCTION c_tensilesource(lambda,mu,s,x,y,z,L,W,strike,dip,beta,sx1,sx2,sx3,dx1,dx2,dx3, &
          f1,f2,f3,al) BIND(C,name="relax_tensilesource")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: lambda,mu,s,x,y,z,L,W,strike,dip,beta,dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,al
       TYPE(c_ptr), VALUE :: f1,f2,f3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_mogisource(lambda,mu,o,x,y,z,sx1,sx2,sx3,dx1,dx2,dx3,f1,f2,f3,al) &
          BIND(C,name="relax_mogisource")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: lambda,mu,o,x,y,z,dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3,al
       TYPE(c_ptr), VALUE :: f1,f2,f3
     END FUNCTION
     INTEGER(c_int) FUNCTION c_mdtensile(lambda,mu,s,x,y,z,L,W,strike,dip,beta,sx1,sx2,sx3,dx1,dx2,dx3,sig) &
          BIND(C,name="relax_momentdensitytensile")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: lambda,mu,s,x,y,z,L,W,strike,dip,beta,dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3
       TYPE(c_ptr), VALUE :: sig
     END FUNCTION
     INTEGER(c_int) FUNCTION c_mdmogi(lambda,mu,o,x,y,z,sx1,sx2,sx3,dx1,dx2,dx3,sig) &
          BIND(C,name="relax_momentdensitymogi")
       IMPORT :: c_int, c_double, c_ptr
       REAL(c_double), VALUE :: lambda,mu,o,x,y,z,dx1,dx2,dx3
       INTEGER(c_int), VALUE :: sx1,sx2,sx3
       TYPE(c_ptr), VALUE :: sig
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_coseismic(h) BIND(C,name="relax_model_coseismic")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_okada(h,sc,ns,tsc,nt,nyquist) BIND(C,name="relax_model_okada")
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: h,sc,tsc
       INTEGER(c_int), VALUE :: ns,nt
       REAL(c_double), VALUE :: nyquist
     END FUNCTION
     INTEGER(c_int) FUNCTION c_model_set_interseismic(h,ev) BIND(C,name="relax_model_set_interseismic")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h,ev
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

  !> epsilonik and epsilonikdot may be the same array (the corrector's call,
  !! relax.f90:768-790)
  SUBROUTINE transienteigenstress(mu,structure,sig,epsilonik,sx1,sx2,sx3, &
       dx1,dx2,dx3,moment,epsilonikdot,maxwelltime,dgammadot0)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*4, DIMENSION(sx1,sx2,sx3), INTENT(IN), OPTIONAL, TARGET :: dgammadot0
    REAL*8, INTENT(IN) :: mu,dx1,dx2,dx3
    TYPE(LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: structure
    TYPE(TENSOR), INTENT(IN), DIMENSION(sx1,sx2,sx3), TARGET :: sig
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: moment
    REAL*8, OPTIONAL, INTENT(INOUT), TARGET :: maxwelltime
    TYPE(TENSOR), INTENT(IN), DIMENSION(sx1,sx2,sx3), TARGET :: epsilonik
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: epsilonikdot
    TYPE(c_ptr) :: pm,pd
    pm=C_NULL_PTR; pd=C_NULL_PTR
    IF (PRESENT(maxwelltime)) pm=C_LOC(maxwelltime)
    IF (PRESENT(dgammadot0)) pd=C_LOC(dgammadot0)
    CALL check(c_transient(mu,C_LOC(structure),SIZE(structure),C_LOC(sig),C_LOC(epsilonik), &
         sx1,sx2,sx3,dx1,dx2,dx3,C_LOC(moment),C_LOC(epsilonikdot),pm,pd),"transienteigenstress")
  END SUBROUTINE transienteigenstress

  SUBROUTINE builddgammadot0(sx1,sx2,sx3,dx1,dx2,dx3,beta,nz,ductilezones,dgammadot0)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3,nz
    REAL*4, DIMENSION(sx1,sx2,sx3), INTENT(OUT), TARGET :: dgammadot0
    TYPE(WEAK_STRUCT), DIMENSION(nz), INTENT(IN), TARGET :: ductilezones
    REAL*8, INTENT(IN) :: dx1,dx2,dx3,beta
    TYPE(c_ptr) :: pz
    pz=C_NULL_PTR
    IF (nz .GT. 0) pz=C_LOC(ductilezones)
    CALL check(c_builddg(sx1,sx2,sx3,dx1,dx2,dx3,beta,nz,pz,C_LOC(dgammadot0)),"builddgammadot0")
  END SUBROUTINE builddgammadot0

  SUBROUTINE source(mu,s,x,y,z,L,W,strike,dip,rake,beta,sx1,sx2,sx3,dx1,dx2,dx3,f1,f2,f3,t1,t2,t3)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: mu,s,x,y,z,L,W,strike,dip,rake,beta,dx1,dx2,dx3
    REAL*4, INTENT(INOUT), DIMENSION(*), TARGET :: f1,f2,f3,t1,t2,t3
    CALL check(c_source(mu,s,x,y,z,L,W,strike,dip,rake,beta,sx1,sx2,sx3,dx1,dx2,dx3, &
         C_LOC(f1),C_LOC(f2),C_LOC(f3),C_LOC(t1),C_LOC(t2),C_LOC(t3),relax_hip_align),"source")
  END SUBROUTINE source

  SUBROUTINE tensilesource(lambda,mu,s,x,y,z,L,W,strike,dip,beta,sx1,sx2,sx3,dx1,dx2,dx3,f1,f2,f3)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: lambda,mu,s,x,y,z,L,W,strike,dip,beta,dx1,dx2,dx3
    REAL*4, INTENT(INOUT), DIMENSION(*), TARGET :: f1,f2,f3
    CALL check(c_tensilesource(lambda,mu,s,x,y,z,L,W,strike,dip,beta,sx1,sx2,sx3,dx1,dx2,dx3, &
         C_LOC(f1),C_LOC(f2),C_LOC(f3),relax_hip_align),"tensilesource")
  END SUBROUTINE tensilesource

  SUBROUTINE mogisource(lambda,mu,o,xs,ys,zs,sx1,sx2,sx3,dx1,dx2,dx3,f1,f2,f3)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: lambda,mu,o,xs,ys,zs,dx1,dx2,dx3
    REAL*4, INTENT(INOUT), DIMENSION(*), TARGET :: f1,f2,f3
    CALL check(c_mogisource(lambda,mu,o,xs,ys,zs,sx1,sx2,sx3,dx1,dx2,dx3,C_LOC(f1),C_LOC(f2),C_LOC(f3), &
         relax_hip_align),"mogisource")
  END SUBROUTINE mogisource

  SUBROUTINE momentdensitytensile(lambda,mu,slip,x,y,z,L,W,strike,dip,rake,beta,sx1,sx2,sx3,dx1,dx2,dx3,sig)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: lambda,mu,slip,x,y,z,L,W,strike,dip,rake,beta,dx1,dx2,dx3
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: sig
    CALL check(c_mdtensile(lambda,mu,slip,x,y,z,L,W,strike,dip,beta,sx1,sx2,sx3,dx1,dx2,dx3,C_LOC(sig)), &
         "momentdensitytensile")
  END SUBROUTINE momentdensitytensile

  SUBROUTINE momentdensitymogi(lambda,mu,o,xs,ys,zs,sx1,sx2,sx3,dx1,dx2,dx3,sig)
    INTEGER, INTENT(IN) :: sx1,sx2,sx3
    REAL*8, INTENT(IN) :: lambda,mu,o,xs,ys,zs,dx1,dx2,dx3
    TYPE(TENSOR), INTENT(INOUT), DIMENSION(sx1,sx2,sx3), TARGET :: sig
    CALL check(c_mdmogi(lambda,mu,o,xs,ys,zs,sx1,sx2,sx3,dx1,dx2,dx3,C_LOC(sig)),"momentdensitymogi")
  END SUBROUTINE momentdensitymogi

  !> traction (elastic3d.f90:2517-2590) over a list of surface loads
  SUBROUTINE relax_hip_traction(mu,l,sx1,sx2,dx1,dx2,t,Dt,t3,rate)
    TYPE(SOURCE_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: l
    INTEGER, INTENT(IN) :: sx1,sx2
    REAL*8, INTENT(IN) :: mu,dx1,dx2,t,Dt
    REAL*4, INTENT(INOUT), DIMENSION(*), TARGET :: t3
    LOGICAL, INTENT(IN), OPTIONAL :: rate
    INTEGER(c_int) :: r
    TYPE(c_ptr) :: pl
    r=0
    IF (PRESENT(rate)) THEN
       IF (rate) r=1
    END IF
    pl=C_NULL_PTR
    IF (SIZE(l) .GT. 0) pl=C_LOC(l)
    CALL check(c_traction(mu,pl,INT(SIZE(l),c_int),sx1,sx2,dx1,dx2,t,Dt,C_LOC(t3),r,relax_hip_align),"traction")
  END SUBROUTINE relax_hip_traction

#ifdef RELAX_REFERENCE_TYPES
  SUBROUTINE traction(mu,e,sx1,sx2,dx1,dx2,t,Dt,t3,rate)
    TYPE(EVENT_STRUC), INTENT(IN) :: e
    INTEGER, INTENT(IN) :: sx1,sx2
    REAL*8, INTENT(IN) :: mu,dx1,dx2,t,Dt
    REAL*4, INTENT(INOUT), DIMENSION(*), TARGET :: t3
    LOGICAL, INTENT(IN), OPTIONAL :: rate
    IF (e%nl .LE. 0) RETURN
    CALL relax_hip_traction(mu,e%l(1:e%nl),sx1,sx2,dx1,dx2,t,Dt,t3,rate)
  END SUBROUTINE traction
#endif

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

  !------------------------------------------------ device-resident model --
  SUBROUTINE relax_hip_model_create(desc,model)
    TYPE(relax_model_desc), INTENT(IN) :: desc
    TYPE(relax_model), INTENT(OUT) :: model
    CALL check(c_model_create(desc,model%h),"relax_model_create")
  END SUBROUTINE relax_hip_model_create

  SUBROUTINE relax_hip_model_destroy(model)
    TYPE(relax_model), INTENT(INOUT) :: model
    CALL check(c_model_destroy(model%h),"relax_model_destroy")
    model%h=C_NULL_PTR
  END SUBROUTINE relax_hip_model_destroy

  !> batches (no counterpart in the reference, whose driver runs one model
  !! per process): models added to a batch step together, one RK2 step of
  !! every member per relax_hip_batch_step; a member's own step calls are
  !! refused while it is in the batch
  SUBROUTINE relax_hip_batch_create(batch)
    TYPE(relax_batch), INTENT(OUT) :: batch
    CALL check(c_batch_create(batch%h),"relax_batch_create")
  END SUBROUTINE relax_hip_batch_create

  SUBROUTINE relax_hip_batch_destroy(batch)
    TYPE(relax_batch), INTENT(INOUT) :: batch
    CALL check(c_batch_destroy(batch%h),"relax_batch_destroy")
    batch%h=C_NULL_PTR
  END SUBROUTINE relax_hip_batch_destroy

  SUBROUTINE relax_hip_batch_add(batch,model)
    TYPE(relax_batch), INTENT(IN) :: batch
    TYPE(relax_model), INTENT(IN) :: model
    CALL check(c_batch_add(batch%h,model%h),"relax_batch_add")
  END SUBROUTINE relax_hip_batch_add

  SUBROUTINE relax_hip_batch_remove(batch,model)
    TYPE(relax_batch), INTENT(IN) :: batch
    TYPE(relax_model), INTENT(IN) :: model
    CALL check(c_batch_remove(batch%h,model%h),"relax_batch_remove")
  END SUBROUTINE relax_hip_batch_remove

  !> one RK2 step of every member; infos (one entry per member, in the order
  !! of addition) receives their scalars when given
  SUBROUTINE relax_hip_batch_step(batch,infos)
    TYPE(relax_batch), INTENT(IN) :: batch
    TYPE(relax_step_info), DIMENSION(:), INTENT(OUT), OPTIONAL, TARGET, CONTIGUOUS :: infos
    TYPE(c_ptr) :: pi
    pi=C_NULL_PTR
    IF (PRESENT(infos)) pi=C_LOC(infos)
    CALL check(c_batch_step(batch%h,pi),"relax_batch_step")
  END SUBROUTINE relax_hip_batch_step

  !> launch shapes timed on the batch (a batch of one grid class)
  SUBROUTINE relax_hip_batch_autotune(batch)
    TYPE(relax_batch), INTENT(IN) :: batch
    CALL check(c_batch_autotune(batch%h,C_NULL_PTR,C_NULL_PTR),"relax_batch_autotune")
  END SUBROUTINE relax_hip_batch_autotune

  !> which: RELAX_LINEAR (in%linearstruc), RELAX_POWERLAW (in%nonlinearstruc)
  !! or RELAX_FAULTCREEP (in%faultcreepstruc), sampled by viscoelasticstructure
  SUBROUTINE relax_hip_model_set_structure(model,which,structure)
    TYPE(relax_model), INTENT(IN) :: model
    INTEGER, INTENT(IN) :: which
    TYPE(LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: structure
    CALL check(c_model_set_structure(model%h,INT(which,c_int),C_LOC(structure),INT(SIZE(structure),c_int)), &
         "relax_model_set_structure")
  END SUBROUTINE relax_hip_model_set_structure

  !> the weak zones of a mechanism (RELAX_LINEAR, RELAX_POWERLAW,
  !! RELAX_LTRANSIENT, RELAX_NLTRANSIENT): in%linearweakzone, ...
  SUBROUTINE relax_hip_model_set_weakzones(model,which,zones)
    TYPE(relax_model), INTENT(IN) :: model
    INTEGER, INTENT(IN) :: which
    TYPE(WEAK_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: zones
    TYPE(c_ptr) :: pz
    pz=C_NULL_PTR
    IF (SIZE(zones) .GT. 0) pz=C_LOC(zones)
    CALL check(c_model_set_weakzones(model%h,INT(which,c_int),pz,INT(SIZE(zones),c_int)), &
         "relax_model_set_weakzones")
  END SUBROUTINE relax_hip_model_set_weakzones

  !> the transient strain epsilonik (sx1,sx2,sx3/2) of a model with transient creep
  SUBROUTINE relax_hip_model_transient_strain(model,epsilonik)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(TENSOR), DIMENSION(:,:,:), INTENT(OUT), TARGET, CONTIGUOUS :: epsilonik
    CALL check(c_model_transient_strain(model%h,C_LOC(epsilonik)),"relax_model_get_transient_strain")
  END SUBROUTINE relax_hip_model_transient_strain

  SUBROUTINE relax_hip_model_set_prestress(model,prestress)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(TENSOR_LAYER_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: prestress
    CALL check(c_model_set_prestress(model%h,C_LOC(prestress),INT(SIZE(prestress),c_int)), &
         "relax_model_set_prestress")
  END SUBROUTINE relax_hip_model_set_prestress

  !> the driver's friction planes (in%n), geometry in the internal
  !! convention as the input routines leave it; their patches start at zero
  SUBROUTINE relax_hip_model_set_planes(model,n)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(PLANE_STRUCT), DIMENSION(:), INTENT(IN) :: n
    TYPE(relax_plane_c), DIMENSION(:), ALLOCATABLE, TARGET :: p
    INTEGER :: k
    ALLOCATE(p(MAX(1,SIZE(n))))
    DO k=1,SIZE(n)
       p(k)=relax_plane_c(n(k)%x,n(k)%y,n(k)%z,n(k)%width,n(k)%length,n(k)%strike,n(k)%dip,n(k)%rake, &
            INT(n(k)%px2,c_int),INT(n(k)%px3,c_int))
    END DO
    CALL check(c_model_set_planes(model%h,C_LOC(p),INT(SIZE(n),c_int),C_NULL_PTR),"relax_model_set_planes")
    DEALLOCATE(p)
  END SUBROUTINE relax_hip_model_set_planes

  !> fields in the driver's layout (relax_hip_align); sig may be omitted: it
  !! is then -tau + C:eps(u) (relax.f90:369-371)
  SUBROUTINE relax_hip_model_upload(model,u1,u2,u3,tau,t,oi,i,sig,gamma)
    TYPE(relax_model), INTENT(IN) :: model
    REAL*4, DIMENSION(*), INTENT(IN), TARGET :: u1,u2,u3
    TYPE(TENSOR), DIMENSION(*), INTENT(IN), TARGET :: tau
    REAL*8, INTENT(IN) :: t
    INTEGER, INTENT(IN) :: oi,i
    TYPE(TENSOR), DIMENSION(*), INTENT(IN), OPTIONAL, TARGET :: sig
    REAL*4, DIMENSION(*), INTENT(IN), OPTIONAL, TARGET :: gamma
    TYPE(c_ptr) :: ps,pg
    ps=C_NULL_PTR; pg=C_NULL_PTR
    IF (PRESENT(sig)) ps=C_LOC(sig)
    IF (PRESENT(gamma)) pg=C_LOC(gamma)
    CALL check(c_model_upload(model%h,C_LOC(u1),C_LOC(u2),C_LOC(u3),C_LOC(tau),ps,pg,t,INT(oi,c_int), &
         INT(i,c_int),INT(relax_hip_align,c_int)),"relax_model_upload")
  END SUBROUTINE relax_hip_model_upload

  SUBROUTINE relax_hip_model_download(model,u1,u2,u3,tau,sig,gamma,t)
    TYPE(relax_model), INTENT(IN) :: model
    REAL*4, DIMENSION(*), INTENT(OUT), TARGET :: u1,u2,u3
    TYPE(TENSOR), DIMENSION(*), INTENT(OUT), OPTIONAL, TARGET :: tau,sig
    REAL*4, DIMENSION(*), INTENT(OUT), OPTIONAL, TARGET :: gamma
    REAL*8, INTENT(OUT), OPTIONAL, TARGET :: t
    TYPE(c_ptr) :: pt,ps,pg,pp
    pt=C_NULL_PTR; ps=C_NULL_PTR; pg=C_NULL_PTR; pp=C_NULL_PTR
    IF (PRESENT(tau)) pt=C_LOC(tau)
    IF (PRESENT(sig)) ps=C_LOC(sig)
    IF (PRESENT(gamma)) pg=C_LOC(gamma)
    IF (PRESENT(t)) pp=C_LOC(t)
    CALL check(c_model_download(model%h,C_LOC(u1),C_LOC(u2),C_LOC(u3),pt,ps,pg,pp,INT(relax_hip_align,c_int)), &
         "relax_model_download")
  END SUBROUTINE relax_hip_model_download

  !> one RK2 step (relax.f90:555-890); dt replaces the adaptive step when given
  SUBROUTINE relax_hip_model_step(model,info,dt)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(relax_step_info), INTENT(OUT), TARGET :: info
    REAL*8, INTENT(IN), OPTIONAL, TARGET :: dt
    TYPE(c_ptr) :: pd
    pd=C_NULL_PTR
    IF (PRESENT(dt)) pd=C_LOC(dt)
    CALL check(c_model_step(model%h,pd,C_LOC(info)),"relax_model_step")
  END SUBROUTINE relax_hip_model_step

  !> momentdensityshear of each patch (internal convention) into tau, then
  !! the elastic u and sigma of that tau (FIR divergence as body force)
  SUBROUTINE relax_hip_model_coseismic(model,patches,slip)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(relax_plane_c), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: patches
    REAL*8, DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: slip
    CALL check(c_model_add_shear(model%h,C_LOC(patches),C_LOC(slip),INT(SIZE(patches),c_int)), &
         "relax_model_add_shear")
    CALL check(c_model_coseismic(model%h),"relax_model_coseismic")
  END SUBROUTINE relax_hip_model_coseismic

  !> a coseismic event: the first at t = 0 (relax.f90:316-371) or a later
  !! one once a step has reached its time (info%e advanced, the event switch
  !! of relax.f90:858-889): dislocations, traction, Green, stress
  !! (shear patches s, surface loads l; optional tensile cracks ts, point
  !! dilatation sources m, eigenstrain cuboids and tetrahedra with their
  !! vertices, as EVENT_STRUC holds them, types.f90:99-106)
  SUBROUTINE relax_hip_model_event(model,s,l,nyquist,ts,m,cuboid,tetrahedron,vertex)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(SOURCE_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: s,l
    REAL*8, INTENT(IN) :: nyquist
    TYPE(SOURCE_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS, OPTIONAL :: ts,m
    TYPE(WEAK_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS, OPTIONAL :: cuboid
    TYPE(TETRAHEDRON_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS, OPTIONAL :: tetrahedron
    REAL*8, DIMENSION(:,:), INTENT(IN), TARGET, CONTIGUOUS, OPTIONAL :: vertex
    TYPE(relax_event_c), TARGET :: ev
    ev%ns=INT(SIZE(s),c_int)
    IF (ev%ns .GT. 0) ev%s=C_LOC(s)
    ev%nl=INT(SIZE(l),c_int)
    IF (ev%nl .GT. 0) ev%l=C_LOC(l)
    IF (PRESENT(ts)) THEN
       ev%nt=INT(SIZE(ts),c_int)
       IF (ev%nt .GT. 0) ev%ts=C_LOC(ts)
    END IF
    IF (PRESENT(m)) THEN
       ev%nm=INT(SIZE(m),c_int)
       IF (ev%nm .GT. 0) ev%m=C_LOC(m)
    END IF
    IF (PRESENT(cuboid)) THEN
       ev%nc=INT(SIZE(cuboid),c_int)
       IF (ev%nc .GT. 0) ev%c=C_LOC(cuboid)
    END IF
    IF (PRESENT(tetrahedron) .AND. PRESENT(vertex)) THEN
       ev%nth=INT(SIZE(tetrahedron),c_int)
       IF (ev%nth .GT. 0) ev%th=C_LOC(tetrahedron)
       ev%nv=INT(SIZE(vertex,2),c_int)
       IF (ev%nv .GT. 0) ev%v=C_LOC(vertex)
    END IF
    CALL check(c_model_event(model%h,C_LOC(ev),nyquist),"relax_model_event")
  END SUBROUTINE relax_hip_model_event

  !> Okada's displacement of the shear segments sc and tensile segments tsc
  !! at or below the Nyquist size (dislocations_disp, relax.f90:357-360,
  !! :1300-1349), from the deck's own copies (in%events(1)%sc, %tsc)
  SUBROUTINE relax_hip_model_okada(model,sc,tsc,nyquist)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(SOURCE_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: sc,tsc
    REAL*8, INTENT(IN) :: nyquist
    TYPE(c_ptr) :: ps,pt
    ps=C_NULL_PTR
    pt=C_NULL_PTR
    IF (SIZE(sc) .GT. 0) ps=C_LOC(sc)
    IF (SIZE(tsc) .GT. 0) pt=C_LOC(tsc)
    CALL check(c_model_okada(model%h,ps,INT(SIZE(sc),c_int),pt,INT(SIZE(tsc),c_int),nyquist),"relax_model_okada")
  END SUBROUTINE relax_hip_model_okada

  !> inter-seismic loading (in%inter: shear segments s and tensile segments
  !! ts whose slip or opening is a rate; relax.f90:669-674, :795-799)
  SUBROUTINE relax_hip_model_set_interseismic(model,s,ts)
    TYPE(relax_model), INTENT(IN) :: model
    TYPE(SOURCE_STRUCT), DIMENSION(:), INTENT(IN), TARGET, CONTIGUOUS :: s,ts
    TYPE(relax_event_c), TARGET :: ev
    ev%ns=INT(SIZE(s),c_int)
    IF (ev%ns .GT. 0) ev%s=C_LOC(s)
    ev%nt=INT(SIZE(ts),c_int)
    IF (ev%nt .GT. 0) ev%ts=C_LOC(ts)
    CALL check(c_model_set_interseismic(model%h,C_LOC(ev)),"relax_model_set_interseismic")
  END SUBROUTINE relax_hip_model_set_interseismic

END MODULE relax_hip
