/*
 * relax_hip.h -- C ABI of the MI355X (gfx950) implementation of RELAX's
 * per-time-step elastic + relaxation solve.
 *
 * Two layers:
 *  1. Per-routine entry points with host pointers, one per reference
 *     subroutine of the hot path.  Arguments keep the reference's order and
 *     meaning; arrays are plain pointers plus explicit sizes.  Each call copies
 *     its arrays to the GPU, runs the HIP kernels and copies the results back.
 *     `align_data` selects the reference's space-domain layout: 1 = padded rows
 *     of sx1+2 floats (ALIGN_DATA builds: FFTW, MKL, SGI; src/include.f90:45-56),
 *     0 = packed rows of sx1 floats (the built-in ctfft build).
 *  2. A device-resident model handle for the RK2 loop (relax.f90:555-890):
 *     the whole state stays in HBM between steps and one call advances it by
 *     one predictor/corrector step.
 *
 * Every function returns a status: RELAX_OK or one of the RELAX_ERR_* codes.
 * Nothing in this library exits the process; the reference's STOP
 * conditions map to the codes below.
 *
 * Paths in citations are relative to the reference's src/ directory.
 */
#ifndef RELAX_HIP_H
#define RELAX_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define RELAX_OK 0
#define RELAX_ERR_POWER 1      /* stress exponent < 1 (viscoelastic3d.f90:262-267, STOP 1) */
#define RELAX_ERR_FILTER 2     /* dimension too small for a FIR filter (elastic3d.f90:886, STOP 2) */
#define RELAX_ERR_FRICTION 5   /* tau/tauc > 20 in fault creep (friction3d.f90:204-215, STOP 5) */
#define RELAX_ERR_ARG 10       /* invalid argument (null pointer, bad flag) */
#define RELAX_ERR_SIZE 11      /* grid size not supported (powers of two, 16..1024) */
#define RELAX_ERR_NOMEM 12     /* device allocation failed */
#define RELAX_ERR_HIP 13       /* HIP runtime error */
#define RELAX_ERR_ADMISSION 14 /* model does not fit the HBM budget (cuelastic.cu:2096-2137 analogue) */
#define RELAX_ERR_RANGE 15     /* patch centre outside the stress field (elastic3d.f90:4068-4095) */
#define RELAX_ERR_NODEVICE 16  /* no HIP device */

/* ---- ABI structures: the reference's SEQUENCE types (types.f90) ---------- */
typedef struct { float s11, s12, s13, s22, s23, s33; } relax_tensor_t;          /* types.f90:41-44 */
typedef struct { double z, gammadot0, stressexponent, cohesion, friction, Gk; } relax_layer_t; /* types.f90:31-34 */
typedef struct { float z, dum; relax_tensor_t t; } relax_tensor_layer_t;        /* types.f90:52-56 */
typedef struct {                                                                 /* types.f90:76-90 */
  double x1, x2, x3, lx, lz, slip, ss, ds, v, vss, vds, taus;
  relax_tensor_t sig;
} relax_slippatch_t;
typedef struct {                                                                 /* types.f90:92-97 */
  double x, y, z, width, length, strike, dip, rake;
  int px2, px3; /* the patch array is passed separately */
} relax_plane_t;
typedef struct {                                                                 /* types.f90:46-50 */
  double dgammadot0, x, y, z, width, length, thickness, strike, dip;  /* angles in radians */
  relax_tensor_t e;
} relax_weak_t;
typedef struct {                                                                 /* types.f90:26-29 */
  /* a shear patch or a surface load of an event, as input.f90 leaves it
   * (angles in radians, the internal convention of wangconvention) */
  double slip, x, y, z, width, length, strike, dip, rake, period, phase, beta, opening;
} relax_source_t;
typedef struct {                                                                 /* types.f90:70-74 */
  /* a tetrahedron of uniform eigenstrain e between four vertices (1-based
   * indices into the event's vertex list) */
  relax_tensor_t e;
  int i1, i2, i3, i4;
} relax_tetrahedron_t;
typedef struct {                                                                 /* types.f90:99-106 */
  /* the sources of one event (EVENT_STRUC without its time and copies) */
  const relax_source_t* s;  int ns;   /* shear patches (slip) */
  const relax_source_t* ts; int nt;   /* tensile cracks (opening) */
  const relax_source_t* m;  int nm;   /* point dilatation sources (slip: volume change) */
  const relax_source_t* l;  int nl;   /* surface loads */
  /* distributed eigenstrain (--with-eigenstrain, input.f90:1496-1671):
   * cuboids as WEAK_STRUCT (the eigenstrain in e; centre, width, length,
   * thickness, strike and dip after Wang's convention), tetrahedra and their
   * vertices (nv rows of x1, x2, x3; not moved by the frame) */
  const relax_weak_t* c; int nc;
  const relax_tetrahedron_t* th; int nth;
  const double* v; int nv;
} relax_event_t;

/* ---- library ------------------------------------------------------------ */
const char* relax_hip_version(void);
const char* relax_hip_status_string(int status);
int relax_hip_init(int device); /* selects the device; optional */
/* Free and total device memory of the current device (for admission). */
int relax_hip_meminfo(size_t* free_bytes, size_t* total_bytes);

/* ---- fourier.f90 --------------------------------------------------------- */
/* fft3, fourier.f90:177-201 (FFTW) / :330-355 (ctfft): in-place R2C
 * (direction -1, scaled by dx1*dx2*dx3) or C2R (direction +1, scaled by
 * 1/(sx1*dx1*sx2*dx2*sx3*dx3)) of a (sx1+2)*sx2*sx3 float buffer. */
int relax_fft3(float* data, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, int direction,
               int align_data);
/* fft2, fourier.f90:372-396 / :510-535, the same in two dimensions. */
int relax_fft2(float* data, int sx1, int sx2, double dx1, double dx2, int direction, int align_data);

/* ---- green.f90 (Fourier-domain arrays, complex interleaved, sx1/2+1 per row) */
/* elasticresponse, green.f90:51-101 */
int relax_elasticresponse(double lambda, double mu, float* f1, float* f2, float* f3, int sx1, int sx2, int sx3,
                          double dx1, double dx2, double dx3);
/* surfacetractioncowling, green.f90:327-377; p1..p3 are (sx1+2)*sx2 outputs */
int relax_surfacetractioncowling(double lambda, double mu, double gamma, const float* u1, const float* u2,
                                 const float* u3, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,
                                 float* p1, float* p2, float* p3);
/* cerrutimodified, green.f90:520-612 */
int relax_cerrutimodified(const float* p1, const float* p2, const float* p3, double lambda, double mu,
                          double gamma, float* u1, float* u2, float* u3, int sx1, int sx2, int sx3, double dx1,
                          double dx2, double dx3);
/* greenfunctioncowling, green.f90:1008-1129: body forces c1..c3 (full grid,
 * (sx1+2)*sx2*sx3) and surface tractions t1..t3 ((sx1+2)*sx2) in, the
 * displacement field out in c1..c3 (in place). */
int relax_greenfunctioncowling(float* c1, float* c2, float* c3, float* t1, float* t2, float* t3, int sx1,
                               int sx2, int sx3, double dx1, double dx2, double dx3, double lambda, double mu,
                               double gamma, int align_data);

/* greenfunctioncowling on device buffers, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  c1..c3 and t1..t3 are device pointers
 * in the padded layout ((sx1+2)*sx2*sx3 and (sx1+2)*sx2 floats), transformed
 * in place as relax_greenfunctioncowling does; the work is enqueued on
 * `stream` (a hipStream_t, or NULL for the null stream) and the call returns
 * without waiting.  The three traction planes it needs live in a static
 * device array of 12.0 MiB (sx1, sx2 up to 1024; larger grids return
 * RELAX_ERR_SIZE), one per device, shared by every call on that device:
 * calls that overlap on one device race.  relax_hip_init sets it up for the
 * device it selects; otherwise the first call on a device does. */
int relax_greenfunctioncowling_noalloc(float* c1, float* c2, float* c3, float* t1, float* t2, float* t3, int sx1,
                                       int sx2, int sx3, double dx1, double dx2, double dx3, double lambda,
                                       double mu, double gamma, void* stream);

/* ---- elastic3d.f90 ------------------------------------------------------- */
/* stressupdate, elastic3d.f90:956-1180: sig += C:eps(v); sx3 is the depth
 * extent of the tensor field (half the FFT grid in the RK2 loop). */
int relax_stressupdate(const float* v1, const float* v2, const float* v3, double lambda, double mu, double dx1,
                       double dx2, double dx3, int sx1, int sx2, int sx3, relax_tensor_t* sig, int align_data);
/* equivalentbodyforce, elastic3d.f90:1214-1506: c -= div(sig),
 * t += sig(:,:,1).e3; mask (sx3 floats) may be NULL. */
int relax_equivalentbodyforce(const relax_tensor_t* sig, double dx1, double dx2, double dx3, int sx1, int sx2,
                              int sx3, float* c1, float* c2, float* c3, float* t1, float* t2, float* t3,
                              const float* mask, int align_data);
/* fieldadd, elastic3d.f90:3774-3812: u = c1*u + c2*v over n floats
 * (c1, c2 may be NULL = absent optional argument). */
int relax_fieldadd(float* u, const float* v, long n, const float* c1, const float* c2);
/* tensorfieldadd, elastic3d.f90:3880-3959: t1 = c1*t1 + c2*t2 over n tensors */
int relax_tensorfieldadd(relax_tensor_t* t1, const relax_tensor_t* t2, long n, const float* c1, const float* c2);
/* tensoramplitude, elastic3d.f90:446-466 */
int relax_tensoramplitude(const relax_tensor_t* t, long n, double dx1, double dx2, double dx3, double* out);
/* momentdensityshear, elastic3d.f90:2661-2750 (coseismic set-up of tau) */
int relax_momentdensityshear(double mu, double slip, double x, double y, double z, double L, double W,
                             double strike, double dip, double rake, double beta, int sx1, int sx2, int sx3,
                             double dx1, double dx2, double dx3, relax_tensor_t* sig);

/* source, elastic3d.f90:2064-2252: the body force of a tapered shear
 * dislocation (L down dip, W along strike) added to the top sx3/2 planes of
 * f1..f3 ((sx1+2*align) x sx2 x sx3 arrays) and its surface traction to
 * t1..t3 ((sx1+2*align) x sx2). */
int relax_source(double mu, double slip, double x, double y, double z, double L, double W, double strike,
                 double dip, double rake, double beta, int sx1, int sx2, int sx3, double dx1, double dx2,
                 double dx3, float* f1, float* f2, float* f3, float* t1, float* t2, float* t3, int align);
/* traction, elastic3d.f90:2517-2590: the surface loads l[0..nl) of an event
 * subtracted from t3 ((sx1+2*align) x sx2) at time t, or their rate over
 * [t, t+Dt] when rate is nonzero. */
int relax_traction(double mu, const relax_source_t* l, int nl, int sx1, int sx2, double dx1, double dx2, double t,
                   double Dt, float* t3, int rate, int align);
/* tensilesource, elastic3d.f90:2315-2411, and mogisource, :2458-2509: the
 * body force of a tapered opening crack (opening s, L down dip, W along
 * strike) or of a point dilatation source (volume change o) added to the
 * whole depth of f1..f3 ((sx1+2*align) x sx2 x sx3). */
int relax_tensilesource(double lambda, double mu, double s, double x, double y, double z, double L, double W,
                        double strike, double dip, double beta, int sx1, int sx2, int sx3, double dx1, double dx2,
                        double dx3, float* f1, float* f2, float* f3, int align);
int relax_mogisource(double lambda, double mu, double o, double xs, double ys, double zs, int sx1, int sx2, int sx3,
                     double dx1, double dx2, double dx3, float* f1, float* f2, float* f3, int align);
/* momentdensitytensile, elastic3d.f90:3137-3215, and momentdensitymogi,
 * :3240-3286: their moment densities added to sig (sx1 x sx2 x sx3). */
int relax_momentdensitytensile(double lambda, double mu, double s, double x, double y, double z, double L, double W,
                               double strike, double dip, double beta, int sx1, int sx2, int sx3, double dx1,
                               double dx2, double dx3, relax_tensor_t* sig);
int relax_momentdensitymogi(double lambda, double mu, double o, double xs, double ys, double zs, int sx1, int sx2,
                            int sx3, double dx1, double dx2, double dx3, relax_tensor_t* sig);

/* grnfct_okada, okada/green_space.f90:119-146: Okada's displacement of one
 * rectangular dislocation (slip, opening; upper corner x, y, z at the start
 * of the strike, north-east-down; width, length; strike, dip, rake in
 * degrees) added to u1..u3 ((sx1+2*align) x sx2 x sx3, depth i3 at
 * (i3-1) dx3). */
int relax_grnfct_okada(double lambda, double mu, double slip, double opening, double x, double y, double z,
                       double width, double length, double strike, double dip, double rake, int sx1, int sx2, int sx3,
                       double dx1, double dx2, double dx3, float* u1, float* u2, float* u3, int align_data);

/* ---- viscoelastic3d.f90 / friction3d.f90 --------------------------------- */
/* viscouseigenstress, viscoelastic3d.f90:217-314.  maxwelltime, dgammadot0
 * and gamma are the optional arguments (NULL = absent). */
int relax_viscouseigenstress(double mu, const relax_layer_t* structure, int nstructure,
                             const relax_tensor_layer_t* prestress, int nprestress, const relax_tensor_t* sig,
                             int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, relax_tensor_t* moment,
                             double* maxwelltime, const float* dgammadot0, float* gamma, int align_data);
/* transienteigenstress, viscoelastic3d.f90:123-202 (Burgers transient
 * creep; structure Gk is the Kelvin shear modulus).  epsilonik and
 * epsilonikdot may be the same array, as in the corrector's call
 * (relax.f90:768-790).  maxwelltime and dgammadot0 are optional (NULL).  A
 * structure whose size is not sx3 is ignored (returns RELAX_OK). */
int relax_transienteigenstress(double mu, const relax_layer_t* structure, int nstructure,
                               const relax_tensor_t* sig, const relax_tensor_t* epsilonik, int sx1, int sx2,
                               int sx3, double dx1, double dx2, double dx3, relax_tensor_t* moment,
                               relax_tensor_t* epsilonikdot, double* maxwelltime, const float* dgammadot0);
/* builddgammadot0, viscoelastic3d.f90:320-435: the fluidity perturbation
 * (sx1, sx2, sx3) of nz weak zones, tapered by beta. */
int relax_builddgammadot0(int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, double beta, int nz,
                          const relax_weak_t* zones, float* dgammadot0);
/* frictioneigenstress, friction3d.f90:82-241 (CPU path semantics) */
int relax_frictioneigenstress(double x, double y, double z, double L, double W, double strike, double dip,
                              double rake, double beta, const relax_tensor_t* sig,
                              const relax_tensor_layer_t* prestress, double mu, const relax_layer_t* structure,
                              int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,
                              relax_tensor_t* moment, double* maxwelltime, float* vel, int align_data);
/* monitorfriction, friction3d.f90:263-416; sx3 is the depth bound used for
 * the range test (the driver passes the full grid), sx3_alloc the depth of
 * the stress array. patch is px2*px3 (j2 fastest). */
int relax_monitorfriction(double x, double y, double z, double L, double W, double strike, double dip,
                          double rake, double beta, int sx1, int sx2, int sx3, int sx3_alloc, double dx1,
                          double dx2, double dx3, const relax_tensor_t* sig, const relax_tensor_layer_t* prestress,
                          const relax_layer_t* structure, relax_slippatch_t* patch, int px2, int px3);
/* frictionadd, friction3d.f90:424-453, for the patches of one plane */
int relax_frictionadd(relax_slippatch_t* patch, long npatch, double dt);

/* ---- device-resident model: the RK2 loop --------------------------------- */
typedef struct relax_model relax_model_t;

typedef struct {
  int sx1, sx2, sx3;                      /* full FFT grid (powers of two) */
  double dx1, dx2, dx3;
  double lambda, mu, gamma;               /* gamma: Cowling coefficient (in%gam) */
  double beta;                            /* taper width of the friction planes */
  double interval, odt, tscale;           /* time line (relax.f90:1420-1452) */
  int skip;
  int ne;                                 /* number of event times */
  const double* event_times;              /* ne entries; event 1 is at t=0 */
} relax_model_desc_t;

typedef struct {
  int i;              /* iteration index after the step (relax.f90:555) */
  double Dt, tm;      /* time step and relaxation time of the step */
  double maxwell[5];  /* linear, power-law, fault creep, linear and power-law
                         transient creep (STEP_MAX when inactive), relax.f90:565 */
  double power;       /* tensoramplitude(moment), relax.f90:1027-1046 */
  double cei;         /* tensoramplitude(tau) */
  double t;           /* time after the step */
  int oi;             /* output counter after the step */
  int status;         /* RELAX_OK or the error of the step */
  int e;              /* index of the last event reached (0: the one at t = 0); it
                         advances when t reaches the next event time (relax.f90:858-863) */
} relax_step_info_t;

/* Bytes of HBM one model of this size keeps resident (for admission). */
size_t relax_model_footprint(const relax_model_desc_t* desc);
int relax_model_create(const relax_model_desc_t* desc, relax_model_t** out);
int relax_model_destroy(relax_model_t* m);
/* Back to the state of a new model (zero fields, t = 0, no mechanisms, no
 * planes) without releasing or allocating device memory, so that a server
 * can reuse the model for another request of the same grid. */
int relax_model_reset(relax_model_t* m);
/* which: 0 = linear viscosity (in%linearstruc), 1 = power law
 * (in%nonlinearstruc), 2 = fault creep (in%faultcreepstruc), 3 = linear
 * transient creep (in%ltransientstruc), 4 = power-law transient creep
 * (in%nltransientstruc; relax.f90:614-650).  layers are the per-depth
 * samples of viscoelasticstructure (sx3/2 entries), or NULL to remove the
 * mechanism.  The first transient structure allocates the transient strain
 * epsilonik (48 bytes per point with its mid-step copy), set to zero. */
#define RELAX_LINEAR 0
#define RELAX_POWERLAW 1
#define RELAX_FAULTCREEP 2
#define RELAX_LTRANSIENT 3
#define RELAX_NLTRANSIENT 4
int relax_model_set_structure(relax_model_t* m, int which, const relax_layer_t* layers, int n);
/* Weak zones of mechanism which (0, 1, 3 or 4): the model adds the field
 * builddgammadot0 makes of them (taper 0, as relax.f90:487-541 passes) to
 * the mechanism's depth-dependent fluidity, as the DGAMMADOT0 argument does
 * (viscoelastic3d.f90:277-280, :164-167).  The field takes 4 bytes per
 * point; nz = 0 removes it. */
int relax_model_set_weakzones(relax_model_t* m, int which, const relax_weak_t* zones, int nz);
/* The transient strain epsilonik (TYPE(TENSOR) (sx1, sx2, sx3/2)): set or
 * read.  RELAX_ERR_ARG without a transient structure. */
int relax_model_set_transient_strain(relax_model_t* m, const relax_tensor_t* epsilonik);
int relax_model_get_transient_strain(relax_model_t* m, relax_tensor_t* epsilonik);
/* The background stress (TENSOR_LAYER_STRUCT, sx3/2 depths, as
 * tensorstructure leaves in%stressstruc, elastic3d.f90:630-702): the
 * deviatoric prestress the viscous rheologies read, and the initial
 * tau(:,:,k) = -prestress(k) of relax.f90:286-297, which overwrites tau --
 * call it before the first event, as the reference's set-up does. */
int relax_model_set_prestress(relax_model_t* m, const relax_tensor_layer_t* prestress, int n);
/* friction planes (in%n), each with its px2*px3 patches (may be NULL) */
int relax_model_set_planes(relax_model_t* m, const relax_plane_t* planes, int np,
                           const relax_slippatch_t* patches);
int relax_model_get_patches(relax_model_t* m, relax_slippatch_t* patches);
/* State in the reference's layout: u1..u3 and gamma are sx3/2 planes of
 * rows of sx1+2 (align_data=1) or sx1 (align_data=0) floats; tau and sig are
 * TENSOR fields (sx1, sx2, sx3/2).  sig may be NULL on upload: it is then
 * computed as -tau + C:eps(u) (relax.f90:369-371). */
int relax_model_upload(relax_model_t* m, const float* u1, const float* u2, const float* u3,
                       const relax_tensor_t* tau, const relax_tensor_t* sig, const float* gamma, double t,
                       int oi, int i, int align_data);
int relax_model_download(relax_model_t* m, float* u1, float* u2, float* u3, relax_tensor_t* tau,
                         relax_tensor_t* sig, float* gamma, double* t, int align_data);
/* One RK2 predictor/corrector step (relax.f90:555-890).  dt_override
 * replaces the adaptive step when not NULL (replay of a Dt sequence).
 * info may be NULL.  The step is asynchronous on the model's stream unless
 * info is requested. */
int relax_model_step(relax_model_t* m, const double* dt_override, relax_step_info_t* info);
/* Run up to nsteps steps, until t >= interval, or through the step that
 * reaches the next event time (info.e advances: the caller applies the
 * event with relax_model_event and runs on); info holds nsteps entries or
 * is NULL.  *done receives the number of steps taken. */
int relax_model_run(relax_model_t* m, int nsteps, const double* dt_override, relax_step_info_t* info, int* done);
int relax_model_synchronize(relax_model_t* m);
/* A coseismic event: the first one at t = 0 on a new model (relax.f90:
 * 316-371) or a later one once a step has reached its time (info.e, the
 * event switch of relax.f90:858-889).  As dislocations (relax.f90:
 * 1116-1289) and traction (elastic3d.f90:2517-2590) followed by the Green
 * solve: f, t = the source() body force and traction of the shear patches
 * and the tensilesource() body force of the tensile cracks longer or wider
 * than nyquist*min(dx1,dx2,dx3), the mogisource() body force of every point
 * dilatation source, the cuboidSource() and tetrahedronSource() body
 * forces of the eigenstrain volumes (elastic3d.f90:1680-1884, :1916-2039),
 * minus the surface loads at the model's time; u +=
 * greenfunctioncowling(f, t); tau += the moment density of every source
 * (momentDensityCuboid, momentDensityTetrahedron, elastic3d.f90:2927-3029,
 * :2788-2873, with the model's taper); sig = -tau + C:eps(u).  A
 * tetrahedron's vertex index outside 1..nv is RELAX_ERR_ARG.  Tensile cracks and point sources reach
 * above the surface (their images), so such an event takes the full-depth
 * Green solve on temporary buffers; shear patches and loads alone take the
 * step's pruned solve.  Sources below the Nyquist size get no displacement
 * (the reference adds Okada's solution for shear patches at t = 0 only,
 * relax.f90:357-360).  The event's surface loads stay with the model: the
 * following steps add the rate of the periodic ones to the surface traction
 * in both stages (relax.f90:684-685, :805-806) until the next event.  Not
 * on a member of a batch. */
int relax_model_event(relax_model_t* m, const relax_event_t* ev, double nyquist);
/* Inter-seismic loading (in%inter: relax.f90:669-674, :795-799, with
 * dislocations' eigenstress form :1116-1289): ev->s are shear segments whose
 * slip is a rate (each with its own taper beta), ev->ts tensile segments
 * whose opening is a rate (the model's taper); ev->m and ev->l must be
 * empty, as the deck has no loading of those kinds (input.f90:1114-1250).
 * Every later corrector adds their moment density m to the moment of the
 * rheology and to tau, unscaled and times Dt (tau += m + Dt (moment + m)),
 * and a step where nothing relaxes takes tm = interval/20.  The model keeps
 * m (24 bytes per point, beyond relax_model_footprint) and steps with the
 * extended stencil kernels.  NULL or no segments removes the loading.  Not
 * on a member of a batch. */
int relax_model_set_interseismic(relax_model_t* m, const relax_event_t* ev);
/* dislocations_disp (relax.f90:357-360, :1300-1349) after the first event:
 * Okada's (1992) closed-form displacement (okada/dc3d.f through
 * okada/green_space.f90) of the shear patches sc[0..ns) and tensile cracks
 * tsc[0..nt) at or below nyquist*min(dx1,dx2,dx3) in length or width,
 * added to u, then sig = -tau + C:eps(u).  The sources are the deck's
 * copies (in%events(1)%sc and %tsc, input.f90:1319): the upper corner at the
 * start of the strike, north-east-down, strike, dip and rake in degrees and
 * the slip as the deck gives them; the reference evaluates them in the grid's
 * frame without the deck's origin shift or rotation.  A grid point on a
 * source's edge gets no displacement from it (dc3d.f's singular case). */
int relax_model_okada(relax_model_t* m, const relax_source_t* sc, int ns, const relax_source_t* tsc, int nt,
                      double nyquist);
/* Utilities of the same set-up: add_shear adds momentdensityshear
 * (elastic3d.f90:2661-2750) of n rectangular patches (centre x,y,z, width,
 * length, strike, dip, rake in the internal convention) with slip[k] to tau
 * (taper beta of the model); coseismic then sets u = greenfunctioncowling(
 * equivalentbodyforce(tau)) and sig = -tau + C:eps(u), the elastic state of
 * a given tau with the FIR divergence as body force. */
int relax_model_add_shear(relax_model_t* m, const relax_plane_t* patches, const double* slip, int n);
int relax_model_coseismic(relax_model_t* m);
/* Phase timers (the HIP-event counterpart of the reference's PAPI phase
 * timers, papi_prof.c): runs nsteps steps with an event after every launch
 * and returns, per kernel kind, the summed milliseconds and the number of
 * launches (arrays of RELAX_PROF_KINDS entries). */
#define RELAX_PROF_BODYFORCE 0    /* rheology + divergence stencil */
#define RELAX_PROF_FFT_X_FWD 1    /* x pass, R2C */
#define RELAX_PROF_FFT_Y 2        /* x2 pass, both directions */
#define RELAX_PROF_GREEN_Z 3      /* x3 pass + Green + traction + Cerruti */
#define RELAX_PROF_FFT_X_INV 4    /* x pass, C2R + stage update */
#define RELAX_PROF_STRESS_PRED 5  /* strain stencil + predictor update */
#define RELAX_PROF_STRESS_CORR 6  /* strain stencil + corrector update */
#define RELAX_PROF_SMALL 7        /* traction planes, Nyquist planes, scalars */
#define RELAX_PROF_KINDS 8
int relax_model_profile(relax_model_t* m, int nsteps, double* ms, int* launches);
/* The surface plane (x3 = 0) of u1..u3 only, each sx2 rows of sx1+2
 * (align_data=1) or sx1 floats: the slice the reference's exportgrd writes
 * at each output step for observation depth 0 (relax.f90:955,
 * export.f90:917). */
int relax_model_surface(relax_model_t* m, float* u1, float* u2, float* u3, int align_data);
/* Observation points (exportpoints, util.f90:59-139): u and sig at the
 * shiftedindex (elastic3d.f90:4063-4101) of each point of x (n x 3, x1 x2 x3
 * in the computational frame, over the half grid as relax.f90:894 passes
 * it).  u (n x 3) is turned into the user frame by rotation(v1, v2, -rot),
 * rot in degrees (input.f90:1809-1822); sig (n, may be NULL) is returned as
 * it is.  RELAX_ERR_RANGE for a point outside the grid, where the reference
 * stops. */
int relax_model_points(relax_model_t* m, const double* x, int n, double rot, double* u, relax_tensor_t* sig);
/* Autotuner: times the launch-shape variants of the tunable kernels (columns
 * per workgroup of the y transform, columns/threads/radix of the Green
 * column pass, divergence tile and depth chunk) on this model's scratch buffers and keeps the
 * fastest for this device and grid; later models of the grid in the process
 * use the choice.  choice (3 ints, may be NULL) receives the variants;
 * ms (may be NULL) the average time of one launch of each variant in
 * milliseconds, 3 rows of RELAX_TUNE_VARIANTS (columns, Green, divergence),
 * with -1 for a variant that does not apply to the grid.  The model's state
 * is not changed.  Not on a member of a batch. */
#define RELAX_TUNE_VARIANTS 6
int relax_model_autotune(relax_model_t* m, int* choice, float* ms);
/* Sets the three variants (as relax_model_autotune reports them) without
 * timing, e.g. a choice kept from an earlier run; RELAX_ERR_ARG when out of
 * range.  A variant that does not fit the grid makes the step return
 * RELAX_ERR_SIZE. */
int relax_model_set_tuning(relax_model_t* m, const int* choice);
/* The model's HIP stream (hipStream_t) for callers that order their own work. */
void* relax_model_stream(relax_model_t* m);

/* ---- batches: many models stepped by shared launches --------------------- */
/* A batch steps all of its models together, and the step is replayed from a
 * captured hipGraph while the membership is unchanged (batches of 2^25 grid
 * points or more run their launches eagerly: there a graph saves nothing).
 * Members of one grid class (grid sx1, sx2, sx3, spacing dx1, dx2, dx3 and
 * lambda, mu, gamma) share every launch of the step (the model number is
 * part of the launch grid).  A batch may hold several classes (uneven
 * grids): the divergence and strain passes are then one launch over all
 * members that walks a work list of (model, tile) items, the transforms and
 * Green passes one launch per class.  Members share the device; rheology,
 * prestress, fault planes, time line and state are per model.  There is no
 * counterpart in the reference, whose driver runs one model per process
 * (relax.f90:185-1454); each member's step is the reference's RK2 step of
 * that model. */
typedef struct relax_batch relax_batch_t;
int relax_batch_create(relax_batch_t** out);
/* Members leave the batch when it is destroyed; the models themselves stay. */
int relax_batch_destroy(relax_batch_t* b);
/* RELAX_ERR_ARG if the model is already in a batch or on another device. */
int relax_batch_add(relax_batch_t* b, relax_model_t* m);
int relax_batch_remove(relax_batch_t* b, relax_model_t* m);
int relax_batch_size(const relax_batch_t* b);
/* One RK2 step of every member.  infos (NULL, or one entry per member in the
 * order of addition) receives each member's scalars; the call then waits for
 * the step and returns the first non-zero member status.  Without infos it
 * returns after enqueueing.  While a model is a member, its own step and run
 * calls are refused (RELAX_ERR_ARG); its download, get_patches and
 * synchronize calls first wait for the batch. */
int relax_batch_step(relax_batch_t* b, relax_step_info_t* infos);
int relax_batch_synchronize(relax_batch_t* b);
/* 1 while the batch still has queued work, 0 when it is idle (no waiting). */
int relax_batch_busy(relax_batch_t* b);
/* Each member's current scalars (waits for the queued work). */
int relax_batch_scalars(relax_batch_t* b, relax_step_info_t* infos);
/* Autotuner of a batch: as relax_model_autotune, but each timing is one
 * launch over all current members (their scratch buffers), so the choice
 * fits the batch's size; the batch keeps it, and later batches of the same
 * device, grid and rheology class in the process start from it.  A batch
 * without a choice launches with its first member's shapes.  Members equal
 * their single runs bitwise when those use the same shapes.  RELAX_ERR_ARG
 * for a batch of several grid classes: tune each class as a batch of its
 * own, and a mixed batch takes those choices per class.  No counterpart in
 * the reference (it has no batches). */
int relax_batch_autotune(relax_batch_t* b, int* choice, float* ms);

#ifdef __cplusplus
}
#endif
#endif
