// kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once
#include "common.h"
#include "modeldev.h"

enum { EPI_PLAIN = 0, EPI_PRED = 1, EPI_CORR = 2, EPI_ADD = 3 };

const float2* twiddle_table(int L);

// fft.hip -- each launch covers nm models (grid y); the MPtr arguments give
// each model's buffer
// ys_n2 > 0 splits the following y transform of planes of ys_n2 rows: the x
// kernels do its first radix-8 stage across rows j1 + (ys_n2/8)*j2, and
// k_fft_cols of length ys_n2/8 over blocks of ys_n2/8 rows completes it,
// leaving y frequency 8*k1 + k2 at row (ys_n2/8)*k2 + k1 of a plane
// side: a second array of rows that rides along in the same launch (plain
// rows, no split y stage; nb is set by the launcher)
struct XSide {
  MPtr<const float> in;
  MPtr<float2> out, nyq;
  long in_pitch = 0, out_pitch = 0, nyq_stride = 0, nrows = 0;
  float scale = 0.f;
  int nb = 0;
};
int launch_fft_x_r2c(int n1, MPtr<const float> in, long in_pitch, MPtr<float2> out, long out_pitch,
                     MPtr<float2> nyq, long nyq_stride, long nrows, float scale, int nm, hipStream_t st,
                     int ys_n2 = 0, const XSide* side = nullptr);
int launch_fft_x_c2r(int n1, int epi, MPtr<const float2> in, long in_pitch, MPtr<const float2> nyq,
                     long nyq_stride, MPtr<float> out, long out_pitch, long nrows, float scale,
                     MPtr<const float> base, long base_pitch, MPtr<const StepScalars> sc, int nm, hipStream_t st,
                     int ys_n2 = 0);
// launch-shape variants of the tunable kernels (0 = default)
constexpr int COLS_VARIANTS = 3, GREEN_VARIANTS = 4, BODY_VARIANTS = 6;

int launch_fft_cols(int L, int sign, MPtr<float2> data, long jstride, long pstride, int ncols, int nplanes, int nm,
                    hipStream_t st, int variant = 0);
int launch_fft_rows(int L, int sign, MPtr<float2> data, long pitch, long nrows, int nm, hipStream_t st);
// The small transforms of a Green solve that ride along with its y pass:
// tnp planes of tnc columns of LS points (column j at j*tjs, plane at tps)
// and two sets of LS-point rows (pitch rp, rn rows).  Zero counts leave a
// job out.
struct YRide {
  MPtr<float2> tc;
  long tjs = 0, tps = 0;
  int tnc = 0, tnp = 0;
  MPtr<float2> r0, r1;
  long rp0 = 0, rn0 = 0, rp1 = 0, rn1 = 0;
};
// the launch's copy: workgroups of each job and the LS-point twiddles
struct YRideDev {
  YRide r;
  const float2* tw;
  int nbt, nb0, nb1, nside;
};
// The y pass: L-point columns of data (as launch_fft_cols) and the ride's
// LS-point transforms in one launch.
int launch_fft_ypass(int L, int LS, int sign, MPtr<float2> data, long jstride, long pstride, int ncols, int nplanes,
                     const YRide& ride, int nm, hipStream_t st, int variant = 0);
// one buffer
inline int launch_fft_x_r2c(int n1, const float* in, long in_pitch, float2* out, long out_pitch, float2* nyq,
                            long nyq_stride, long nrows, float scale, hipStream_t st) {
  return launch_fft_x_r2c(n1, mptr(in), in_pitch, mptr(out), out_pitch, mptr(nyq), nyq_stride, nrows, scale, 1, st);
}
inline int launch_fft_x_c2r(int n1, int epi, const float2* in, long in_pitch, const float2* nyq, long nyq_stride,
                            float* out, long out_pitch, long nrows, float scale, const float* base, long base_pitch,
                            const StepScalars* sc, hipStream_t st) {
  return launch_fft_x_c2r(n1, epi, mptr(in), in_pitch, mptr(nyq), nyq_stride, mptr(out), out_pitch, nrows, scale,
                          mptr(base), base_pitch, mptr(sc), 1, st);
}
inline int launch_fft_cols(int L, int sign, float2* data, long jstride, long pstride, int ncols, int nplanes,
                           hipStream_t st) {
  return launch_fft_cols(L, sign, mptr(data), jstride, pstride, ncols, nplanes, 1, st);
}
inline int launch_fft_rows(int L, int sign, float2* data, long pitch, long nrows, hipStream_t st) {
  return launch_fft_rows(L, sign, mptr(data), pitch, nrows, 1, st);
}

// green.hip
struct GreenConst {
  double lambda, mu, gamma;  // Lame parameters and Cowling coefficient
  double dx1, dx2, dx3;
  int n1, n2, n3;
};
// fused pruned column pass: spectra of the half grid (x and y already
// transformed) -> Green + surface traction + Cerruti -> back to x3 space
// l2 > 0: the spectrum rows of W are in the split-y order (row l2*k2 + k1
// holds y frequency 8*k1 + k2); the traction planes TM are in natural order
int launch_green_columns(const GreenConst& gc, int h3, MPtr<float2> W, MPtr<float2> NQ, MPtr<const float2> TM,
                         MPtr<const float2> TN, int nm, hipStream_t st, int l2 = 0, int variant = 0);
// per-routine kernels on padded full spectra (row pitch n1/2+1 complex)
int launch_elastic_padded(const GreenConst& gc, float2* f1, float2* f2, float2* f3, hipStream_t st);
int launch_surface_padded(const GreenConst& gc, const float2* u1, const float2* u2, const float2* u3,
                          const float2* t1, const float2* t2, const float2* t3, float2* p1, float2* p2,
                          float2* p3, hipStream_t st);
int launch_cerruti_padded(const GreenConst& gc, const float2* p1, const float2* p2, const float2* p3,
                          float2* u1, float2* u2, float2* u3, hipStream_t st);

// model.hip: the body force of source() (elastic3d.f90:2064-2252) of the
// np planes P (slips slip) added to the top half of f (3 planes of
// n1*n2*h3, packed rows of ld floats, plane stride fs) and their surface
// traction to t (3 planes of n1*n2, row stride ld, plane stride ts).
struct PlaneDev;
int launch_source(const PlaneDev* P, const double* slip, int np, double mu, int n1, int n2, int h3, double dx1,
                  double dx2, double dx3, float* f, long ld, long fs, float* t, long ts, hipStream_t st);

// model.hip: traction() of the surface loads l subtracted from t3 (n1*n2,
// row stride ld) at time t, or their rate over [t, t+Dt] when rate != 0
int launch_traction(const relax_source_t* l, int nl, int n1, int n2, double dx1, double dx2, double t, double Dt,
                    int rate, float* t3, long ld, hipStream_t st);

// compat.hip: greenfunctioncowling over the full depth on device buffers in
// the padded layout (v: fields of (sx1+2) sx2 sx3, t and p: planes of
// (sx1+2) sx2), enqueued on stream st (and waited for with sync)
int green_full_device(float* const v[3], float* const t[3], float* const p[3], int sx1, int sx2, int sx3,
                      double dx1, double dx2, double dx3, double lambda, double mu, double gamma, hipStream_t st,
                      bool sync);

// model.hip: tensilesource (elastic3d.f90:2315-2411) and mogisource
// (:2458-2509) of np cracks / point sources into f (3 fields of n1 n2 n3
// floats in rows of ld, component stride fs) over the full depth, and their
// moment densities (:3137-3215, :3240-3286) into a tensor field of the
// top half (6 planes of H)
int launch_tensile_source(const PlaneDev* P, const double* opening, int np, double lambda, double mu, int n1, int n2,
                          int n3, double dx1, double dx2, double dx3, float* f, long ld, long fs, hipStream_t st);
int launch_mogi_source(const relax_source_t* M, int nm, double lambda, double mu, int n1, int n2, int n3,
                       double dx1, double dx2, double dx3, float* f, long ld, long fs, hipStream_t st);
int launch_add_tensile(const PlaneDev* P, const double* opening, int np, double lambda, double mu, int n1, int n2,
                       int h3, double dx1, double dx2, double dx3, float* sig, long H, hipStream_t st);
int launch_add_mogi(const relax_source_t* M, int nm, double lambda, double mu, int n1, int n2, int h3, double dx1,
                    double dx2, double dx3, float* sig, long H, hipStream_t st);

// compat.hip: the weak-zone fluidity field of builddgammadot0
// (viscoelastic3d.f90:320-435) on the device.  A zone in device form holds
// what depends on the zone only, computed on the host in double as the
// reference does (its trigonometry from the host libm), so the kernel is
// left with additions, products and quotients in the reference's order.
struct WeakDev {
  double dg, x, y, z;       // dgammadot0 and centre
  double W, L, D, LM;       // x2, x3, x1 extents and the gate half-width
  double cs, ss, cd, sd;    // cos/sin of strike and dip
  double xr, yr, zr;        // rotated centre
};
WeakDev make_weak(const relax_weak_t& z, double beta);
int launch_builddgammadot0(int n1, int n2, int n3, double dx1, double dx2, double dx3, double beta, int nz,
                           const WeakDev* zones, float* dg, hipStream_t st);
