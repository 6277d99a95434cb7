// kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once
#include "common.h"

enum { EPI_PLAIN = 0, EPI_PRED = 1, EPI_CORR = 2 };

const float2* twiddle_table(int L);

// fft.hip
int launch_fft_x_r2c(int n1, const float* in, long in_pitch, float2* out, long out_pitch, float2* nyq,
                     long nyq_stride, long nrows, float scale, hipStream_t st);
int launch_fft_x_c2r(int n1, int epi, const float2* in, long in_pitch, const float2* nyq, long nyq_stride,
                     float* out, long out_pitch, long nrows, float scale, const float* base, long base_pitch,
                     const StepScalars* sc, hipStream_t st);
int launch_fft_cols(int L, int sign, float2* data, long jstride, long pstride, int ncols, int nplanes,
                    hipStream_t st);
int launch_fft_rows(int L, int sign, float2* data, long pitch, long nrows, hipStream_t st);

// green.hip
struct GreenConst {
  double lambda, mu, gamma;  // Lame parameters and Cowling coefficient
  double dx1, dx2, dx3;
  int n1, n2, n3;
};
// fused pruned column pass: spectra of the half grid (x and y already
// transformed) -> Green + surface traction + Cerruti -> back to x3 space
int launch_green_columns(const GreenConst& gc, int h3, float2* W, float2* NQ, const float2* TM, const float2* TN,
                         hipStream_t st);
// per-routine kernels on padded full spectra (row pitch n1/2+1 complex)
int launch_elastic_padded(const GreenConst& gc, float2* f1, float2* f2, float2* f3, hipStream_t st);
int launch_surface_padded(const GreenConst& gc, const float2* u1, const float2* u2, const float2* u3,
                          const float2* t1, const float2* t2, const float2* t3, float2* p1, float2* p2,
                          float2* p3, hipStream_t st);
int launch_cerruti_padded(const GreenConst& gc, const float2* p1, const float2* p2, const float2* p3,
                          float2* u1, float2* u2, float2* u3, hipStream_t st);
