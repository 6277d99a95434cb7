// modeldev.h -- the per-model device record that the step kernels read, and
// the addressing of a launch that covers one model or a batch of models of
// the same grid.
//
// A launch over nm models carries the model number in its grid (blockIdx.y,
// or folded into blockIdx.z for the 3-D stencil grids) and finds model m's
// buffers and rheology in entry m of a device array of ModelDev records.  A
// single model is a batch of one: its own one-entry array.
#pragma once
#include <cstddef>

#include "rheology.h"

// Time line of one model (relax.f90:659-677 step size, :1364-1376 isoutput,
// :1420-1452 integrationstep).
struct TimeLine {
  double interval, odt, tscale;
  const double* ev;  // device array of the ne event times (event 0 at t = 0)
  int ne;
  int skip;
  int mech[NMECH];
  int inter;  // inter-seismic loading: tm = interval/20 when nothing relaxes
};

struct ModelDev {
  float* u;          // 3 planes of H
  float* sig;        // 6 planes of H
  float* tau;        // 6 planes of H
  float* gam;        // H
  float* E;          // 6 planes of H: transient strain epsilonik (or NULL)
  float* Em;         // 6 planes of H: its mid-step value (predictor output)
  float* W;          // 3H work field / spectrum
  float2* NQ;        // Nyquist bins of the work spectrum
  float* T;          // surface traction planes
  float2* TN;        // their Nyquist bins
  StepScalars* sc;   // t, Dt, tm, ...
  double2* partials; // amplitude partials (two levels)
  double dv;         // dx1 dx2 dx3
  TimeLine tl;
  RheoParams R;
};

// One pointer argument of a launch: the pointer itself (one buffer), or the
// field of every record of a ModelDev array (entry m for model m).
template <class T>
struct MPtr {
  T* p = nullptr;
  const char* rec = nullptr;  // address of the field in record 0
  int stride = 0;             // sizeof(ModelDev)
  __device__ __forceinline__ T* at(int m) const {
    return rec ? *reinterpret_cast<T* const*>(rec + (long)m * stride) : p;
  }
};

// Loads and stores through the global address space: pointers taken from
// records in memory are otherwise generic, and the compiler emits flat
// accesses, which also count against the LDS wait counter.
typedef float relax_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gload(const float2* p) {
  const relax_f2v v = *(const __attribute__((address_space(1))) relax_f2v*)p;
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void gstore(float2* p, float2 v) {
  relax_f2v w;
  w.x = v.x;
  w.y = v.y;
  *(__attribute__((address_space(1))) relax_f2v*)p = w;
}

template <class T>
inline MPtr<T> mptr(T* p) {
  MPtr<T> r;
  r.p = p;
  return r;
}

// MFIELD(T, md, field): the field of every record of the device array md.
#define MFIELD(T, md, field) \
  mfield<T>(reinterpret_cast<const char*>(md) + offsetof(ModelDev, field))
template <class T>
inline MPtr<T> mfield(const char* rec) {
  MPtr<T> r;
  r.rec = rec;
  r.stride = (int)sizeof(ModelDev);
  return r;
}
