// common.h -- shared definitions of the gfx950 implementation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/relax_hip.h"

#define RELAX_PI 3.141592653589793115997963468544185161
#define RELAX_PI2 6.28318530717958623199592693708837032318
#define RELAX_DEG2RAD 0.01745329251994329547437168059786927  // DEG2RAD of input.f90:28

// Return a status from a HIP call.
#define HIP_TRY(expr)                                     \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) return relax_hip_error(e_, #expr); \
  } while (0)

int relax_hip_error(hipError_t e, const char* what);

// Memory management and other device-wide calls, serialised with the graph
// capture of batches (relax_batch_step).  Freeing memory or destroying a
// stream synchronises the device, and a synchronisation that reaches a stream
// another thread is capturing breaks that capture; every such call of the
// library takes this process-wide (recursive) lock, as does the capture.
void capture_lock();
void capture_unlock();
struct CaptureGuard {
  CaptureGuard() { capture_lock(); }
  ~CaptureGuard() { capture_unlock(); }
  CaptureGuard(const CaptureGuard&) = delete;
  CaptureGuard& operator=(const CaptureGuard&) = delete;
};
template <class T>
inline hipError_t rx_malloc(T** p, size_t bytes) {
  CaptureGuard g;
  return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
template <class T>
inline hipError_t rx_host_malloc(T** p, size_t bytes) {
  CaptureGuard g;
  return hipHostMalloc(reinterpret_cast<void**>(p), bytes, hipHostMallocDefault);
}
inline hipError_t rx_free(void* p) {
  CaptureGuard g;
  return hipFree(p);
}
inline hipError_t rx_host_free(void* p) {
  CaptureGuard g;
  return hipHostFree(p);
}
inline hipError_t rx_device_sync() {
  CaptureGuard g;
  return hipDeviceSynchronize();
}
inline hipError_t rx_stream_create(hipStream_t* s) {
  CaptureGuard g;
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}
inline hipError_t rx_stream_destroy(hipStream_t s) {
  CaptureGuard g;
  return hipStreamDestroy(s);
}

// ----------------------------------------------------------------- complex --
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by +i or -i
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// ------------------------------------------------------ workgroup order ----
// The dispatcher hands workgroup b to XCD b % 8.  Tiles that share cache
// lines or halo rows should meet in one XCD's L2, so workgroup b takes tile
// (b % 8) * q + b / 8: each XCD walks a contiguous range of q tiles (the
// last nb % 8 tiles keep their own index).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int NX = 8;
  const int q = nb / NX;
  if (b >= q * NX) return b;
  return (b % NX) * q + b / NX;
}

// --------------------------------------------------------------- grid info --
// Wavenumber in cycles per unit length (fourier.f90:64-81), 0-based index.
__host__ __device__ __forceinline__ float wavenumber23f(int i, int n, float inv_ndx) {
  return (float)(i < n / 2 ? i : i - n) * inv_ndx;
}

// Per-model constant block passed by value to the kernels.
struct Grid {
  int n1, n2, n3, h3;  // full FFT grid and the depth of the half grid
  int m1;              // n1/2: complex samples of an x row without the Nyquist bin
  long H;              // n1*n2*h3 points of the half grid
  float dx1, dx2, dx3;
};

// Tensor fields are stored as six planes (SoA) of H floats:
// s11, s12, s13, s22, s23, s33.
struct T6 {
  float s11, s12, s13, s22, s23, s33;
};
// Accesses through the global address space.  Pointers read from a record in
// device memory (a batch member's fields) are generic, and the compiler would
// emit flat accesses for them, which also wait on the LDS counter.
typedef __attribute__((address_space(1))) float relax_gfloat;
__device__ __forceinline__ float gld(const float* p) { return *(const relax_gfloat*)p; }
__device__ __forceinline__ void gst(float* p, float v) { *(relax_gfloat*)p = v; }
__device__ __forceinline__ float gld_nt(const float* p) { return __builtin_nontemporal_load((const relax_gfloat*)p); }
__device__ __forceinline__ void gst_nt(float* p, float v) { __builtin_nontemporal_store(v, (relax_gfloat*)p); }

__device__ __forceinline__ T6 load_t6(const float* __restrict__ t, long H, long o) {
  return T6{gld(t + o), gld(t + H + o), gld(t + 2 * H + o), gld(t + 3 * H + o), gld(t + 4 * H + o),
            gld(t + 5 * H + o)};
}
__device__ __forceinline__ void store_t6(float* __restrict__ t, long H, long o, const T6& v) {
  gst(t + o, v.s11); gst(t + H + o, v.s12); gst(t + 2 * H + o, v.s13);
  gst(t + 3 * H + o, v.s22); gst(t + 4 * H + o, v.s23); gst(t + 5 * H + o, v.s33);
}
__device__ __forceinline__ T6 load_t6_nt(const float* __restrict__ t, long H, long o) {
  return T6{gld_nt(t + o), gld_nt(t + H + o), gld_nt(t + 2 * H + o), gld_nt(t + 3 * H + o),
            gld_nt(t + 4 * H + o), gld_nt(t + 5 * H + o)};
}
__device__ __forceinline__ void store_t6_nt(float* __restrict__ t, long H, long o, const T6& v) {
  gst_nt(t + o, v.s11); gst_nt(t + H + o, v.s12); gst_nt(t + 2 * H + o, v.s13);
  gst_nt(t + 3 * H + o, v.s22); gst_nt(t + 4 * H + o, v.s23); gst_nt(t + 5 * H + o, v.s33);
}

// Device tables of the rheology, one entry per depth of the half grid.
struct Mechanism {
  const float* gammadot0;  // h3
  const float* power;      // h3 (stress exponent n)
  const float* cohesion;   // h3 (viscous); Gk (transient)
  const float* gdp;        // h3: gammadot0 (taup/mu)^(n-1), the prestress rate
  const float* rp;         // h3: (taup/mu)^(n-1), for points with a weak zone
  const float* dg;         // H: weak-zone fluidity perturbation, or NULL
  int active;
};

struct Prestress {
  const float* sp;    // 6*h3 deviatoric prestress per depth (AoS per depth)
  const float* taup;  // h3 norm of the deviatoric prestress
};

// Friction planes in device form (friction3d.f90:82-241); angles in radians.
struct PlaneDev {
  double x, y, z, L, W, beta;
  double cstrike, sstrike, cdip, sdip, rake, dip;
  double Wp, Lp, xr, yr, zr;
  double n[3], r[3];
  double minlw;  // min(L,W)/sqrt(dx1*dx3) of the relaxation-time estimate
};
struct FrictionTables {
  // double precision, as LAYER_STRUCT (types.f90:31-34)
  const double* v0;        // h3 (gammadot0 of the friction structure)
  const double* tauc;      // h3 ((a-b) sigma, stored in stressexponent)
  const double* friction;  // h3
  const double* cohesion;  // h3
  const PlaneDev* planes;
  int np;
  int active;
};

// Device-side scalars of one model's time line.
// Relaxation mechanisms of the time step (relax.f90:565-660): linear and
// power-law viscosity, fault creep, linear and power-law transient creep.
constexpr int NMECH = 5;
struct StepScalars {
  double t, Dt, tm;
  double maxwell[NMECH];
  double power, cei;
  int oi, it, status, e;          // e: index of the last event (0 at t = 0)
  unsigned int tmin_bits[NMECH];  // running minima (positive floats as ordered ints)
  unsigned int pad2[3];
};
