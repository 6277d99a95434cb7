// compat.hip -- per-routine entry points with host pointers: one per reference
// subroutine of the hot path, same argument order and meaning.  Each call
// copies its arrays to the GPU, runs the HIP kernels and copies the results
// back; the device-resident model (model.hip) is the fast path.
//
// Stencils here take any FIR length chosen by optimalfilter
// (elastic3d.f90:846-907): fir1 for 15+ samples, fir7 for 5..14.
#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kernels.h"
#include "model.h"
#include "rheology.h"

// ------------------------------------------------------------ library ----
// A failed HIP call becomes a status; with RELAX_HIP_VERBOSE set in the
// environment the call and the runtime's error name go to stderr as well.
int relax_hip_error(hipError_t e, const char* what) {
  static const bool verbose = std::getenv("RELAX_HIP_VERBOSE") != nullptr;
  if (verbose) std::fprintf(stderr, "relax-hip: %s failed: %s\n", what, hipGetErrorName(e));
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return RELAX_ERR_NOMEM;
  return RELAX_ERR_HIP;
}

const char* relax_hip_version(void) { return "relax-hip 0.1 (gfx950)"; }

const char* relax_hip_status_string(int s) {
  switch (s) {
    case RELAX_OK: return "ok";
    case RELAX_ERR_POWER: return "invalid power exponent (stress exponent < 1)";
    case RELAX_ERR_FILTER: return "dimension too small for a FIR differentiator";
    case RELAX_ERR_FRICTION: return "wrong value of (a-b)sigma: tau/tauc > 20";
    case RELAX_ERR_ARG: return "invalid argument";
    case RELAX_ERR_SIZE: return "grid size not supported";
    case RELAX_ERR_NOMEM: return "device allocation failed";
    case RELAX_ERR_HIP: return "HIP runtime error";
    case RELAX_ERR_ADMISSION: return "model does not fit the device memory";
    case RELAX_ERR_RANGE: return "patch centre outside the stress field";
    case RELAX_ERR_NODEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

int relax_hip_meminfo(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return RELAX_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return RELAX_ERR_NODEVICE;
  HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
  return RELAX_OK;
}

int noalloc_setup();

int relax_hip_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return RELAX_ERR_NODEVICE;
  if (device < 0 || device >= n) return RELAX_ERR_ARG;
  HIP_TRY(hipSetDevice(device));
  return noalloc_setup();
}

namespace {

// device buffer owned for the duration of one call
struct Buf {
  void* p = nullptr;
  ~Buf() { if (p) (void)rx_free(p); }
  int alloc(size_t bytes) { return rx_malloc(&p, bytes ? bytes : 16) == hipSuccess ? RELAX_OK : RELAX_ERR_NOMEM; }
  float* f() const { return (float*)p; }
  float2* c() const { return (float2*)p; }
};

bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

int has_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return RELAX_ERR_NODEVICE;
  return RELAX_OK;
}

dim3 g1d(long n, int bs = 256) {
  long b = (n + bs - 1) / bs;
  return dim3((unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b)));
}

// optimalfilter, elastic3d.f90:846-907 (taps divided by the spacing)
struct Taps {
  double k[3][16];
  int len[3];
};
int make_taps(int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, Taps& T) {
  static const float fir7f[7] = {8.77856e-01f, -2.81913e-01f, +6.22696e-02f, +2.82441e-02f,
                                 -5.09029e-02f, +4.20471e-02f, -1.59409e-02f};
  const int sx[3] = {sx1, sx2, sx3};
  const double dx[3] = {dx1, dx2, dx3};
  std::memset(&T, 0, sizeof(T));
  for (int a = 0; a < 3; ++a) {
    const int n = sx[a];
    if (a < 2 && n >= 2 && n <= 4) { T.len[a] = 1; T.k[a][0] = 0.5 / dx[a]; continue; }
    if (n >= 5 && n <= 14) { T.len[a] = 7; for (int l = 0; l < 7; ++l) T.k[a][l] = (double)fir7f[l] / dx[a]; continue; }
    if (n >= 15) { T.len[a] = 1; T.k[a][0] = 0.5 / dx[a]; continue; }
    return RELAX_ERR_FILTER;
  }
  return RELAX_OK;
}

// ---------------------------------------------- generic-tap stencils ----
__global__ void k_stressupdate_generic(const float* __restrict__ v1, const float* __restrict__ v2,
                                       const float* __restrict__ v3, long vp, int n1, int n2, int n3, Taps T,
                                       double dx3, double lambda, double mu, float* __restrict__ sig) {
  const long N = (long)n1 * n2 * n3;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < N; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
    auto V = [&](const float* v, int a, int b, int c) { return v[a + vp * (b + (long)n2 * c)]; };
    const int len3 = T.len[2];
    const bool boundary = !((i3 + 1 > len3) && (i3 + 1 < n3 - len3 + 1));
    float e11 = 0, e12 = 0, e13 = 0, e22 = 0, e23 = 0, e33 = 0;
    for (int l = 1; l <= T.len[0]; ++l) {
      const int m = (n1 + i1 - l) % n1, p = (i1 + l) % n1;
      e11 = (float)(e11 + (double)(V(v1, p, i2, i3) - V(v1, m, i2, i3)) * T.k[0][l - 1]);
      e12 = (float)(e12 + (double)(V(v2, p, i2, i3) - V(v2, m, i2, i3)) * T.k[0][l - 1]);
      e13 = (float)(e13 + (double)(V(v3, p, i2, i3) - V(v3, m, i2, i3)) * T.k[0][l - 1]);
    }
    for (int l = 1; l <= T.len[1]; ++l) {
      const int m = (n2 + i2 - l) % n2, p = (i2 + l) % n2;
      e12 = (float)(e12 + (double)(V(v1, i1, p, i3) - V(v1, i1, m, i3)) * T.k[1][l - 1]);
      e22 = (float)(e22 + (double)(V(v2, i1, p, i3) - V(v2, i1, m, i3)) * T.k[1][l - 1]);
      e23 = (float)(e23 + (double)(V(v3, i1, p, i3) - V(v3, i1, m, i3)) * T.k[1][l - 1]);
    }
    if (boundary) {
      int zp, zm;
      double px3;
      if (i3 == 0) { px3 = dx3; zp = 1; zm = 0; }
      else if (i3 == n3 - 1) { px3 = dx3; zp = n3 - 1; zm = n3 - 2; }
      else { px3 = 2.0 * dx3; zp = i3 + 1; zm = i3 - 1; }
      e13 = (float)(e13 + (double)(V(v1, i1, i2, zp) - V(v1, i1, i2, zm)) / px3);
      e23 = (float)(e23 + (double)(V(v2, i1, i2, zp) - V(v2, i1, i2, zm)) / px3);
      e33 = (float)((double)(V(v3, i1, i2, zp) - V(v3, i1, i2, zm)) / px3);
    } else {
      for (int l = 1; l <= len3; ++l) {
        e13 = (float)(e13 + (double)(V(v1, i1, i2, i3 + l) - V(v1, i1, i2, i3 - l)) * T.k[2][l - 1]);
        e23 = (float)(e23 + (double)(V(v2, i1, i2, i3 + l) - V(v2, i1, i2, i3 - l)) * T.k[2][l - 1]);
        e33 = (float)(e33 + (double)(V(v3, i1, i2, i3 + l) - V(v3, i1, i2, i3 - l)) * T.k[2][l - 1]);
      }
    }
    e12 *= 0.5f; e13 *= 0.5f; e23 *= 0.5f;
    const double tr = (double)(e11 + e22 + e33);
    const float m2 = (float)(2.0 * mu);
    float* s = sig + 6 * o;
    s[0] += (float)(e11 * m2 + lambda * tr);
    s[1] += e12 * m2;
    s[2] += e13 * m2;
    s[3] += (float)(e22 * m2 + lambda * tr);
    s[4] += e23 * m2;
    s[5] += (float)(e33 * m2 + lambda * tr);
  }
}

__global__ void k_bodyforce_generic(const float* __restrict__ sig, int n1, int n2, int n3, Taps T, double dx3,
                                    float* c1, float* c2, float* c3, long cp, const float* __restrict__ mask) {
  const long N = (long)n1 * n2 * n3;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < N; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
    if (mask && mask[i3] == 0) continue;
    auto S = [&](int a, int b, int c, int k) { return sig[6 * (a + (long)n1 * (b + (long)n2 * c)) + k]; };
    const int len3 = T.len[2];
    const bool boundary = !((i3 + 1 > len3) && (i3 + 1 < n3 - len3 + 1));
    double f1 = 0, f2 = 0, f3 = 0;
    for (int l = 1; l <= T.len[0]; ++l) {
      const int m = (n1 + i1 - l) % n1, p = (i1 + l) % n1;
      f1 += (double)(S(p, i2, i3, 0) - S(m, i2, i3, 0)) * T.k[0][l - 1];
      f2 += (double)(S(p, i2, i3, 1) - S(m, i2, i3, 1)) * T.k[0][l - 1];
      f3 += (double)(S(p, i2, i3, 2) - S(m, i2, i3, 2)) * T.k[0][l - 1];
    }
    for (int l = 1; l <= T.len[1]; ++l) {
      const int m = (n2 + i2 - l) % n2, p = (i2 + l) % n2;
      f1 += (double)(S(i1, p, i3, 1) - S(i1, m, i3, 1)) * T.k[1][l - 1];
      f2 += (double)(S(i1, p, i3, 3) - S(i1, m, i3, 3)) * T.k[1][l - 1];
      f3 += (double)(S(i1, p, i3, 4) - S(i1, m, i3, 4)) * T.k[1][l - 1];
    }
    if (boundary) {
      int zp, zm;
      double px3;
      if (i3 == 0) { px3 = dx3; zp = 1; zm = 0; }
      else if (i3 == n3 - 1) { px3 = dx3; zp = n3 - 1; zm = n3 - 2; }
      else { px3 = 2.0 * dx3; zp = i3 + 1; zm = i3 - 1; }
      f1 += (double)(S(i1, i2, zp, 2) - S(i1, i2, zm, 2)) / px3;
      f2 += (double)(S(i1, i2, zp, 4) - S(i1, i2, zm, 4)) / px3;
      f3 += (double)(S(i1, i2, zp, 5) - S(i1, i2, zm, 5)) / px3;
    } else {
      for (int l = 1; l <= len3; ++l) {
        f1 += (double)(S(i1, i2, i3 + l, 2) - S(i1, i2, i3 - l, 2)) * T.k[2][l - 1];
        f2 += (double)(S(i1, i2, i3 + l, 4) - S(i1, i2, i3 - l, 4)) * T.k[2][l - 1];
        f3 += (double)(S(i1, i2, i3 + l, 5) - S(i1, i2, i3 - l, 5)) * T.k[2][l - 1];
      }
    }
    const long q = i1 + cp * (i2 + (long)n2 * i3);
    c1[q] = c1[q] - (float)f1;
    c2[q] = c2[q] - (float)f2;
    c3[q] = c3[q] - (float)f3;
  }
}

__global__ void k_traction_add(const float* __restrict__ sig, int n1, int n2, float* t1, float* t2, float* t3,
                               long tp) {
  const long n = (long)n1 * n2;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % n1, i2 = o / n1;
    const long q = i1 + tp * i2;
    t1[q] += sig[6 * o + 2];
    t2[q] += sig[6 * o + 4];
    t3[q] += sig[6 * o + 5];
  }
}

// ------------------------------------------------------------ rheology ----
__global__ void k_viscous_generic(Mechanism M, Prestress P, float mu, int n1, int n2, int n3,
                                  const float* __restrict__ sig, float* moment, const float* __restrict__ dg0,
                                  float* gamma, long gp, unsigned int* tmin_bits) {
  const long N = (long)n1 * n2 * n3;
  float tmin = __builtin_inff();
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < N; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
    float g0 = M.gammadot0[i3];
    if (dg0) g0 = g0 + dg0[o];
    if (1.0e-20 > (double)g0) continue;
    const float* s = sig + 6 * o;
    const T6 sd = deviatoric(T6{s[0], s[1], s[2], s[3], s[4], s[5]});
    const float* sp = P.sp + 6 * i3;
    const T6 ss{sd.s11 + sp[0], sd.s12 + sp[1], sd.s13 + sp[2], sd.s22 + sp[3], sd.s23 + sp[4], sd.s33 + sp[5]};
    const float tau = tnorm(ss);
    const float tauc = fmaxf(0.f, tau - M.cohesion[i3]);
    if ((double)tauc <= 1.0e-20) continue;
    const float ex = M.power[i3] - 1.f;
    const float gd = g0 * powf(tauc / mu, ex), gdp = g0 * powf(P.taup[i3] / mu, ex);
    const float a = 2.f * gd, b = 2.f * gdp;
    float* m = moment + 6 * o;
    m[0] = m[0] + (a * sd.s11 - b * sp[0]);
    m[1] = m[1] + (a * sd.s12 - b * sp[1]);
    m[2] = m[2] + (a * sd.s13 - b * sp[2]);
    m[3] = m[3] + (a * sd.s22 - b * sp[3]);
    m[4] = m[4] + (a * sd.s23 - b * sp[4]);
    m[5] = m[5] + (a * sd.s33 - b * sp[5]);
    tmin = fminf(tmin, 1.f / gd);
    if (gamma) gamma[i1 + gp * (i2 + (long)n2 * i3)] = gd * tauc / mu;
  }
  for (int off = 32; off > 0; off >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, off));
  if ((threadIdx.x & 63) == 0 && tmin < __builtin_inff()) atomicMin(tmin_bits, __float_as_uint(tmin));
}

// transienteigenstress, viscoelastic3d.f90:123-202.  E and Ed may be one
// buffer (the corrector's call): each point reads its strain before it
// writes its rate.
__global__ void k_transient_generic(Mechanism M, float mu, int n1, int n2, int n3, const float* __restrict__ sig,
                                    const float* E, float* moment, float* Ed, const float* __restrict__ dg0,
                                    unsigned int* tmin_bits) {
  const long N = (long)n1 * n2 * n3;
  float tmin = __builtin_inff();
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < N; o += (long)gridDim.x * blockDim.x) {
    const int i3 = o / ((long)n1 * n2);
    float g0 = M.gammadot0[i3];
    if (dg0) g0 = g0 + dg0[o];
    const float* s = sig + 6 * o;
    const float* e = E + 6 * o;
    const T6 ev{e[0], e[1], e[2], e[3], e[4], e[5]};
    T6 m{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, r;
    const float t = transient_point(M, mu, i3, g0, deviatoric(T6{s[0], s[1], s[2], s[3], s[4], s[5]}), ev, m, r);
    float* d = Ed + 6 * o;
    d[0] = r.s11; d[1] = r.s12; d[2] = r.s13; d[3] = r.s22; d[4] = r.s23; d[5] = r.s33;
    if (t < __builtin_inff()) {
      float* mo = moment + 6 * o;
      mo[0] = mo[0] + m.s11; mo[1] = mo[1] + m.s12; mo[2] = mo[2] + m.s13;
      mo[3] = mo[3] + m.s22; mo[4] = mo[4] + m.s23; mo[5] = mo[5] + m.s33;
    }
    tmin = fminf(tmin, t);
  }
  for (int off = 32; off > 0; off >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, off));
  if ((threadIdx.x & 63) == 0 && tmin < __builtin_inff()) atomicMin(tmin_bits, __float_as_uint(tmin));
}

// builddgammadot0, viscoelastic3d.f90:320-435 (fdgammadot0 :351-433): one
// thread per point, the zones' contributions summed in their order.  The
// gate |x_i - c_i| <= LM uses the largest tapered half extent, so zones
// with all three extents negative (spheres) add nothing, as in the
// reference.  No contraction into fused multiply-adds: a grid point on a
// zone face then lands on the same side as in the reference's arithmetic.
__global__ void k_builddgammadot0(int n1, int n2, int n3, double dx1, double dx2, double dx3, double beta, int nz,
                                  const WeakDev* __restrict__ zones, float* dg) {
#pragma clang fp contract(off)
  const long N = (long)n1 * n2 * n3;
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= N) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = (double)i3 * dx3;
  double f = 0.0;
  for (int i = 0; i < nz; ++i) {
    const WeakDev& Z = zones[i];
    if ((fabs(x3 - Z.z) > Z.LM) || (fabs(x1 - Z.x) > Z.LM) || (fabs(x2 - Z.y) > Z.LM)) continue;
    const double x2r = Z.cs * x1 - Z.ss * x2;
    const double x1s = Z.cd * x2r - Z.sd * x3, x2s = Z.ss * x1 + Z.cs * x2, x3s = Z.sd * x2r + Z.cd * x3;
    const double a = (x1s - Z.xr) / Z.D, b = (x2s - Z.yr) / Z.W, c = (x3s - Z.zr) / Z.L;
    const double D = Z.D, W = Z.W, L = Z.L;
    if (D >= 0 && W >= 0 && L >= 0) f = f + d_omega(a, beta) * d_omega(b, beta) * d_omega(c, beta) * Z.dg;
    else if (D < 0 && W < 0 && L < 0) f = f + d_omega(sqrt(a * a + b * b + c * c), beta) * Z.dg;
    else if (D < 0 && W < 0) f = f + d_omega(sqrt(a * a + b * b), beta) * d_omega(c, beta) * Z.dg;
    else if (D < 0 && L < 0) f = f + d_omega(sqrt(a * a + c * c), beta) * d_omega(b, beta) * Z.dg;
    else if (W < 0 && L < 0) f = f + d_omega(sqrt(b * b + c * c), beta) * d_omega(a, beta) * Z.dg;
  }
  dg[o] = (float)f;
}

__global__ void k_friction_generic(RheoParams R, int plane, int n3, const float* __restrict__ sig, float* moment,
                                   float* vel, long vp, unsigned int* tmin_bits, int* status) {
  const int n1 = R.n1, n2 = R.n2;
  const long N = (long)n1 * n2 * n3;
  float tmin = __builtin_inff();
  int st = 0;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < N; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
    const float* s = sig + 6 * o;
    T6 m{0, 0, 0, 0, 0, 0};
    float* v = vel ? vel + (i1 + vp * (i2 + (long)n2 * i3)) : nullptr;
    const float c = friction_point(R, R.fr.planes[plane], i1, i2, i3, T6{s[0], s[1], s[2], s[3], s[4], s[5]}, m, v,
                                   &st);
    float* q = moment + 6 * o;  // skipped points add zeros
    q[0] += m.s11; q[1] += m.s12; q[2] += m.s13; q[3] += m.s22; q[4] += m.s23; q[5] += m.s33;
    tmin = fminf(tmin, c);
  }
  for (int off = 32; off > 0; off >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, off));
  if ((threadIdx.x & 63) == 0 && tmin < __builtin_inff()) atomicMin(tmin_bits, __float_as_uint(tmin));
  if (st) atomicMax(status, st);
}

// ------------------------------------------------------- field algebra ----
__global__ void k_fieldadd(float* u, const float* v, long n, int has1, float c1, int has2, float c2) {
#pragma clang fp contract(off)
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
    // each product is rounded before the sum, as in the reference
    const float a = u[o], b = v[o];
    if (has1 && has2) u[o] = c1 * a + c2 * b;
    else if (has1) u[o] = c1 * a + b;
    else if (has2) u[o] = a + c2 * b;
    else u[o] = a + b;
  }
}

// tensorfieldadd, elastic3d.f90:3880-3959 (the last branch, with neither
// coefficient present, returns t2 + t2 as the reference does)
__global__ void k_tensorfieldadd(float* t1, const float* t2, long n, int has1, float c1, int has2, float c2) {
#pragma clang fp contract(off)
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < 6 * n; o += (long)gridDim.x * blockDim.x) {
    const float a = t1[o], b = t2[o];
    float r;
    if (has1 && has2) {
      if (0.f == c1) r = (0.f == c2) ? 0.f : c2 * b;
      else r = c1 * a + c2 * b;
    } else if (has1) {
      r = c1 * a + b;
    } else if (has2) {
      r = a + c2 * b;
    } else {
      r = b + b;
    }
    t1[o] = r;
  }
}

__global__ void k_tensoramplitude(const float* t, long n, double* partial) {
  __shared__ double red[4];
  double a = 0;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
    const float* s = t + 6 * o;
    a += tnorm_d(T6{s[0], s[1], s[2], s[3], s[4], s[5]});
  }
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// momentdensityshear, elastic3d.f90:2661-2750
__global__ void k_momentdensityshear(PlaneDev P, double slip, float rmu, int n1, int n2, int n3, double dx1,
                                     double dx2, double dx3, float* sig) {
  const long N = (long)n1 * n2 * n3;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < N; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
    const double x3 = (double)i3 * dx3;
    if (fabs(x3 - P.z) > P.Lp) continue;
    const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2);
    if ((fabs(x1 - P.x) > P.Wp) || (fabs(x2 - P.y) > P.Wp)) continue;
    const double x2r = P.cstrike * x1 - P.sstrike * x2;
    const double x1s = P.cdip * x2r - P.sdip * x3, x1i = P.cdip * x2r + P.sdip * x3;
    if ((fabs(x1s - P.xr) > 7.01 * dx1) && (fabs(x1i - P.xr) > 7.01 * dx1)) continue;
    const double x2s = P.sstrike * x1 + P.cstrike * x2;
    const double x3s = P.sdip * x2r + P.cdip * x3, x3i = -P.sdip * x2r + P.cdip * x3;
    const double temp2 = d_omega((x2s - P.yr) / P.W, P.beta);
    double ap = d_gauss(x1s - P.xr, dx1) * temp2 * d_omega((x3s - P.zr) / P.L, P.beta);
    ap = ap + d_gauss(x1i - P.xr, dx1) * temp2 * d_omega((x3i + P.zr) / P.L, P.beta);
    const double n[3] = {P.n[0] * ap, P.n[1] * ap, P.n[2] * ap};
    const double s[3] = {P.r[0] * slip, P.r[1] * slip, P.r[2] * slip};
    const float e[6] = {(float)(n[0] * s[0]), (float)((n[0] * s[1] + n[1] * s[0]) / 2.0),
                        (float)((n[0] * s[2] + n[2] * s[0]) / 2.0), (float)(n[1] * s[1]),
                        (float)((n[1] * s[2] + n[2] * s[1]) / 2.0), (float)(n[2] * s[2])};
    float* q = sig + 6 * o;
    for (int k = 0; k < 6; ++k) q[k] += rmu * e[k];
  }
}

int sync_status(int rc) {
  if (rc) return rc;
  if (rx_device_sync() != hipSuccess) return RELAX_ERR_HIP;
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// 3-D (nz > 1) or 2-D (nz == 1) transform of a padded device buffer
int fftn_padded(float* d, int n1, int n2, int nz, double dx1, double dx2, double dx3, int dir, hipStream_t st) {
  const int M = n1 / 2;
  const long P = M + 1;  // complex row pitch
  float2* c = reinterpret_cast<float2*>(d);
  int rc;
  if (dir == -1) {
    const double sc = nz > 1 ? dx1 * dx2 * dx3 : dx1 * dx2;
    if ((rc = launch_fft_x_r2c(n1, d, n1 + 2, c, P, c + M, P, (long)n2 * nz, (float)sc, st))) return rc;
    if ((rc = launch_fft_cols(n2, -1, c, P, (long)n2 * P, M + 1, nz, st))) return rc;
    if (nz > 1 && (rc = launch_fft_cols(nz, -1, c, (long)n2 * P, P, M + 1, n2, st))) return rc;
  } else {
    const double sc = nz > 1 ? 1.0 / (n1 * dx1 * n2 * dx2 * nz * dx3) : 1.0 / (n1 * dx1 * n2 * dx2);
    if (nz > 1 && (rc = launch_fft_cols(nz, +1, c, (long)n2 * P, P, M + 1, n2, st))) return rc;
    if ((rc = launch_fft_cols(n2, +1, c, P, (long)n2 * P, M + 1, nz, st))) return rc;
    if ((rc = launch_fft_x_c2r(n1, EPI_PLAIN, c, P, c + M, P, d, n1 + 2, (long)n2 * nz, (float)sc, nullptr, 0,
                               nullptr, st)))
      return rc;
  }
  return RELAX_OK;
}

// host <-> device copies between the reference layout (padded rows of n1+2 or
// packed rows of n1) and the device's padded layout
int put_real(float* dev, const float* host, int n1, long rows, int align) {
  if (align) HIP_TRY(hipMemcpy(dev, host, (size_t)(n1 + 2) * rows * 4, hipMemcpyHostToDevice));
  else HIP_TRY(hipMemcpy2D(dev, (n1 + 2) * 4, host, n1 * 4, n1 * 4, rows, hipMemcpyHostToDevice));
  return RELAX_OK;
}
int get_real(float* host, const float* dev, int n1, long rows, int align) {
  if (align) HIP_TRY(hipMemcpy(host, dev, (size_t)(n1 + 2) * rows * 4, hipMemcpyDeviceToHost));
  else HIP_TRY(hipMemcpy2D(host, n1 * 4, dev, (n1 + 2) * 4, n1 * 4, rows, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

GreenConst green_const(int n1, int n2, int n3, double dx1, double dx2, double dx3, double l, double m, double g) {
  GreenConst gc;
  gc.lambda = l; gc.mu = m; gc.gamma = g;
  gc.dx1 = dx1; gc.dx2 = dx2; gc.dx3 = dx3;
  gc.n1 = n1; gc.n2 = n2; gc.n3 = n3;
  return gc;
}

}  // namespace

// =========================================================== fourier.f90 ==
int relax_fft3(float* data, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, int direction,
               int align_data) {
  if (!data || (direction != -1 && direction != 1)) return RELAX_ERR_ARG;
  if (!pow2_in(sx1, 16, 1024) || !pow2_in(sx2, 16, 1024) || !pow2_in(sx3, 16, 1024)) return RELAX_ERR_SIZE;
  int rc;
  if ((rc = has_device())) return rc;
  const long rows = (long)sx2 * sx3;
  Buf d;
  if ((rc = d.alloc((size_t)(sx1 + 2) * rows * 4))) return rc;
  const bool real_in = direction == -1;
  if (real_in) { if ((rc = put_real(d.f(), data, sx1, rows, align_data))) return rc; }
  else HIP_TRY(hipMemcpy(d.p, data, (size_t)(sx1 + 2) * rows * 4, hipMemcpyHostToDevice));
  if ((rc = sync_status(fftn_padded(d.f(), sx1, sx2, sx3, dx1, dx2, dx3, direction, 0)))) return rc;
  if (real_in) HIP_TRY(hipMemcpy(data, d.p, (size_t)(sx1 + 2) * rows * 4, hipMemcpyDeviceToHost));
  else if ((rc = get_real(data, d.f(), sx1, rows, align_data))) return rc;
  return RELAX_OK;
}

int relax_fft2(float* data, int sx1, int sx2, double dx1, double dx2, int direction, int align_data) {
  if (!data || (direction != -1 && direction != 1)) return RELAX_ERR_ARG;
  if (!pow2_in(sx1, 16, 1024) || !pow2_in(sx2, 16, 1024)) return RELAX_ERR_SIZE;
  int rc;
  if ((rc = has_device())) return rc;
  Buf d;
  if ((rc = d.alloc((size_t)(sx1 + 2) * sx2 * 4))) return rc;
  const bool real_in = direction == -1;
  if (real_in) { if ((rc = put_real(d.f(), data, sx1, sx2, align_data))) return rc; }
  else HIP_TRY(hipMemcpy(d.p, data, (size_t)(sx1 + 2) * sx2 * 4, hipMemcpyHostToDevice));
  if ((rc = sync_status(fftn_padded(d.f(), sx1, sx2, 1, dx1, dx2, 1.0, direction, 0)))) return rc;
  if (real_in) HIP_TRY(hipMemcpy(data, d.p, (size_t)(sx1 + 2) * sx2 * 4, hipMemcpyDeviceToHost));
  else if ((rc = get_real(data, d.f(), sx1, sx2, align_data))) return rc;
  return RELAX_OK;
}

// ============================================================= green.f90 ==
int relax_elasticresponse(double lambda, double mu, float* f1, float* f2, float* f3, int sx1, int sx2, int sx3,
                          double dx1, double dx2, double dx3) {
  if (!f1 || !f2 || !f3) return RELAX_ERR_ARG;
  if (!pow2_in(sx1, 2, 4096) || sx2 < 1 || sx3 < 1) return RELAX_ERR_SIZE;
  int rc;
  if ((rc = has_device())) return rc;
  const size_t bytes = (size_t)(sx1 + 2) * sx2 * sx3 * 4;
  Buf a, b, c;
  if ((rc = a.alloc(bytes)) || (rc = b.alloc(bytes)) || (rc = c.alloc(bytes))) return rc;
  HIP_TRY(hipMemcpy(a.p, f1, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, f2, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c.p, f3, bytes, hipMemcpyHostToDevice));
  const GreenConst gc = green_const(sx1, sx2, sx3, dx1, dx2, dx3, lambda, mu, 0.0);
  if ((rc = sync_status(launch_elastic_padded(gc, a.c(), b.c(), c.c(), 0)))) return rc;
  HIP_TRY(hipMemcpy(f1, a.p, bytes, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(f2, b.p, bytes, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(f3, c.p, bytes, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_surfacetractioncowling(double lambda, double mu, double gamma, const float* u1, const float* u2,
                                 const float* u3, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,
                                 float* p1, float* p2, float* p3) {
  if (!u1 || !u2 || !u3 || !p1 || !p2 || !p3) return RELAX_ERR_ARG;
  if (!pow2_in(sx1, 2, 4096) || sx2 < 1 || sx3 < 1) return RELAX_ERR_SIZE;
  int rc;
  if ((rc = has_device())) return rc;
  const size_t bytes = (size_t)(sx1 + 2) * sx2 * sx3 * 4, pb = (size_t)(sx1 + 2) * sx2 * 4;
  Buf a, b, c, q1, q2, q3;
  if ((rc = a.alloc(bytes)) || (rc = b.alloc(bytes)) || (rc = c.alloc(bytes)) || (rc = q1.alloc(pb)) ||
      (rc = q2.alloc(pb)) || (rc = q3.alloc(pb)))
    return rc;
  HIP_TRY(hipMemcpy(a.p, u1, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, u2, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c.p, u3, bytes, hipMemcpyHostToDevice));
  const GreenConst gc = green_const(sx1, sx2, sx3, dx1, dx2, dx3, lambda, mu, gamma);
  if ((rc = sync_status(launch_surface_padded(gc, a.c(), b.c(), c.c(), nullptr, nullptr, nullptr, q1.c(), q2.c(),
                                              q3.c(), 0))))
    return rc;
  HIP_TRY(hipMemcpy(p1, q1.p, pb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(p2, q2.p, pb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(p3, q3.p, pb, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_cerrutimodified(const float* p1, const float* p2, const float* p3, double lambda, double mu,
                          double gamma, float* u1, float* u2, float* u3, int sx1, int sx2, int sx3, double dx1,
                          double dx2, double dx3) {
  if (!u1 || !u2 || !u3 || !p1 || !p2 || !p3) return RELAX_ERR_ARG;
  if (!pow2_in(sx1, 2, 4096) || sx2 < 1 || sx3 < 1) return RELAX_ERR_SIZE;
  int rc;
  if ((rc = has_device())) return rc;
  const size_t bytes = (size_t)(sx1 + 2) * sx2 * sx3 * 4, pb = (size_t)(sx1 + 2) * sx2 * 4;
  Buf a, b, c, q1, q2, q3;
  if ((rc = a.alloc(bytes)) || (rc = b.alloc(bytes)) || (rc = c.alloc(bytes)) || (rc = q1.alloc(pb)) ||
      (rc = q2.alloc(pb)) || (rc = q3.alloc(pb)))
    return rc;
  HIP_TRY(hipMemcpy(a.p, u1, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, u2, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c.p, u3, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(q1.p, p1, pb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(q2.p, p2, pb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(q3.p, p3, pb, hipMemcpyHostToDevice));
  const GreenConst gc = green_const(sx1, sx2, sx3, dx1, dx2, dx3, lambda, mu, gamma);
  if ((rc = sync_status(launch_cerruti_padded(gc, q1.c(), q2.c(), q3.c(), a.c(), b.c(), c.c(), 0)))) return rc;
  HIP_TRY(hipMemcpy(u1, a.p, bytes, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(u2, b.p, bytes, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(u3, c.p, bytes, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

// greenfunctioncowling on device buffers in the padded layout: v, three
// fields of (sx1+2) sx2 sx3 floats over the full depth; t, three planes of
// (sx1+2) sx2; p, three planes of scratch.  Enqueued on stream st; with sync
// the call waits for it and reports the kernels' status.
int green_full_device(float* const v[3], float* const t[3], float* const p[3], int sx1, int sx2, int sx3,
                      double dx1, double dx2, double dx3, double lambda, double mu, double gamma, hipStream_t st,
                      bool sync) {
  int rc;
  const GreenConst gc = green_const(sx1, sx2, sx3, dx1, dx2, dx3, lambda, mu, gamma);
  float2* vc[3] = {(float2*)v[0], (float2*)v[1], (float2*)v[2]};
  float2* tc[3] = {(float2*)t[0], (float2*)t[1], (float2*)t[2]};
  float2* pc[3] = {(float2*)p[0], (float2*)p[1], (float2*)p[2]};
  for (int k = 0; k < 3; ++k) {
    if ((rc = fftn_padded(v[k], sx1, sx2, sx3, dx1, dx2, dx3, -1, st))) return rc;
    if ((rc = fftn_padded(t[k], sx1, sx2, 1, dx1, dx2, dx3, -1, st))) return rc;
  }
  if ((rc = launch_elastic_padded(gc, vc[0], vc[1], vc[2], st))) return rc;
  if ((rc = launch_surface_padded(gc, vc[0], vc[1], vc[2], tc[0], tc[1], tc[2], pc[0], pc[1], pc[2], st))) return rc;
  if ((rc = launch_cerruti_padded(gc, pc[0], pc[1], pc[2], vc[0], vc[1], vc[2], st))) return rc;
  for (int k = 0; k < 3; ++k) {
    if ((rc = fftn_padded(v[k], sx1, sx2, sx3, dx1, dx2, dx3, +1, st))) return rc;
    if ((rc = fftn_padded(t[k], sx1, sx2, 1, dx1, dx2, dx3, +1, st))) return rc;
  }
  if (!sync) return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
  HIP_TRY(hipStreamSynchronize(st));
  return sync_status(RELAX_OK);
}

int relax_greenfunctioncowling(float* c1, float* c2, float* c3, float* t1, float* t2, float* t3, int sx1,
                               int sx2, int sx3, double dx1, double dx2, double dx3, double lambda, double mu,
                               double gamma, int align_data) {
  if (!c1 || !c2 || !c3 || !t1 || !t2 || !t3) return RELAX_ERR_ARG;
  if (!pow2_in(sx1, 16, 1024) || !pow2_in(sx2, 16, 1024) || !pow2_in(sx3, 16, 1024)) return RELAX_ERR_SIZE;
  int rc;
  if ((rc = has_device())) return rc;
  const long rows = (long)sx2 * sx3;
  const size_t vb = (size_t)(sx1 + 2) * rows * 4, pb = (size_t)(sx1 + 2) * sx2 * 4;
  Buf v[3], t[3], p[3];
  for (int k = 0; k < 3; ++k)
    if ((rc = v[k].alloc(vb)) || (rc = t[k].alloc(pb)) || (rc = p[k].alloc(pb))) return rc;
  float* hc[3] = {c1, c2, c3};
  float* ht[3] = {t1, t2, t3};
  for (int k = 0; k < 3; ++k) {
    if ((rc = put_real(v[k].f(), hc[k], sx1, rows, align_data))) return rc;
    if ((rc = put_real(t[k].f(), ht[k], sx1, sx2, align_data))) return rc;
  }
  float* vd[3] = {v[0].f(), v[1].f(), v[2].f()};
  float* td[3] = {t[0].f(), t[1].f(), t[2].f()};
  float* pd[3] = {p[0].f(), p[1].f(), p[2].f()};
  if ((rc = green_full_device(vd, td, pd, sx1, sx2, sx3, dx1, dx2, dx3, lambda, mu, gamma, 0, true))) return rc;
  for (int k = 0; k < 3; ++k) {
    if ((rc = get_real(hc[k], v[k].f(), sx1, rows, align_data))) return rc;
    if ((rc = get_real(ht[k], t[k].f(), sx1, sx2, align_data))) return rc;
  }
  return RELAX_OK;
}

// ------------------------------------------ greenfunctioncowling, no allocation --
// The three traction planes of the surface correction, for grids of up to
// 1024 x 1024 horizontally, in a static device array: every device the
// library runs on holds one (12.0 MiB) until the process ends, and every
// call on a device shares it.  Host code uses the address that
// hipGetSymbolAddress returns for the current device, never the array's
// name.
static const long NOALLOC_MAX_N = 1024;
static const long NOALLOC_PLANE = (NOALLOC_MAX_N + 2) * NOALLOC_MAX_N;  // floats per plane
static __device__ __align__(256) float scratch[3 * NOALLOC_PLANE];
static std::mutex g_scratch_mu;
static std::map<int, float*> g_scratch;  // device -> address of its scratch

// The current device's scratch address, looked up once per device.
static int scratch_address(float** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return RELAX_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it == g_scratch.end()) {
    float* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RELAX_ERR_HIP;
    it = g_scratch.emplace(dev, p).first;
  }
  *out = it->second;
  return RELAX_OK;
}

// Per-device set-up of the allocation-free entry: the scratch address and
// the twiddle tables of every transform length (relax_hip_init calls it).
int noalloc_setup() {
  float* p = nullptr;
  if (int rc = scratch_address(&p)) return rc;
  return twiddle_table(2) ? RELAX_OK : RELAX_ERR_HIP;
}

int relax_greenfunctioncowling_noalloc(float* c1, float* c2, float* c3, float* t1, float* t2, float* t3, int sx1,
                                       int sx2, int sx3, double dx1, double dx2, double dx3, double lambda,
                                       double mu, double gamma, void* stream) {
  if (!c1 || !c2 || !c3 || !t1 || !t2 || !t3) return RELAX_ERR_ARG;
  if (!pow2_in(sx1, 16, NOALLOC_MAX_N) || !pow2_in(sx2, 16, NOALLOC_MAX_N) || !pow2_in(sx3, 16, 1024))
    return RELAX_ERR_SIZE;
  if ((long)(sx1 + 2) * sx2 > NOALLOC_PLANE) return RELAX_ERR_SIZE;
  int rc;
  if ((rc = has_device())) return rc;
  float* base = nullptr;
  if ((rc = noalloc_setup()) || (rc = scratch_address(&base))) return rc;
  const long plane = (long)(sx1 + 2) * sx2;
  float* v[3] = {c1, c2, c3};
  float* t[3] = {t1, t2, t3};
  float* pp[3] = {base, base + plane, base + 2 * plane};
  return green_full_device(v, t, pp, sx1, sx2, sx3, dx1, dx2, dx3, lambda, mu, gamma, (hipStream_t)stream, false);
}

// ========================================================= elastic3d.f90 ==
int relax_stressupdate(const float* v1, const float* v2, const float* v3, double lambda, double mu, double dx1,
                       double dx2, double dx3, int sx1, int sx2, int sx3, relax_tensor_t* sig, int align_data) {
  if (!v1 || !v2 || !v3 || !sig) return RELAX_ERR_ARG;
  Taps T;
  int rc;
  if ((rc = make_taps(sx1, sx2, sx3, dx1, dx2, dx3, T))) return rc;
  if ((rc = has_device())) return rc;
  const long vp = align_data ? sx1 + 2 : sx1;
  const long N = (long)sx1 * sx2 * sx3;
  const size_t vb = (size_t)vp * sx2 * sx3 * 4;
  Buf a, b, c, s;
  if ((rc = a.alloc(vb)) || (rc = b.alloc(vb)) || (rc = c.alloc(vb)) || (rc = s.alloc(N * 24))) return rc;
  HIP_TRY(hipMemcpy(a.p, v1, vb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, v2, vb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c.p, v3, vb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.p, sig, N * 24, hipMemcpyHostToDevice));
  k_stressupdate_generic<<<g1d(N), 256>>>(a.f(), b.f(), c.f(), vp, sx1, sx2, sx3, T, dx3, lambda, mu, s.f());
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(sig, s.p, N * 24, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_equivalentbodyforce(const relax_tensor_t* sig, double dx1, double dx2, double dx3, int sx1, int sx2,
                              int sx3, float* c1, float* c2, float* c3, float* t1, float* t2, float* t3,
                              const float* mask, int align_data) {
  if (!sig || !c1 || !c2 || !c3 || !t1 || !t2 || !t3) return RELAX_ERR_ARG;
  Taps T;
  int rc;
  if ((rc = make_taps(sx1, sx2, sx3, dx1, dx2, dx3, T))) return rc;
  if ((rc = has_device())) return rc;
  const long cp = align_data ? sx1 + 2 : sx1;
  const long N = (long)sx1 * sx2 * sx3;
  const size_t cb = (size_t)cp * sx2 * sx3 * 4, tb = (size_t)cp * sx2 * 4;
  Buf s, a, b, c, d, e, f, mk;
  if ((rc = s.alloc(N * 24)) || (rc = a.alloc(cb)) || (rc = b.alloc(cb)) || (rc = c.alloc(cb)) ||
      (rc = d.alloc(tb)) || (rc = e.alloc(tb)) || (rc = f.alloc(tb)))
    return rc;
  HIP_TRY(hipMemcpy(s.p, sig, N * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(a.p, c1, cb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, c2, cb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c.p, c3, cb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d.p, t1, tb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e.p, t2, tb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(f.p, t3, tb, hipMemcpyHostToDevice));
  if (mask) {
    if ((rc = mk.alloc(sx3 * 4))) return rc;
    HIP_TRY(hipMemcpy(mk.p, mask, sx3 * 4, hipMemcpyHostToDevice));
  }
  k_traction_add<<<g1d((long)sx1 * sx2), 256>>>(s.f(), sx1, sx2, d.f(), e.f(), f.f(), cp);
  k_bodyforce_generic<<<g1d(N), 256>>>(s.f(), sx1, sx2, sx3, T, dx3, a.f(), b.f(), c.f(), cp,
                                       mask ? mk.f() : nullptr);
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(c1, a.p, cb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(c2, b.p, cb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(c3, c.p, cb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(t1, d.p, tb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(t2, e.p, tb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(t3, f.p, tb, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_fieldadd(float* u, const float* v, long n, const float* c1, const float* c2) {
  if (!u || !v || n < 0) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  Buf a, b;
  if ((rc = a.alloc(n * 4)) || (rc = b.alloc(n * 4))) return rc;
  HIP_TRY(hipMemcpy(a.p, u, n * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, v, n * 4, hipMemcpyHostToDevice));
  k_fieldadd<<<g1d(n), 256>>>(a.f(), b.f(), n, c1 != nullptr, c1 ? *c1 : 0.f, c2 != nullptr, c2 ? *c2 : 0.f);
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(u, a.p, n * 4, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_tensorfieldadd(relax_tensor_t* t1, const relax_tensor_t* t2, long n, const float* c1, const float* c2) {
  if (!t1 || !t2 || n < 0) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  Buf a, b;
  if ((rc = a.alloc(n * 24)) || (rc = b.alloc(n * 24))) return rc;
  HIP_TRY(hipMemcpy(a.p, t1, n * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, t2, n * 24, hipMemcpyHostToDevice));
  k_tensorfieldadd<<<g1d(6 * n), 256>>>(a.f(), b.f(), n, c1 != nullptr, c1 ? *c1 : 0.f, c2 != nullptr,
                                        c2 ? *c2 : 0.f);
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(t1, a.p, n * 24, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_tensoramplitude(const relax_tensor_t* t, long n, double dx1, double dx2, double dx3, double* out) {
  if (!t || !out || n < 0) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  const dim3 g = g1d(n);
  Buf a, p;
  if ((rc = a.alloc(n * 24)) || (rc = p.alloc(g.x * 8))) return rc;
  HIP_TRY(hipMemcpy(a.p, t, n * 24, hipMemcpyHostToDevice));
  k_tensoramplitude<<<g, 256>>>(a.f(), n, (double*)p.p);
  if ((rc = sync_status(RELAX_OK))) return rc;
  std::vector<double> h(g.x);
  HIP_TRY(hipMemcpy(h.data(), p.p, g.x * 8, hipMemcpyDeviceToHost));
  double s = 0;
  for (double x : h) s += x;
  *out = s * (dx1 * dx2 * dx3);
  return RELAX_OK;
}

int relax_momentdensityshear(double mu, double slip, double x, double y, double z, double L, double W,
                             double strike, double dip, double rake, double beta, int sx1, int sx2, int sx3,
                             double dx1, double dx2, double dx3, relax_tensor_t* sig) {
  if (!sig || sx1 < 1 || sx2 < 1 || sx3 < 1) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  const long N = (long)sx1 * sx2 * sx3;
  Buf s;
  if ((rc = s.alloc(N * 24))) return rc;
  HIP_TRY(hipMemcpy(s.p, sig, N * 24, hipMemcpyHostToDevice));
  relax_plane_t p{x, y, z, /*width*/ L, /*length*/ W, strike, dip, rake, 0, 0};
  const PlaneDev P = make_plane(p, beta, dx1, dx3);
  k_momentdensityshear<<<g1d(N), 256>>>(P, slip, 2.0f * (float)mu, sx1, sx2, sx3, dx1, dx2, dx3, s.f());
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(sig, s.p, N * 24, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_source(double mu, double slip, double x, double y, double z, double L, double W, double strike,
                 double dip, double rake, double beta, int sx1, int sx2, int sx3, double dx1, double dx2,
                 double dx3, float* f1, float* f2, float* f3, float* t1, float* t2, float* t3, int align) {
  if (!f1 || !f2 || !f3 || !t1 || !t2 || !t3 || sx1 < 1 || sx2 < 1 || sx3 < 2) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  // the top sx3/2 planes are the first ld*sx2*(sx3/2) values of each array
  const long ld = sx1 + (align ? 2 : 0), fs = ld * sx2 * (sx3 / 2), ts = ld * sx2;
  Buf f, t, pl, sl;
  if ((rc = f.alloc(3 * fs * 4)) || (rc = t.alloc(3 * ts * 4)) || (rc = pl.alloc(sizeof(PlaneDev))) ||
      (rc = sl.alloc(sizeof(double))))
    return rc;
  float* fh[3] = {f1, f2, f3};
  float* th[3] = {t1, t2, t3};
  for (int c = 0; c < 3; ++c) {
    HIP_TRY(hipMemcpy(f.f() + c * fs, fh[c], fs * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.f() + c * ts, th[c], ts * 4, hipMemcpyHostToDevice));
  }
  // the driver passes (width, length) as (L, W)
  const relax_plane_t p{x, y, z, /*width*/ L, /*length*/ W, strike, dip, rake, 0, 0};
  const PlaneDev P = make_plane(p, beta, dx1, dx3);
  HIP_TRY(hipMemcpy(pl.p, &P, sizeof(PlaneDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(sl.p, &slip, sizeof(double), hipMemcpyHostToDevice));
  if ((rc = launch_source((const PlaneDev*)pl.p, (const double*)sl.p, 1, mu, sx1, sx2, sx3 / 2, dx1, dx2, dx3,
                          f.f(), ld, fs, t.f(), ts, 0)))
    return rc;
  if ((rc = sync_status(RELAX_OK))) return rc;
  for (int c = 0; c < 3; ++c) {
    HIP_TRY(hipMemcpy(fh[c], f.f() + c * fs, fs * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(th[c], t.f() + c * ts, ts * 4, hipMemcpyDeviceToHost));
  }
  return RELAX_OK;
}

int relax_tensilesource(double lambda, double mu, double s, double x, double y, double z, double L, double W,
                        double strike, double dip, double beta, int sx1, int sx2, int sx3, double dx1, double dx2,
                        double dx3, float* f1, float* f2, float* f3, int align) {
  if (!f1 || !f2 || !f3 || sx1 < 1 || sx2 < 1 || sx3 < 2) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  const long ld = sx1 + (align ? 2 : 0), fs = ld * sx2 * sx3;
  Buf f, pl, op;
  if ((rc = f.alloc(3 * fs * 4)) || (rc = pl.alloc(sizeof(PlaneDev))) || (rc = op.alloc(sizeof(double)))) return rc;
  float* fh[3] = {f1, f2, f3};
  for (int c = 0; c < 3; ++c) HIP_TRY(hipMemcpy(f.f() + c * fs, fh[c], fs * 4, hipMemcpyHostToDevice));
  const relax_plane_t p{x, y, z, /*width*/ L, /*length*/ W, strike, dip, 0.0, 0, 0};
  const PlaneDev P = make_plane(p, beta, dx1, dx3);
  HIP_TRY(hipMemcpy(pl.p, &P, sizeof(PlaneDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(op.p, &s, sizeof(double), hipMemcpyHostToDevice));
  if ((rc = launch_tensile_source((const PlaneDev*)pl.p, (const double*)op.p, 1, lambda, mu, sx1, sx2, sx3, dx1, dx2,
                                  dx3, f.f(), ld, fs, 0)))
    return rc;
  if ((rc = sync_status(RELAX_OK))) return rc;
  for (int c = 0; c < 3; ++c) HIP_TRY(hipMemcpy(fh[c], f.f() + c * fs, fs * 4, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_mogisource(double lambda, double mu, double o, double xs, double ys, double zs, int sx1, int sx2, int sx3,
                     double dx1, double dx2, double dx3, float* f1, float* f2, float* f3, int align) {
  if (!f1 || !f2 || !f3 || sx1 < 1 || sx2 < 1 || sx3 < 2) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  const long ld = sx1 + (align ? 2 : 0), fs = ld * sx2 * sx3;
  Buf f, ms;
  if ((rc = f.alloc(3 * fs * 4)) || (rc = ms.alloc(sizeof(relax_source_t)))) return rc;
  float* fh[3] = {f1, f2, f3};
  for (int c = 0; c < 3; ++c) HIP_TRY(hipMemcpy(f.f() + c * fs, fh[c], fs * 4, hipMemcpyHostToDevice));
  relax_source_t M{};
  M.slip = o; M.x = xs; M.y = ys; M.z = zs;
  HIP_TRY(hipMemcpy(ms.p, &M, sizeof(M), hipMemcpyHostToDevice));
  if ((rc = launch_mogi_source((const relax_source_t*)ms.p, 1, lambda, mu, sx1, sx2, sx3, dx1, dx2, dx3, f.f(), ld,
                               fs, 0)))
    return rc;
  if ((rc = sync_status(RELAX_OK))) return rc;
  for (int c = 0; c < 3; ++c) HIP_TRY(hipMemcpy(fh[c], f.f() + c * fs, fs * 4, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

// a TENSOR field (AoS) through the device's six planes
static int tensor_roundtrip(relax_tensor_t* sig, long N, const std::function<int(float*)>& body) {
  int rc;
  Buf s;
  if ((rc = s.alloc(N * 24))) return rc;
  std::vector<float> h(6 * N);
  for (long i = 0; i < N; ++i) {
    const relax_tensor_t& t = sig[i];
    h[i] = t.s11; h[N + i] = t.s12; h[2 * N + i] = t.s13; h[3 * N + i] = t.s22; h[4 * N + i] = t.s23; h[5 * N + i] = t.s33;
  }
  HIP_TRY(hipMemcpy(s.p, h.data(), N * 24, hipMemcpyHostToDevice));
  if ((rc = body(s.f()))) return rc;
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(h.data(), s.p, N * 24, hipMemcpyDeviceToHost));
  for (long i = 0; i < N; ++i) sig[i] = relax_tensor_t{h[i], h[N + i], h[2 * N + i], h[3 * N + i], h[4 * N + i], h[5 * N + i]};
  return RELAX_OK;
}

int relax_momentdensitytensile(double lambda, double mu, double s, double x, double y, double z, double L, double W,
                               double strike, double dip, double beta, int sx1, int sx2, int sx3, double dx1,
                               double dx2, double dx3, relax_tensor_t* sig) {
  if (!sig || sx1 < 1 || sx2 < 1 || sx3 < 1) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  Buf pl, op;
  if ((rc = pl.alloc(sizeof(PlaneDev))) || (rc = op.alloc(sizeof(double)))) return rc;
  const relax_plane_t p{x, y, z, /*width*/ L, /*length*/ W, strike, dip, 0.0, 0, 0};
  const PlaneDev P = make_plane(p, beta, dx1, dx3);
  HIP_TRY(hipMemcpy(pl.p, &P, sizeof(PlaneDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(op.p, &s, sizeof(double), hipMemcpyHostToDevice));
  const long N = (long)sx1 * sx2 * sx3;
  return tensor_roundtrip(sig, N, [&](float* d) {
    return launch_add_tensile((const PlaneDev*)pl.p, (const double*)op.p, 1, lambda, mu, sx1, sx2, sx3, dx1, dx2, dx3,
                              d, N, 0);
  });
}

int relax_momentdensitymogi(double lambda, double mu, double o, double xs, double ys, double zs, int sx1, int sx2,
                            int sx3, double dx1, double dx2, double dx3, relax_tensor_t* sig) {
  if (!sig || sx1 < 1 || sx2 < 1 || sx3 < 1) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  Buf ms;
  if ((rc = ms.alloc(sizeof(relax_source_t)))) return rc;
  relax_source_t M{};
  M.slip = o; M.x = xs; M.y = ys; M.z = zs;
  HIP_TRY(hipMemcpy(ms.p, &M, sizeof(M), hipMemcpyHostToDevice));
  const long N = (long)sx1 * sx2 * sx3;
  return tensor_roundtrip(sig, N, [&](float* d) {
    return launch_add_mogi((const relax_source_t*)ms.p, 1, lambda, mu, sx1, sx2, sx3, dx1, dx2, dx3, d, N, 0);
  });
}

int relax_traction(double mu, const relax_source_t* l, int nl, int sx1, int sx2, double dx1, double dx2, double t,
                   double Dt, float* t3, int rate, int align) {
  (void)mu;  // the reference's argument list carries mu; the loads do not use it
  if (!t3 || nl < 0 || (nl > 0 && !l) || sx1 < 1 || sx2 < 1) return RELAX_ERR_ARG;
  if (rate && Dt == 0.0) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  const long ld = sx1 + (align ? 2 : 0), ts = ld * sx2;
  Buf tb, lb;
  if ((rc = tb.alloc(ts * 4)) || (rc = lb.alloc(sizeof(relax_source_t) * (nl ? nl : 1)))) return rc;
  HIP_TRY(hipMemcpy(tb.p, t3, ts * 4, hipMemcpyHostToDevice));
  if (nl) HIP_TRY(hipMemcpy(lb.p, l, sizeof(relax_source_t) * nl, hipMemcpyHostToDevice));
  if ((rc = launch_traction((const relax_source_t*)lb.p, nl, sx1, sx2, dx1, dx2, t, Dt, rate, tb.f(), ld, 0)))
    return rc;
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(t3, tb.p, ts * 4, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

// ============================================ viscoelastic3d / friction3d ==
namespace {
// per-depth tables of one mechanism and the deviatoric prestress
int upload_mech(const relax_layer_t* st, const relax_tensor_layer_t* pre, int n, Buf& tab, Mechanism& M,
                Prestress& P) {
  std::vector<float> h(10 * n, 0.f);
  for (int i = 0; i < n; ++i) {
    h[i] = (float)st[i].gammadot0;
    h[n + i] = (float)st[i].stressexponent;
    h[2 * n + i] = (float)st[i].cohesion;
    float d[6] = {0, 0, 0, 0, 0, 0};
    if (pre) {
      const relax_tensor_t& t = pre[i].t;
      const float diag = (float)((double)(t.s11 + t.s22 + t.s33) / 3.0);
      const float dd[6] = {t.s11 - diag, t.s12, t.s13, t.s22 - diag, t.s23, t.s33 - diag};
      for (int k = 0; k < 6; ++k) d[k] = dd[k];
    }
    for (int k = 0; k < 6; ++k) h[3 * n + 6 * i + k] = d[k];
    h[9 * n + i] = (float)std::sqrt(((double)(d[0] * d[0]) + 2.0 * (double)(d[1] * d[1]) +
                                     2.0 * (double)(d[2] * d[2]) + (double)(d[3] * d[3]) +
                                     2.0 * (double)(d[4] * d[4]) + (double)(d[5] * d[5])) / 2.0);
  }
  int rc;
  if ((rc = tab.alloc(h.size() * 4))) return rc;
  HIP_TRY(hipMemcpy(tab.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  M.gammadot0 = tab.f(); M.power = tab.f() + n; M.cohesion = tab.f() + 2 * n; M.active = 1;
  P.sp = tab.f() + 3 * n; P.taup = tab.f() + 9 * n;
  return RELAX_OK;
}
}  // namespace

int relax_viscouseigenstress(double mu, const relax_layer_t* structure, int nstructure,
                             const relax_tensor_layer_t* prestress, int nprestress, const relax_tensor_t* sig,
                             int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, relax_tensor_t* moment,
                             double* maxwelltime, const float* dgammadot0, float* gamma, int align_data) {
  (void)dx1; (void)dx2; (void)dx3;
  if (!structure || !prestress || !sig || !moment) return RELAX_ERR_ARG;
  // viscoelastic3d.f90:240-241: a structure of the wrong size is ignored
  if (nstructure != sx3 || nprestress != sx3) return RELAX_OK;
  for (int i = 0; i < sx3; ++i)
    if (structure[i].stressexponent < 0.999999) return RELAX_ERR_POWER;
  int rc;
  if ((rc = has_device())) return rc;
  const long N = (long)sx1 * sx2 * sx3;
  const long gp = align_data ? sx1 + 2 : sx1;
  const size_t gb = (size_t)gp * sx2 * sx3 * 4;
  Buf tab, s, m, dg, ga, tb;
  Mechanism M;
  Prestress P;
  if ((rc = upload_mech(structure, prestress, sx3, tab, M, P))) return rc;
  if ((rc = s.alloc(N * 24)) || (rc = m.alloc(N * 24)) || (rc = tb.alloc(4))) return rc;
  HIP_TRY(hipMemcpy(s.p, sig, N * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m.p, moment, N * 24, hipMemcpyHostToDevice));
  const unsigned int inf_bits = 0x7f800000u;
  HIP_TRY(hipMemcpy(tb.p, &inf_bits, 4, hipMemcpyHostToDevice));
  if (dgammadot0) {
    if ((rc = dg.alloc(N * 4))) return rc;
    HIP_TRY(hipMemcpy(dg.p, dgammadot0, N * 4, hipMemcpyHostToDevice));
  }
  if (gamma) {
    if ((rc = ga.alloc(gb))) return rc;
    HIP_TRY(hipMemcpy(ga.p, gamma, gb, hipMemcpyHostToDevice));
  }
  k_viscous_generic<<<g1d(N), 256>>>(M, P, (float)mu, sx1, sx2, sx3, s.f(), m.f(), dgammadot0 ? dg.f() : nullptr,
                                     gamma ? ga.f() : nullptr, gp, (unsigned int*)tb.p);
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(moment, m.p, N * 24, hipMemcpyDeviceToHost));
  if (gamma) HIP_TRY(hipMemcpy(gamma, ga.p, gb, hipMemcpyDeviceToHost));
  if (maxwelltime) {
    unsigned int bits;
    HIP_TRY(hipMemcpy(&bits, tb.p, 4, hipMemcpyDeviceToHost));
    float tm;
    std::memcpy(&tm, &bits, 4);
    tm = std::fmin(tm, (float)*maxwelltime);
    *maxwelltime = std::fmin((double)tm, *maxwelltime);
  }
  return RELAX_OK;
}

WeakDev make_weak(const relax_weak_t& z, double beta) {
  WeakDev w;
  w.dg = z.dgammadot0; w.x = z.x; w.y = z.y; w.z = z.z;
  // the zone's length is its x2 extent W, its width the x3 extent L
  w.W = z.length; w.L = z.width; w.D = z.thickness;
  const double Wp = w.W * (1.0 + 2.0 * beta) / 2.0, Lp = w.L * (1.0 + 2.0 * beta) / 2.0,
               Dp = w.D * (1.0 + 2.0 * beta) / 2.0;
  w.LM = std::max(Wp, std::max(Lp, Dp));
  w.cs = std::cos(z.strike); w.ss = std::sin(z.strike);
  w.cd = std::cos(z.dip); w.sd = std::sin(z.dip);
  const double x2r = w.cs * z.x - w.ss * z.y;
  w.xr = w.cd * x2r - w.sd * z.z;
  w.yr = w.ss * z.x + w.cs * z.y;
  w.zr = w.sd * x2r + w.cd * z.z;
  return w;
}

int launch_builddgammadot0(int n1, int n2, int n3, double dx1, double dx2, double dx3, double beta, int nz,
                           const WeakDev* zones, float* dg, hipStream_t st) {
  const long N = (long)n1 * n2 * n3;
  k_builddgammadot0<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(n1, n2, n3, dx1, dx2, dx3, beta, nz, zones, dg);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int relax_transienteigenstress(double mu, const relax_layer_t* structure, int nstructure,
                               const relax_tensor_t* sig, const relax_tensor_t* epsilonik, int sx1, int sx2,
                               int sx3, double dx1, double dx2, double dx3, relax_tensor_t* moment,
                               relax_tensor_t* epsilonikdot, double* maxwelltime, const float* dgammadot0) {
  (void)dx1; (void)dx2; (void)dx3;
  if (!structure || !sig || !epsilonik || !moment || !epsilonikdot) return RELAX_ERR_ARG;
  // viscoelastic3d.f90:144: a structure of the wrong size is ignored
  if (nstructure != sx3) return RELAX_OK;
  for (int i = 0; i < sx3; ++i)
    if (structure[i].stressexponent < 0.999999) return RELAX_ERR_POWER;
  int rc;
  if ((rc = has_device())) return rc;
  const long N = (long)sx1 * sx2 * sx3;
  Buf tab, s, e, m, ed, dg, tb;
  std::vector<float> h(3 * sx3);
  for (int i = 0; i < sx3; ++i) {
    h[i] = (float)structure[i].gammadot0;
    h[sx3 + i] = (float)structure[i].stressexponent;
    h[2 * sx3 + i] = (float)structure[i].Gk;
  }
  if ((rc = tab.alloc(h.size() * 4)) || (rc = s.alloc(N * 24)) || (rc = e.alloc(N * 24)) ||
      (rc = m.alloc(N * 24)) || (rc = tb.alloc(4)))
    return rc;
  const bool alias = epsilonik == epsilonikdot;
  if (!alias && (rc = ed.alloc(N * 24))) return rc;
  HIP_TRY(hipMemcpy(tab.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.p, sig, N * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e.p, epsilonik, N * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m.p, moment, N * 24, hipMemcpyHostToDevice));
  const unsigned int inf_bits = 0x7f800000u;
  HIP_TRY(hipMemcpy(tb.p, &inf_bits, 4, hipMemcpyHostToDevice));
  if (dgammadot0) {
    if ((rc = dg.alloc(N * 4))) return rc;
    HIP_TRY(hipMemcpy(dg.p, dgammadot0, N * 4, hipMemcpyHostToDevice));
  }
  Mechanism M{};
  M.gammadot0 = tab.f(); M.power = tab.f() + sx3; M.cohesion = tab.f() + 2 * sx3; M.active = 1;
  float* edp = alias ? e.f() : ed.f();
  k_transient_generic<<<g1d(N), 256>>>(M, (float)mu, sx1, sx2, sx3, s.f(), e.f(), m.f(), edp,
                                       dgammadot0 ? dg.f() : nullptr, (unsigned int*)tb.p);
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(moment, m.p, N * 24, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(epsilonikdot, edp, N * 24, hipMemcpyDeviceToHost));
  if (maxwelltime) {
    unsigned int bits;
    HIP_TRY(hipMemcpy(&bits, tb.p, 4, hipMemcpyDeviceToHost));
    float tm;
    std::memcpy(&tm, &bits, 4);
    tm = std::fmin(tm, (float)*maxwelltime);
    *maxwelltime = std::fmin((double)tm, *maxwelltime);
  }
  return RELAX_OK;
}

int relax_builddgammadot0(int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, double beta, int nz,
                          const relax_weak_t* zones, float* dgammadot0) {
  if (!dgammadot0 || nz < 0 || (nz > 0 && !zones) || sx1 < 1 || sx2 < 1 || sx3 < 1) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  const long N = (long)sx1 * sx2 * sx3;
  Buf z, out;
  if ((rc = out.alloc(N * 4))) return rc;
  if (nz > 0) {
    std::vector<WeakDev> wz(nz);
    for (int i = 0; i < nz; ++i) wz[i] = make_weak(zones[i], beta);
    if ((rc = z.alloc((size_t)nz * sizeof(WeakDev)))) return rc;
    HIP_TRY(hipMemcpy(z.p, wz.data(), (size_t)nz * sizeof(WeakDev), hipMemcpyHostToDevice));
  }
  if ((rc = launch_builddgammadot0(sx1, sx2, sx3, dx1, dx2, dx3, beta, nz, (const WeakDev*)z.p, out.f(), 0)))
    return rc;
  if ((rc = sync_status(RELAX_OK))) return rc;
  HIP_TRY(hipMemcpy(dgammadot0, out.p, N * 4, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

int relax_frictioneigenstress(double x, double y, double z, double L, double W, double strike, double dip,
                              double rake, double beta, const relax_tensor_t* sig,
                              const relax_tensor_layer_t* prestress, double mu, const relax_layer_t* structure,
                              int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,
                              relax_tensor_t* moment, double* maxwelltime, float* vel, int align_data) {
  (void)prestress;
  if (!sig || !structure || !moment) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  const long N = (long)sx1 * sx2 * sx3;
  const long vp = align_data ? sx1 + 2 : sx1;
  const size_t vb = (size_t)vp * sx2 * sx3 * 4;
  std::vector<double> h(4 * sx3);
  for (int i = 0; i < sx3; ++i) {
    h[i] = structure[i].gammadot0;
    h[sx3 + i] = structure[i].stressexponent;
    h[2 * sx3 + i] = structure[i].friction;
    h[3 * sx3 + i] = structure[i].cohesion;
  }
  relax_plane_t pl{x, y, z, /*width*/ L, /*length*/ W, strike, dip, rake, 0, 0};
  const PlaneDev P = make_plane(pl, beta, dx1, dx3);
  Buf tab, pd, s, m, v, tb, stt;
  if ((rc = tab.alloc(h.size() * 8)) || (rc = pd.alloc(sizeof(PlaneDev))) || (rc = s.alloc(N * 24)) ||
      (rc = m.alloc(N * 24)) || (rc = tb.alloc(4)) || (rc = stt.alloc(4)))
    return rc;
  HIP_TRY(hipMemcpy(tab.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(pd.p, &P, sizeof(P), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.p, sig, N * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m.p, moment, N * 24, hipMemcpyHostToDevice));
  const unsigned int inf_bits = 0x7f800000u;
  const int zero = 0;
  HIP_TRY(hipMemcpy(tb.p, &inf_bits, 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(stt.p, &zero, 4, hipMemcpyHostToDevice));
  if (vel) {
    if ((rc = v.alloc(vb))) return rc;
    HIP_TRY(hipMemcpy(v.p, vel, vb, hipMemcpyHostToDevice));
  }
  RheoParams R;
  std::memset(&R, 0, sizeof(R));
  R.mu = (float)mu;
  const double* td = (const double*)tab.p;
  R.fr.v0 = td; R.fr.tauc = td + sx3; R.fr.friction = td + 2 * sx3; R.fr.cohesion = td + 3 * sx3;
  R.fr.planes = (const PlaneDev*)pd.p; R.fr.np = 1; R.fr.active = 1;
  R.n1 = sx1; R.n2 = sx2; R.dx1 = dx1; R.dx2 = dx2; R.dx3 = dx3;
  k_friction_generic<<<g1d(N), 256>>>(R, 0, sx3, s.f(), m.f(), vel ? v.f() : nullptr, vp, (unsigned int*)tb.p,
                                      (int*)stt.p);
  if ((rc = sync_status(RELAX_OK))) return rc;
  int status = 0;
  HIP_TRY(hipMemcpy(&status, stt.p, 4, hipMemcpyDeviceToHost));
  if (status) return status;
  HIP_TRY(hipMemcpy(moment, m.p, N * 24, hipMemcpyDeviceToHost));
  if (vel) HIP_TRY(hipMemcpy(vel, v.p, vb, hipMemcpyDeviceToHost));
  if (maxwelltime) {
    unsigned int bits;
    HIP_TRY(hipMemcpy(&bits, tb.p, 4, hipMemcpyDeviceToHost));
    float tm;
    std::memcpy(&tm, &bits, 4);
    tm = std::fmin(tm, (float)*maxwelltime);
    *maxwelltime = std::fmin((double)tm, *maxwelltime);
  }
  return RELAX_OK;
}

int relax_monitorfriction(double x, double y, double z, double L, double W, double strike, double dip,
                          double rake, double beta, int sx1, int sx2, int sx3, int sx3_alloc, double dx1,
                          double dx2, double dx3, const relax_tensor_t* sig, const relax_tensor_layer_t* prestress,
                          const relax_layer_t* structure, relax_slippatch_t* patch, int px2, int px3) {
  (void)prestress;
  if (!sig || !structure || !patch || px2 < 0 || px3 < 0) return RELAX_ERR_ARG;
  int rc;
  if ((rc = has_device())) return rc;
  Grid g;
  g.n1 = sx1; g.n2 = sx2; g.n3 = sx3; g.h3 = sx3_alloc; g.m1 = sx1 / 2;
  g.H = (long)sx1 * sx2 * sx3_alloc;
  g.dx1 = (float)dx1; g.dx2 = (float)dx2; g.dx3 = (float)dx3;
  const long H = g.H, np = (long)px2 * px3;
  std::vector<double> h(4 * sx3_alloc);
  for (int i = 0; i < sx3_alloc; ++i) {
    h[i] = structure[i].gammadot0;
    h[sx3_alloc + i] = structure[i].stressexponent;
    h[2 * sx3_alloc + i] = structure[i].friction;
    h[3 * sx3_alloc + i] = structure[i].cohesion;
  }
  relax_plane_t pl{x, y, z, /*width*/ L, /*length*/ W, strike, dip, rake, px2, px3};
  const PlaneDev P = make_plane(pl, beta, dx1, dx3);
  Buf tab, aos, soa, pa, sc;
  if ((rc = tab.alloc(h.size() * 8)) || (rc = aos.alloc(H * 24)) || (rc = soa.alloc(H * 24)) ||
      (rc = pa.alloc(np * sizeof(relax_slippatch_t))) || (rc = sc.alloc(sizeof(StepScalars))))
    return rc;
  HIP_TRY(hipMemcpy(tab.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(aos.p, sig, H * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(pa.p, patch, np * sizeof(relax_slippatch_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(sc.p, 0, sizeof(StepScalars)));
  if ((rc = launch_aos_soa(aos.f(), soa.f(), H, true, 0))) return rc;
  const double* td = (const double*)tab.p;
  if ((rc = launch_monitorfriction(P, rake, px2, px3, g, dx1, dx2, dx3, sx3, soa.f(), td, td + sx3_alloc,
                                   td + 2 * sx3_alloc, td + 3 * sx3_alloc, (relax_slippatch_t*)pa.p,
                                   (StepScalars*)sc.p, 0)))
    return rc;
  if ((rc = sync_status(RELAX_OK))) return rc;
  StepScalars s;
  HIP_TRY(hipMemcpy(&s, sc.p, sizeof(s), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(patch, pa.p, np * sizeof(relax_slippatch_t), hipMemcpyDeviceToHost));
  return s.status;
}

int relax_frictionadd(relax_slippatch_t* patch, long npatch, double dt) {
  if (!patch || npatch < 0) return RELAX_ERR_ARG;
  if (npatch == 0) return RELAX_OK;
  int rc;
  if ((rc = has_device())) return rc;
  Buf pa;
  if ((rc = pa.alloc(npatch * sizeof(relax_slippatch_t)))) return rc;
  HIP_TRY(hipMemcpy(pa.p, patch, npatch * sizeof(relax_slippatch_t), hipMemcpyHostToDevice));
  if ((rc = sync_status(launch_frictionadd((relax_slippatch_t*)pa.p, npatch, dt, 0)))) return rc;
  HIP_TRY(hipMemcpy(patch, pa.p, npatch * sizeof(relax_slippatch_t), hipMemcpyDeviceToHost));
  return RELAX_OK;
}
