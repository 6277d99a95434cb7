// fft_lds.h -- batched power-of-two complex FFTs held in LDS.
//
// Stockham autosort (decimation in frequency), radix 4 with one trailing
// radix-2 stage when log2(L) is odd.  The transform runs in place: every
// thread reads all inputs of its butterflies into registers, the workgroup
// synchronises, then every thread writes its outputs.
//
// Element j of batch b lives at  buf[(b % B0) * S0 + (b / B0) * S1 + j * JS].
// Twiddles come from a table tw[x] = exp(-2 pi i x / (L * TWS)) in global
// memory; SIGN = -1 is the forward transform, +1 the inverse (unnormalised).
#pragma once
#include "common.h"

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

template <int SIGN>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int idx) {
  float2 w = tw[idx];
  return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

// MAPB: consecutive threads take consecutive batches (use when S0 == 1);
// otherwise consecutive threads take consecutive butterflies of one batch.
template <int L, int NB, int NT, int B0, int S0, int S1, int JS, int SIGN, int TWS, bool MAPB>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* __restrict__ tw, int tid) {
  constexpr int LOG2L = ilog2c(L);
  constexpr int N4 = LOG2L / 2;
  constexpr bool HAS_R2 = (LOG2L & 1) != 0;
  constexpr int Q4 = L / 4;  // butterflies per batch in a radix-4 stage
  constexpr int NBF4 = NB * Q4;
  constexpr int K4 = (NBF4 + NT - 1) / NT;

#pragma unroll
  for (int st = 0; st < N4; ++st) {
    const int n = L >> (2 * st);  // current sub-transform length
    const int s = 1 << (2 * st);  // number of interleaved sub-transforms
    const int m = n >> 2;
    float2 a[K4][4];
    int ob[K4], pp[K4], qq[K4];
#pragma unroll
    for (int k = 0; k < K4; ++k) {
      const int u = tid + k * NT;
      ob[k] = -1;
      if ((NBF4 % NT) == 0 || u < NBF4) {
        int b, r;
        if (MAPB) { b = u % NB; r = u / NB; } else { r = u % Q4; b = u / Q4; }
        const int q = r & (s - 1), p = r >> (2 * st);
        const int o = (b % B0) * S0 + (b / B0) * S1;
        ob[k] = o; pp[k] = p; qq[k] = q;
#pragma unroll
        for (int t = 0; t < 4; ++t) a[k][t] = buf[o + (q + s * (p + t * m)) * JS];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K4; ++k) {
      if (ob[k] < 0) continue;
      const int p = pp[k], q = qq[k], o = ob[k];
      const float2 apc = cadd(a[k][0], a[k][2]), amc = csub(a[k][0], a[k][2]);
      const float2 bpd = cadd(a[k][1], a[k][3]), bmd = csub(a[k][1], a[k][3]);
      const float2 jb = SIGN < 0 ? mul_mi(bmd) : mul_i(bmd);
      float2 y0 = cadd(apc, bpd);
      float2 y1 = cadd(amc, jb);
      float2 y2 = csub(apc, bpd);
      float2 y3 = csub(amc, jb);
      if (st + 1 < N4 || HAS_R2) {  // the last radix-4 stage of a pure radix-4 transform has p == 0
        const int step = (L / n) * TWS;
        y1 = cmul(y1, twiddle<SIGN>(tw, p * step));
        y2 = cmul(y2, twiddle<SIGN>(tw, 2 * p * step));
        y3 = cmul(y3, twiddle<SIGN>(tw, 3 * p * step));
      }
      buf[o + (q + s * (4 * p + 0)) * JS] = y0;
      buf[o + (q + s * (4 * p + 1)) * JS] = y1;
      buf[o + (q + s * (4 * p + 2)) * JS] = y2;
      buf[o + (q + s * (4 * p + 3)) * JS] = y3;
    }
    __syncthreads();
  }
  if (HAS_R2) {
    constexpr int S = L / 2;  // radix-2 stage: n == 2, p == 0
    constexpr int NBF2 = NB * S;
    constexpr int K2 = (NBF2 + NT - 1) / NT;
#pragma unroll
    for (int k = 0; k < K2; ++k) {
      const int u = tid + k * NT;
      if ((NBF2 % NT) == 0 || u < NBF2) {
        int b, q;
        if (MAPB) { b = u % NB; q = u / NB; } else { q = u % S; b = u / S; }
        const int o = (b % B0) * S0 + (b / B0) * S1;
        const float2 x0 = buf[o + q * JS], x1 = buf[o + (q + S) * JS];
        buf[o + q * JS] = cadd(x0, x1);
        buf[o + (q + S) * JS] = csub(x0, x1);
      }
    }
    __syncthreads();
  }
}
