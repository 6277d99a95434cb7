// fft_lds.h -- batched power-of-two complex FFTs held in LDS.
//
// Stockham autosort (decimation in frequency) with radix-8 (or radix-16)
// stages and one trailing smaller stage (512 = 8*8*8, 256 = 8*8*4,
// 128 = 8*8*2, or with radix 16: 128 = 16*8, 256 = 16*16).
// The transform runs in place: every thread reads all inputs of its
// butterflies into registers, the workgroup synchronises, then every thread
// writes its outputs.
//
// Element j of batch b lives at  buf[(b % B0) * S0 + (b / B0) * S1 + j * JS],
// or, with PAD > 0, at buf[lpad<PAD>(...)] of that index: one spare element
// after every 2^PAD, which spreads the strided writes of a Stockham stage
// (element stride R within a batch) over the LDS banks.
// Twiddles come from a table tw[x] = exp(-2 pi i x / (L * TWS)) in global
// memory; SIGN = -1 is the forward transform, +1 the inverse (unnormalised).
#pragma once
#include "common.h"

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

template <int PAD>
__host__ __device__ __forceinline__ constexpr int lpad(int i) { return PAD ? i + (i >> PAD) : i; }

template <int SIGN>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int idx) {
  float2 w = tw[idx];
  return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

// multiply by sigma*i (sigma = SIGN)
template <int SIGN>
__device__ __forceinline__ float2 mul_si(float2 a) { return SIGN < 0 ? mul_mi(a) : mul_i(a); }

// small DFTs in natural order, y_t = sum_r a_r exp(SIGN 2 pi i r t / R)
template <int SIGN>
__device__ __forceinline__ void dft2(float2* a) {
  const float2 x = a[0], y = a[1];
  a[0] = cadd(x, y);
  a[1] = csub(x, y);
}
template <int SIGN>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = mul_si<SIGN>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}
template <int SIGN>
__device__ __forceinline__ void dft4(float2* a) { dft4<SIGN>(a[0], a[1], a[2], a[3]); }
template <int SIGN>
__device__ __forceinline__ void dft8(float2* a) {
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<SIGN>(e0, e1, e2, e3);
  dft4<SIGN>(o0, o1, o2, o3);
  const float r = 0.70710678118654752440f;
  // w^t o_t with w = exp(SIGN 2 pi i / 8)
  const float2 w1 = make_float2(r * (o1.x - SIGN * o1.y), r * (o1.y + SIGN * o1.x));
  const float2 w2 = mul_si<SIGN>(o2);
  const float2 w3 = make_float2(r * (-o3.x - SIGN * o3.y), r * (-o3.y + SIGN * o3.x));
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, w1); a[5] = csub(e1, w1);
  a[2] = cadd(e2, w2); a[6] = csub(e2, w2);
  a[3] = cadd(e3, w3); a[7] = csub(e3, w3);
}
// 16 = 4 x 4: 4-point DFTs over r2 of a[r1 + 4 r2], the twiddles
// w16^(r1 t1), then 4-point DFTs over r1; output t1 + 4 t2 in natural order.
template <int SIGN>
__device__ __forceinline__ void dft16(float2* a) {
  // cos and sin of 2 pi k / 16, k = 0..9 (the products r1*t1 that occur)
  constexpr float C[10] = {1.f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f, 0.f,
                           -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f, -1.f,
                           -0.92387953251128674f};
  constexpr float S[10] = {0.f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f, 1.f,
                           0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f, 0.f,
                           -0.38268343236508977f};
  float2 b[4][4];
#pragma unroll
  for (int r1 = 0; r1 < 4; ++r1) {
    b[r1][0] = a[r1]; b[r1][1] = a[r1 + 4]; b[r1][2] = a[r1 + 8]; b[r1][3] = a[r1 + 12];
    dft4<SIGN>(b[r1][0], b[r1][1], b[r1][2], b[r1][3]);
#pragma unroll
    for (int t1 = 1; t1 < 4; ++t1) {
      if (r1 == 0) continue;
      const int k = r1 * t1;
      const float2 w = make_float2(C[k], SIGN * S[k]);
      b[r1][t1] = cmul(b[r1][t1], w);
    }
  }
#pragma unroll
  for (int t1 = 0; t1 < 4; ++t1) {
    float2 c0 = b[0][t1], c1 = b[1][t1], c2 = b[2][t1], c3 = b[3][t1];
    dft4<SIGN>(c0, c1, c2, c3);
    a[t1] = c0; a[t1 + 4] = c1; a[t1 + 8] = c2; a[t1 + 12] = c3;
  }
}

template <int R, int SIGN>
__device__ __forceinline__ void dft(float2* a) {
  if (R == 16) dft16<SIGN>(a);
  else if (R == 8) dft8<SIGN>(a);
  else if (R == 4) dft4<SIGN>(a);
  else dft2<SIGN>(a);
}

// One Stockham stage of radix R on sub-transforms of length N (stride L/N).
template <int L, int N, int R, int NB, int NT, int B0, int S0, int S1, int JS, int SIGN, int TWS, bool MAPB,
          int PAD = 0>
__device__ __forceinline__ void lds_stage(float2* buf, const float2* __restrict__ tw, int tid) {
  constexpr int S = L / N;   // interleaved sub-transforms
  constexpr int M = N / R;   // length of the outputs' sub-transforms
  constexpr int Q = L / R;   // butterflies per batch
  constexpr int NBF = NB * Q;
  constexpr int K = (NBF + NT - 1) / NT;
  float2 a[K][R];
  int ob[K], pp[K], qq[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int u = tid + k * NT;
    ob[k] = -1;
    if ((NBF % NT) == 0 || u < NBF) {
      int b, r;
      if (MAPB) { b = u % NB; r = u / NB; } else { r = u % Q; b = u / Q; }
      const int q = r % S, p = r / S;
      const int o = (b % B0) * S0 + (b / B0) * S1;
      ob[k] = o; pp[k] = p; qq[k] = q;
#pragma unroll
      for (int t = 0; t < R; ++t) a[k][t] = buf[lpad<PAD>(o + (q + S * (p + t * M)) * JS)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (ob[k] < 0) continue;
    const int p = pp[k], q = qq[k], o = ob[k];
    dft<R, SIGN>(a[k]);
    if (M > 1) {
      const int step = (L / N) * TWS;
#pragma unroll
      for (int t = 1; t < R; ++t) a[k][t] = cmul(a[k][t], twiddle<SIGN>(tw, p * t * step));
    }
#pragma unroll
    for (int t = 0; t < R; ++t) buf[lpad<PAD>(o + (q + S * (R * p + t)) * JS)] = a[k][t];
  }
  __syncthreads();
}

template <int L, int N, int NB, int NT, int B0, int S0, int S1, int JS, int SIGN, int TWS, bool MAPB, int RMAX,
          int PAD>
struct LdsStages {
  static __device__ __forceinline__ void run(float2* buf, const float2* __restrict__ tw, int tid) {
    constexpr int R = N >= RMAX ? RMAX : N;
    lds_stage<L, N, R, NB, NT, B0, S0, S1, JS, SIGN, TWS, MAPB, PAD>(buf, tw, tid);
    LdsStages<L, N / R, NB, NT, B0, S0, S1, JS, SIGN, TWS, MAPB, RMAX, PAD>::run(buf, tw, tid);
  }
};
template <int L, int NB, int NT, int B0, int S0, int S1, int JS, int SIGN, int TWS, bool MAPB, int RMAX, int PAD>
struct LdsStages<L, 1, NB, NT, B0, S0, S1, JS, SIGN, TWS, MAPB, RMAX, PAD> {
  static __device__ __forceinline__ void run(float2*, const float2* __restrict__, int) {}
};

// MAPB: consecutive threads take consecutive batches (use when S0 == 1);
// otherwise consecutive threads take consecutive butterflies of one batch.
// RMAX: the largest radix (8, or 16 for fewer stages and LDS passes).
template <int L, int NB, int NT, int B0, int S0, int S1, int JS, int SIGN, int TWS, bool MAPB, int RMAX = 8,
          int PAD = 0>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* __restrict__ tw, int tid) {
  LdsStages<L, L, NB, NT, B0, S0, S1, JS, SIGN, TWS, MAPB, RMAX, PAD>::run(buf, tw, tid);
}
