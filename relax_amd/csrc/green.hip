// green.hip -- Fourier-domain elastic Green's function with the Cowling
// surface correction (green.f90:51-101 elasticresponse, :327-377
// surfacetractioncowling, :520-612 cerrutimodified, :1008-1129
// greenfunctioncowling; the CUDA kernels cuElasticResKernel, cuSurfaceKernel,
// scaleAndSub and cuCerrutiModKernel of cugreen.cu are replaced).
//
// The hot-path kernel is k_green_columns: one workgroup owns TX (k1,k2)
// columns of all three displacement components.  Because the body force lives
// only in the upper half of the FFT grid and only the upper half of the
// displacement is kept (relax.f90:682, :833-835), the x3 transforms are
// pruned: the even and odd k3 halves are two half-length FFTs of the input
// and of the input times w^i3, and the inverse keeps only the upper half.
// Between them the workgroup applies the Green's function, reduces the
// surface traction over k3 in a fixed order (no atomics: the result is
// bitwise reproducible), forms p = t - p and adds the Cerruti correction.
#include <cmath>

#include "fft_lds.h"
#include "kernels.h"

namespace {

struct ColumnConsts {
  float ratio1, ratio2p1;     // elasticresponse: (l+m)/(l+2m)/m/(2pi)^2, 1 + m/(l+m)
  float mu2pi, mod2pi, lam2pi, grav;  // traction: 2 pi mu, 2 pi (l+2m), 2 pi l, 2 mu alpha gamma
  float alpha, inv_alpha;
  float inv_n3dx3;
  float ik1, ik2, ik3;        // 1/(n dx) per axis
};

__host__ ColumnConsts make_consts(const GreenConst& gc) {
  ColumnConsts c;
  const double l = gc.lambda, m = gc.mu;
  const double alpha = (l + m) / (l + 2.0 * m);
  c.ratio1 = (float)((l + m) / (l + 2.0 * m) / m / (RELAX_PI2 * RELAX_PI2));
  c.ratio2p1 = (float)(1.0 + m / (l + m));
  c.mu2pi = (float)(RELAX_PI2 * m);
  c.mod2pi = (float)(RELAX_PI2 * (l + 2.0 * m));
  c.lam2pi = (float)(RELAX_PI2 * l);
  c.grav = (float)(2.0 * m * alpha * gc.gamma);
  c.alpha = (float)alpha;
  c.inv_alpha = (float)(1.0 / alpha);
  c.inv_n3dx3 = (float)(gc.n3 * gc.dx3);  // the reference divides by REAL(sx3*dx3,4)
  c.ik1 = (float)(1.0 / (gc.n1 * gc.dx1));
  c.ik2 = (float)(1.0 / (gc.n2 * gc.dx2));
  c.ik3 = (float)(1.0 / (gc.n3 * gc.dx3));
  return c;
}

// elasticresponse at one wavenumber (green.f90:63-86).  The reference's
// matrix [(|k|^2 (1 + m/(l+m)) - k_i^2) delta_ij - k_i k_j (1 - delta_ij)]
// is applied as s c - k (k.c) with s = |k|^2 (1 + m/(l+m)).
__device__ __forceinline__ void elastic_point(const ColumnConsts& C, float k1, float k2, float k3, float2& c1,
                                              float2& c2, float2& c3) {
  const float r2 = k1 * k1 + k2 * k2 + k3 * k3;
  // the hardware reciprocal (1 ulp) in place of a correctly rounded
  // division: the column pass is bound by its FP32 issue rate
  const float denom = C.ratio1 * __builtin_amdgcn_rcpf(r2 * r2);
  const float s = r2 * C.ratio2p1;
  const float2 kc = make_float2(k1 * c1.x + k2 * c2.x + k3 * c3.x, k1 * c1.y + k2 * c2.y + k3 * c3.y);
  c1 = make_float2((s * c1.x - k1 * kc.x) * denom, (s * c1.y - k1 * kc.y) * denom);
  c2 = make_float2((s * c2.x - k2 * kc.x) * denom, (s * c2.y - k2 * kc.y) * denom);
  c3 = make_float2((s * c3.x - k3 * kc.x) * denom, (s * c3.y - k3 * kc.y) * denom);
}

// summand of surfacetractioncowling (green.f90:357-359)
__device__ __forceinline__ void traction_point(const ColumnConsts& C, float k1, float k2, float k3, float2 c1,
                                               float2 c2, float2 c3, float2& s1, float2& s2, float2& s3) {
  // i 2 pi mu (k3 c1 + k1 c3)
  const float2 a = make_float2(k3 * c1.x + k1 * c3.x, k3 * c1.y + k1 * c3.y);
  s1 = make_float2(-C.mu2pi * a.y, C.mu2pi * a.x);
  const float2 b = make_float2(k3 * c2.x + k2 * c3.x, k3 * c2.y + k2 * c3.y);
  s2 = make_float2(-C.mu2pi * b.y, C.mu2pi * b.x);
  // i 2 pi ((l+2m) k3 c3 + l (k1 c1 + k2 c2)) - grav c3
  const float2 d = make_float2(C.mod2pi * k3 * c3.x + C.lam2pi * (k1 * c1.x + k2 * c2.x),
                               C.mod2pi * k3 * c3.y + C.lam2pi * (k1 * c1.y + k2 * c2.y));
  s3 = make_float2(-d.y - C.grav * c3.x, d.x - C.grav * c3.y);
}

// Per-column part of cerrutisolmodified (green.f90:571-606).  With
// fi = 2 beta / ((2 pi k3)^2 + beta^2) and kf = 2 pi k3 fi, every term the
// reference adds to u (:599-606) is a column constant times fi or fi kf:
//   u_c += fi A_c + fi kf B_c,
// with b3 and P_c = p_c / (2 mu beta^3) (B1,2 of :596-598 without fi), and
// BK = k1 P1 + k2 P2:
//   A_c = i k_c a b 2pi (1 - 1/a) b3 - 2 b^2 P_c + k_c a (2pi)^2 BK   (c = 1, 2)
//   B_c = k_c a b 2pi b3 - i k_c a (2pi)^2 BK
//   A_3 = -b^2 b3,   B_3 = i a b^2 b3 + a b 2pi BK
// (a = alpha, b = beta).  beta, b3 and 1/(2 mu beta^3) are formed in FP64
// and rounded; the constants are then combined in FP32 (few live registers
// while the column's spectrum is held in them).
struct CerrutiCol {
  float beta, beta2;
  float2 A[3], B[3];
  int zero;                // k1 == k2 == 0: no correction
};

__device__ __forceinline__ CerrutiCol cerruti_column(const ColumnConsts& C, double mu, double alpha, double gamma,
                                                     float k1, float k2, float2 p1, float2 p2, float2 p3) {
  CerrutiCol cc;
  cc.zero = (k1 == 0.f && k2 == 0.f);
  if (cc.zero) {
    cc.beta = cc.beta2 = 0.f;
    for (int c = 0; c < 3; ++c) cc.A[c] = cc.B[c] = make_float2(0.f, 0.f);
    return cc;
  }
  const double dk1 = k1, dk2 = k2;
  const double beta = RELAX_PI2 * sqrt(dk1 * dk1 + dk2 * dk2);
  const double h = gamma / beta;
  const double den = 1.0 / (2.0 * alpha * mu * beta * beta * beta * beta * (1.0 + h));
  // b3 = (beta p3 - i (1-alpha) 2 pi (k1 p1 + k2 p2)) / (2 alpha mu beta^4 (1+h))
  const double qx = RELAX_PI2 * (dk1 * p1.x + dk2 * p2.x), qy = RELAX_PI2 * (dk1 * p1.y + dk2 * p2.y);
  const double b3x = (beta * p3.x + (1.0 - alpha) * qy) * den;
  const double b3y = (beta * p3.y - (1.0 - alpha) * qx) * den;
  const float cb1 = (float)(1.0 / (2.0 * mu * beta * beta * beta));
  const float fb3x = (float)b3x, fb3y = (float)b3y;
  const float fbeta = (float)beta, b2 = (float)(beta * beta);
  const float P1x = p1.x * cb1, P1y = p1.y * cb1, P2x = p2.x * cb1, P2y = p2.y * cb1;
  const float BKx = k1 * P1x + k2 * P2x, BKy = k1 * P1y + k2 * P2y;
  const float ab2pi = C.alpha * fbeta * (float)RELAX_PI2, a4pi2 = C.alpha * (float)(RELAX_PI2 * RELAX_PI2);
  const float r = 1.f - C.inv_alpha;
  const float kk[2] = {k1, k2}, Px[2] = {P1x, P2x}, Py[2] = {P1y, P2y};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    // i k ab2pi r b3 = k ab2pi r (-b3y, b3x)
    const float kr = kk[c] * ab2pi * r, ka = kk[c] * a4pi2, kb = kk[c] * ab2pi;
    cc.A[c] = make_float2(-kr * fb3y - 2.f * b2 * Px[c] + ka * BKx, kr * fb3x - 2.f * b2 * Py[c] + ka * BKy);
    // -i k a4pi2 BK = k a4pi2 (BKy, -BKx)
    cc.B[c] = make_float2(kb * fb3x + ka * BKy, kb * fb3y - ka * BKx);
  }
  cc.A[2] = make_float2(-b2 * fb3x, -b2 * fb3y);
  // i alpha b^2 b3 + ab2pi BK
  const float ab2 = C.alpha * b2;
  cc.B[2] = make_float2(-ab2 * fb3y + ab2pi * BKx, ab2 * fb3x + ab2pi * BKy);
  cc.beta = fbeta;
  cc.beta2 = b2;
  return cc;
}

// The k3-dependent part of cerrutisolmodified (green.f90:596-606).
__device__ __forceinline__ void cerruti_point(const ColumnConsts& C, const CerrutiCol& cc, float2 p1, float2 p2,
                                              float k3, float2& u1, float2& u2, float2& u3) {
  (void)C; (void)p1; (void)p2;
  if (cc.zero) return;
  const float tk3 = (float)RELAX_PI2 * k3;
  const float fi = 2.f * cc.beta * __builtin_amdgcn_rcpf(tk3 * tk3 + cc.beta2);
  const float fk = fi * (tk3 * fi);  // fi kf
  u1 = make_float2(u1.x + (fi * cc.A[0].x + fk * cc.B[0].x), u1.y + (fi * cc.A[0].y + fk * cc.B[0].y));
  u2 = make_float2(u2.x + (fi * cc.A[1].x + fk * cc.B[1].x), u2.y + (fi * cc.A[1].y + fk * cc.B[1].y));
  u3 = make_float2(u3.x + (fi * cc.A[2].x + fk * cc.B[2].x), u3.y + (fi * cc.A[2].y + fk * cc.B[2].y));
}

}  // namespace

// ---------------------------------------------------- fused column pass ----
// Addresses and wavenumbers of one tile: TX columns of the spectrum (or of
// its Nyquist column), all depths, all three components.
struct GreenTileAddr {
  float2* base;
  const float2* tbase;
  long cs, zs, tcs;
  int k1i, k2i;
};

template <int TX>
__device__ __forceinline__ GreenTileAddr green_tile(int bid, int tx, int n1, int n2, int h3, int l2, float2* W,
                                                    float2* NQ, const float2* TM, const float2* TN) {
  GreenTileAddr A;
  const int m1 = n1 / 2;
  const int nmain = (m1 / TX) * n2;
  if (bid < nmain) {
    // spectrum row r of W; its y frequency is r, or 8*(r % l2) + r / l2 in
    // the split-y order; the traction planes are in natural order
    const int r = bid / (m1 / TX), k10 = (bid % (m1 / TX)) * TX;
    const int k2 = l2 ? 8 * (r % l2) + r / l2 : r;
    A.base = W + (long)r * m1 + k10;
    A.cs = (long)h3 * n2 * m1; A.zs = (long)n2 * m1;
    A.tbase = TM + (long)k2 * m1 + k10; A.tcs = (long)n2 * m1;
    A.k1i = k10 + tx; A.k2i = k2;
  } else {
    const int k20 = (bid - nmain) * TX;
    A.base = NQ + k20;
    A.cs = (long)h3 * n2; A.zs = n2;
    A.tbase = TN + k20; A.tcs = n2;
    A.k1i = m1; A.k2i = k20 + tx;
  }
  return A;
}

// RM: the largest radix of the column transforms (8 or 16).  One tile per
// workgroup.
template <int H3, int TX, int NT, int RM>
__global__ __launch_bounds__(NT) void k_green_columns(ColumnConsts C, double mu, double alpha, double gamma,
                                                      int n1, int n2, int n3, MPtr<float2> W_m, MPtr<float2> NQ_m,
                                                      MPtr<const float2> TM_m, MPtr<const float2> TN_m,
                                                      const float2* __restrict__ tw3, int l2) {
  // model blockIdx.y of the launch
  float2* W = W_m.at(blockIdx.y);
  float2* NQ = NQ_m.at(blockIdx.y);
  const float2* __restrict__ TM = TM_m.at(blockIdx.y);
  const float2* __restrict__ TN = TN_m.at(blockIdx.y);
  static_assert(NT % 64 == 0 && 64 % TX == 0, "a thread keeps one column; whole waves");
  constexpr int NW = NT / 64;
  constexpr int NE = 3 * H3 * TX;  // input elements of a tile
  // LDS: six sequences (component c, half) of H3 x TX complex.  The stride
  // between sequences is padded by 12 float2 so that the lanes of a wave,
  // which walk batches across sequences in the FFT stages, spread over the
  // LDS banks instead of landing on the same 16.
  constexpr int SB = H3 * TX + 12;
  __shared__ float2 s[6 * SB];
  __shared__ float2 red[3][NW][TX];
  __shared__ float2 pv[3][TX];
  __shared__ float2 stw[2 * H3];  // the n3-point twiddles, staged in LDS
  const int tid = threadIdx.x;
  const int tx = tid % TX;
  const int h3 = H3;
  for (int e = tid; e < 2 * H3; e += NT) stw[e] = tw3[e];

  // transform, Green + traction, p = t - p, Cerruti, inverse, store
  auto body = [&](const GreenTileAddr& A) {
    const float k1 = (float)A.k1i * C.ik1;
    const float k2 = wavenumber23f(A.k2i, n2, C.ik2);
    lds_fft<H3, 6 * TX, NT, TX, 1, SB, TX, -1, 2, true, RM>(s, stw, tid);

    // Green's function and the traction partial sums of this thread's column
    float2 ps1 = make_float2(0.f, 0.f), ps2 = ps1, ps3 = ps1;
    for (int e = tid; e < 2 * H3 * TX; e += NT) {
      const int slot = e / TX, half = slot / H3, m = slot % H3;
      const int k3i = 2 * m + half;
      const float k3 = wavenumber23f(k3i, n3, C.ik3);
      float2& c1 = s[(0 + half) * SB + m * TX + tx];
      float2& c2 = s[(2 + half) * SB + m * TX + tx];
      float2& c3 = s[(4 + half) * SB + m * TX + tx];
      float2 a1 = c1, a2 = c2, a3 = c3;
      if (A.k1i == 0 && A.k2i == 0 && k3i == 0) {
        a1 = a2 = a3 = make_float2(0.f, 0.f);
      } else {
        elastic_point(C, k1, k2, k3, a1, a2, a3);
      }
      c1 = a1; c2 = a2; c3 = a3;
      float2 s1, s2, s3;
      traction_point(C, k1, k2, k3, a1, a2, a3, s1, s2, s3);
      ps1 = cadd(ps1, s1); ps2 = cadd(ps2, s2); ps3 = cadd(ps3, s3);
    }
    // reduce over the lanes of the wave that hold the same column (fixed
    // order: bitwise reproducible), then over the waves
#pragma unroll
    for (int off = 32; off >= TX; off >>= 1) {
      ps1.x += __shfl_xor(ps1.x, off); ps1.y += __shfl_xor(ps1.y, off);
      ps2.x += __shfl_xor(ps2.x, off); ps2.y += __shfl_xor(ps2.y, off);
      ps3.x += __shfl_xor(ps3.x, off); ps3.y += __shfl_xor(ps3.y, off);
    }
    if ((tid & 63) < TX) {
      red[0][tid >> 6][tx] = ps1; red[1][tid >> 6][tx] = ps2; red[2][tid >> 6][tx] = ps3;
    }
    __syncthreads();
    if (tid < 3 * TX) {
      const int c = tid / TX, t = tid % TX;
      float2 acc = red[c][0][t];
      for (int w = 1; w < NW; ++w) acc = cadd(acc, red[c][w][t]);
      const float2 tt = gload(A.tbase + c * A.tcs + t);
      // p = t - p / (sx3 dx3)   (green.f90:373-375, :1082-1084)
      pv[c][t] = make_float2(tt.x - acc.x / C.inv_n3dx3, tt.y - acc.y / C.inv_n3dx3);
    }
    __syncthreads();
    {
      const float2 p1 = pv[0][tx], p2 = pv[1][tx], p3 = pv[2][tx];
      const CerrutiCol cc = cerruti_column(C, mu, alpha, gamma, k1, k2, p1, p2, p3);
      for (int e = tid; e < 2 * H3 * TX; e += NT) {
        const int slot = e / TX, half = slot / H3, m = slot % H3;
        const float k3 = wavenumber23f(2 * m + half, n3, C.ik3);
        float2& c1 = s[(0 + half) * SB + m * TX + tx];
        float2& c2 = s[(2 + half) * SB + m * TX + tx];
        float2& c3 = s[(4 + half) * SB + m * TX + tx];
        float2 a1 = c1, a2 = c2, a3 = c3;
        cerruti_point(C, cc, p1, p2, k3, a1, a2, a3);
        c1 = a1; c2 = a2; c3 = a3;
      }
    }
    __syncthreads();
    lds_fft<H3, 6 * TX, NT, TX, 1, SB, TX, +1, 2, true, RM>(s, stw, tid);
    // upper half of the inverse: y[n] = E'[n] + w^-n O'[n]
    for (int e = tid; e < NE; e += NT) {
      const int t = e % TX, z = (e / TX) % H3, c = e / (TX * H3);
      const float2 ev = s[(2 * c) * SB + z * TX + t];
      const float2 od = s[(2 * c + 1) * SB + z * TX + t];
      gstore(A.base + c * A.cs + z * A.zs + t, cadd(ev, cmul(od, cconj(stw[z]))));
    }
  };

  const GreenTileAddr A = green_tile<TX>(xcd_tile(blockIdx.x, gridDim.x), tx, n1, n2, h3, l2, W, NQ, TM, TN);
  // load the upper half and its w^i3-shifted copy
  for (int e = tid; e < NE; e += NT) {
    const int t = e % TX, z = (e / TX) % H3, c = e / (TX * H3);
    const float2 x = gload(A.base + c * A.cs + z * A.zs + t);
    s[(2 * c) * SB + z * TX + t] = x;
    s[(2 * c + 1) * SB + z * TX + t] = cmul(x, tw3[z]);
  }
  __syncthreads();
  body(A);
}

// Register-resident form of the column pass for H3 = 16 Q (Q = 2, 4, 8):
// 16 columns per workgroup, so each depth row of a tile is 128 contiguous
// bytes.  The H3-point transforms are split as 16 x Q:
//   z = Q m + j,  k = k1 + 16 k2:
//   step A (thread c, j, column): DFT16 over m in registers, twiddle
//     w_H3^(j k1), through LDS ([c][j][k1][column]);
//   step B (thread half, k1, column): DFT_Q over j for the three components
//     in registers -> X[k1 + 16 k2].
// The even (E) and odd (O: x w_n3^z) halves go through the one 52 KB LDS
// buffer in turn; E ends in threads 0..255 and O in 256..511, which then
// apply the Green's function and the traction sums to their own 3 Q
// frequencies, reduce the sums per column, and run the inverse the same way
// back (B^-1 then A^-1, E then O) to y = E' + w^-z O'.
//
// The kernel is bound by its vector instruction issue, so:
//  - the tile is read and written with buffer instructions: 32-bit offsets
//    from the tile's base, with the depth stride in the scalar offset, so a
//    load or store costs no vector address arithmetic (the work spectrum of
//    a grid of this kernel, H3 <= 128 and n1, n2 <= 1024, is under 4 GiB);
//  - the twiddles of the split, w_H3^(a b) for a, b < 16, come from an LDS
//    table whose row is the thread's index, read at constant offsets;
//  - the traction sum of surfacetractioncowling (green.f90:357-359) is
//    linear in u, so each thread sums k3 u1, k3 u2, k3 u3, u3 and
//    k1 u1 + k2 u2 over its frequencies (12 operations per frequency) and
//    the column's traction is formed from the five sums once.
template <int H3>
__global__ __launch_bounds__(512, 4) void k_green_reg(ColumnConsts C, double mu, double alpha, double gamma, int n1,
                                                   int n2, int n3, MPtr<float2> W_m, MPtr<float2> NQ_m,
                                                   MPtr<const float2> TM_m, MPtr<const float2> TN_m,
                                                   const float2* __restrict__ tw3, int l2) {
  constexpr int TX = 16, NT = 512, Q = H3 / 16;
  static_assert(Q == 2 || Q == 4 || Q == 8, "H3 = 32, 64 or 128");
  constexpr int NA = 3 * Q * TX;  // step-A threads
  constexpr int NB = 16 * TX;     // step-B threads per half
  constexpr int RS = 16 * TX + 16;  // float2 per (c, j) group (padded: j groups on alternate banks)
  constexpr int NW = NT / 64;
  constexpr unsigned RSRC3 = 0x00020000u;  // raw buffer, 32-bit data format
  float2* W = W_m.at(blockIdx.y);
  float2* NQ = NQ_m.at(blockIdx.y);
  const float2* __restrict__ TM = TM_m.at(blockIdx.y);
  const float2* __restrict__ TN = TN_m.at(blockIdx.y);
  __shared__ float2 buf[3 * Q * RS];
  __shared__ float2 stw[2 * H3];
  __shared__ float2 tws[16 * 16];  // w_H3^(a b), row a
  // the per-column reductions reuse the transpose buffer
  float2(*red)[NW][TX] = reinterpret_cast<float2(*)[NW][TX]>(buf);
  float2(*pv)[TX] = reinterpret_cast<float2(*)[TX]>(buf + 5 * NW * TX);
  // the Cerruti constants of the tile's columns follow pv
  CerrutiCol* ccs = reinterpret_cast<CerrutiCol*>(buf + 5 * NW * TX + 3 * TX);
  static_assert((5 * NW * TX + 3 * TX) * sizeof(float2) + TX * sizeof(CerrutiCol) <= 3 * Q * RS * sizeof(float2),
                "reductions and column constants fit the buffer");

  const int tid = threadIdx.x;
  const int acol = tid % TX, aj = (tid / TX) % Q, ac = tid / (TX * Q);  // step A
  const int bcol = tid % TX, bk1 = (tid / TX) % 16, half = tid / NB;    // step B
  const bool isA = tid < NA;
  const GreenTileAddr A = green_tile<TX>(xcd_tile(blockIdx.x, gridDim.x), bcol, n1, n2, H3, l2, W, NQ, TM, TN);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(A.base, (short)0, 0x7fffffff, RSRC3);
  const unsigned voff = (unsigned)((ac * A.cs + aj * A.zs + acol) * (long)sizeof(float2));
  const unsigned zstep = (unsigned)(Q * A.zs * (long)sizeof(float2));  // uniform: the scalar offset

  // step A of one half: DFT16, twiddle, into the buffer.  The odd half
  // loads the column again (from L2) rather than keeping it in registers;
  // the loads are issued right after the even half's step A, so that they
  // complete under the even half's step B.
  float2 x[16];
  auto load = [&]() {
#pragma unroll
    for (int m = 0; m < 16; ++m)
      x[m] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, m * zstep, 0));
  };
  if (isA) load();
  for (int e = tid; e < 2 * H3; e += NT) stw[e] = tw3[e];
  if (tid < 256) tws[tid] = tw3[2 * (((tid >> 4) * (tid & 15)) % H3)];
  __syncthreads();

  auto stepA = [&](bool odd) {
    if (!isA) return;
    float2* y = x;
    if (odd) {
#pragma unroll
      for (int m = 0; m < 16; ++m) y[m] = cmul(x[m], stw[Q * m + aj]);
    }
    dft16<-1>(y);
    float2* row = buf + (ac * Q + aj) * RS + acol;
    const float2* tw = tws + 16 * aj;
    row[0] = y[0];
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) row[k1 * TX] = cmul(y[k1], tw[k1]);
  };
  // step B: this thread's 3 Q frequencies of its half
  float2 X[3][Q];
  const float2* twb = tws + 16 * bk1;
  auto stepB = [&]() {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int j = 0; j < Q; ++j) X[c][j] = buf[(c * Q + j) * RS + bk1 * TX + bcol];
      dft<Q, -1>(X[c]);
    }
  };
  stepA(false);
  if (isA) load();  // the odd half's input, in flight during the even half's step B
  __syncthreads();
  if (half == 0) stepB();
  __syncthreads();
  stepA(true);
  __syncthreads();
  if (half == 1) stepB();

  // Green's function and the sums of k3 u and u over this thread's
  // frequencies
  const float k1w = (float)A.k1i * C.ik1;
  const float k2w = wavenumber23f(A.k2i, n2, C.ik2);
  const float kh2 = k1w * k1w + k2w * k2w;
  // the traction sums: k3 u1, k3 u2, k3 u3, u3 and k1 u1 + k2 u2
  float2 S5[5];
#pragma unroll
  for (int a = 0; a < 5; ++a) S5[a] = make_float2(0.f, 0.f);
  const bool origin = A.k1i == 0 && A.k2i == 0 && bk1 == 0 && half == 0;
#pragma unroll
  for (int k2 = 0; k2 < Q; ++k2) {
    const float k3 = wavenumber23f(2 * (bk1 + 16 * k2) + half, n3, C.ik3);
    float2 a1 = X[0][k2], a2 = X[1][k2], a3 = X[2][k2];
    // elasticresponse (green.f90:63-86) as s c - k (k.c); the k = 0
    // frequency is zeroed (:96-99)
    const float r2 = fmaf(k3, k3, kh2);
    const float denom = C.ratio1 * __builtin_amdgcn_rcpf(r2 * r2);
    const float s = r2 * C.ratio2p1;
    const float2 kc = make_float2(k1w * a1.x + k2w * a2.x + k3 * a3.x, k1w * a1.y + k2w * a2.y + k3 * a3.y);
    a1 = make_float2((s * a1.x - k1w * kc.x) * denom, (s * a1.y - k1w * kc.y) * denom);
    a2 = make_float2((s * a2.x - k2w * kc.x) * denom, (s * a2.y - k2w * kc.y) * denom);
    a3 = make_float2((s * a3.x - k3 * kc.x) * denom, (s * a3.y - k3 * kc.y) * denom);
    if (origin && k2 == 0) a1 = a2 = a3 = make_float2(0.f, 0.f);
    X[0][k2] = a1; X[1][k2] = a2; X[2][k2] = a3;
    S5[0] = make_float2(fmaf(k3, a1.x, S5[0].x), fmaf(k3, a1.y, S5[0].y));
    S5[1] = make_float2(fmaf(k3, a2.x, S5[1].x), fmaf(k3, a2.y, S5[1].y));
    S5[2] = make_float2(fmaf(k3, a3.x, S5[2].x), fmaf(k3, a3.y, S5[2].y));
    S5[3] = cadd(S5[3], a3);
    S5[4] = make_float2(fmaf(k1w, a1.x, fmaf(k2w, a2.x, S5[4].x)), fmaf(k1w, a1.y, fmaf(k2w, a2.y, S5[4].y)));
  }
  // reduce over the threads of a column (fixed order): lanes 16 apart in a
  // wave, then the waves
#pragma unroll
  for (int off = 32; off >= TX; off >>= 1) {
#pragma unroll
    for (int a = 0; a < 5; ++a) {
      S5[a].x += __shfl_xor(S5[a].x, off);
      S5[a].y += __shfl_xor(S5[a].y, off);
    }
  }
  __syncthreads();  // the last step B has read the buffer
  if ((tid & 63) < TX) {
#pragma unroll
    for (int a = 0; a < 5; ++a) red[a][tid >> 6][bcol] = S5[a];
  }
  __syncthreads();
  if (tid < 3 * TX) {
    const int c = tid / TX, t = tid % TX;
    // this column's wavenumbers (thread t of the first wave holds column t)
    const float q1 = __shfl(k1w, t), q2 = __shfl(k2w, t);
    float2 S[5];
#pragma unroll
    for (int a = 0; a < 5; ++a) {
      S[a] = red[a][0][t];
      for (int w = 1; w < NW; ++w) S[a] = cadd(S[a], red[a][w][t]);
    }
    // sum_k3 of the traction summands (green.f90:357-359):
    //   i 2 pi mu (k3 u1 + k1 u3), i 2 pi mu (k3 u2 + k2 u3),
    //   i 2 pi ((l+2m) k3 u3 + l (k1 u1 + k2 u2)) - 2 mu alpha gamma u3
    float2 acc;
    if (c < 2) {
      // (S[c] with a constant index: a run-time index would move S to LDS)
      const float qc = c == 0 ? q1 : q2;
      const float2 Sc = c == 0 ? S[0] : S[1];
      const float2 a = make_float2(Sc.x + qc * S[3].x, Sc.y + qc * S[3].y);
      acc = make_float2(-C.mu2pi * a.y, C.mu2pi * a.x);
    } else {
      const float2 d = make_float2(C.mod2pi * S[2].x + C.lam2pi * S[4].x, C.mod2pi * S[2].y + C.lam2pi * S[4].y);
      acc = make_float2(-d.y - C.grav * S[3].x, d.x - C.grav * S[3].y);
    }
    const float2 tt = gload(A.tbase + c * A.tcs + t);
    // p = t - p / (sx3 dx3)   (green.f90:373-375, :1082-1084)
    pv[c][t] = make_float2(tt.x - acc.x / C.inv_n3dx3, tt.y - acc.y / C.inv_n3dx3);
  }
  __syncthreads();
  // one thread per column forms its constants (FP64 divisions and square
  // root), the others read them
  if (tid < TX) ccs[tid] = cerruti_column(C, mu, alpha, gamma, k1w, k2w, pv[0][tid], pv[1][tid], pv[2][tid]);
  __syncthreads();
  {
    const CerrutiCol cc = ccs[bcol];
    const float2 p1 = pv[0][bcol], p2 = pv[1][bcol];
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2)
      cerruti_point(C, cc, p1, p2, wavenumber23f(2 * (bk1 + 16 * k2) + half, n3, C.ik3), X[0][k2], X[1][k2],
                    X[2][k2]);
  }
  __syncthreads();  // pv and the constants are read before the buffer is refilled

  // inverse: B^-1 into the buffer, A^-1 out of it
  auto stepBinv = [&]() {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      dft<Q, +1>(X[c]);
      buf[(c * Q) * RS + bk1 * TX + bcol] = X[c][0];
#pragma unroll
      for (int j = 1; j < Q; ++j) buf[(c * Q + j) * RS + bk1 * TX + bcol] = cmul(X[c][j], cconj(twb[j]));
    }
  };
  auto stepAinv = [&](float2* y) {
    const float2* row = buf + (ac * Q + aj) * RS + acol;
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) y[k1] = row[k1 * TX];
    dft16<+1>(y);
  };
  if (half == 0) stepBinv();
  __syncthreads();
  if (isA) stepAinv(x);  // E'
  __syncthreads();
  if (half == 1) stepBinv();
  __syncthreads();
  if (isA) {
    float2 o[16];
    stepAinv(o);  // O'
    typedef unsigned u2v __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const float2 v = cadd(x[m], cmul(o[m], cconj(stw[Q * m + aj])));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, v), rs, voff, m * zstep, 0);
    }
  }
}

template <int HH>
static int green_reg_launch(const GreenConst& gc, const ColumnConsts& C, double alpha, MPtr<float2> W,
                            MPtr<float2> NQ, MPtr<const float2> TM, MPtr<const float2> TN, int nm,
                            const float2* tw3, int l2, hipStream_t st) {
  constexpr int TX = 16;
  const int m1 = gc.n1 / 2;
  if (HH < 32 || HH > 128 || m1 % TX || gc.n2 % TX) return RELAX_ERR_SIZE;
  const int nb = (m1 / TX) * gc.n2 + gc.n2 / TX;
  k_green_reg<(HH < 32 ? 32 : HH > 128 ? 128 : HH)><<<dim3(nb, nm), 512, 0, st>>>(
      C, gc.mu, alpha, gc.gamma, gc.n1, gc.n2, gc.n3, W, NQ, TM, TN, tw3, l2);
  return RELAX_OK;
}

// Launch shapes of the column pass: columns per workgroup, threads, largest
// radix.  Variant 0 is the default (measured at 512x512x256: 4 columns of
// 128 with 256 threads and radix 16 is 4 % faster than 8 columns with 512
// threads, 25 % faster than radix 8); the autotuner may pick another.
template <int H3, int V>
struct GreenVariant {
  static constexpr int TX0 = H3 >= 128 ? 4 : 16;
  static constexpr int NT0 = H3 >= 256 ? 512 : 256;
  static constexpr int TX = V == 1 ? (H3 >= 256 ? 4 : TX0 * 2 > 16 ? 16 : TX0 * 2) : TX0;
  static constexpr int NT = V == 1 ? (H3 >= 256 ? 256 : 512) : NT0;
  static constexpr int RM = V == 2 ? 8 : 16;
};

template <int HH, int V>
static int green_launch(const GreenConst& gc, const ColumnConsts& C, double alpha, MPtr<float2> W, MPtr<float2> NQ,
                        MPtr<const float2> TM, MPtr<const float2> TN, int nm, const float2* tw3, int l2,
                        hipStream_t st) {
  constexpr int TX = GreenVariant<HH, V>::TX, NT = GreenVariant<HH, V>::NT, RM = GreenVariant<HH, V>::RM;
  const int m1 = gc.n1 / 2;
  if (m1 % TX || gc.n2 % TX) return RELAX_ERR_SIZE;
  const int nb = (m1 / TX) * gc.n2 + gc.n2 / TX;
  k_green_columns<HH, TX, NT, RM><<<dim3(nb, nm), NT, 0, st>>>(C, gc.mu, alpha, gc.gamma, gc.n1, gc.n2, gc.n3, W,
                                                               NQ, TM, TN, tw3, l2);
  return RELAX_OK;
}

int launch_green_columns(const GreenConst& gc, int h3, MPtr<float2> W, MPtr<float2> NQ, MPtr<const float2> TM,
                         MPtr<const float2> TN, int nm, hipStream_t st, int l2, int variant) {
  const float2* tw3 = twiddle_table(gc.n3);
  if (!tw3) return RELAX_ERR_NOMEM;
  if (variant < 0 || variant >= GREEN_VARIANTS) return RELAX_ERR_ARG;
  const ColumnConsts C = make_consts(gc);
  const double alpha = (gc.lambda + gc.mu) / (gc.lambda + 2.0 * gc.mu);
  int rc = RELAX_OK;
#define X(HH)                                                                                          \
  case HH:                                                                                             \
    if (variant == 0) rc = green_launch<HH, 0>(gc, C, alpha, W, NQ, TM, TN, nm, tw3, l2, st);          \
    else if (variant == 1) rc = green_launch<HH, 1>(gc, C, alpha, W, NQ, TM, TN, nm, tw3, l2, st);     \
    else if (variant == 2) rc = green_launch<HH, 2>(gc, C, alpha, W, NQ, TM, TN, nm, tw3, l2, st);     \
    else rc = green_reg_launch<HH>(gc, C, alpha, W, NQ, TM, TN, nm, tw3, l2, st);                      \
    break;
  switch (h3) {
    X(8) X(16) X(32) X(64) X(128) X(256)
    default: return RELAX_ERR_SIZE;
  }
#undef X
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// --------------------------------------- per-routine padded-spectrum kernels --
// Spectra are (sx1/2+1) complex per row, rows (k2) then planes (k3).
__global__ void k_elastic_padded(ColumnConsts C, int n1, int n2, int n3, float2* f1, float2* f2, float2* f3) {
  const int h1 = n1 / 2 + 1;
  const long n = (long)h1 * n2 * n3;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % h1, i2 = (o / h1) % n2, i3 = o / ((long)h1 * n2);
    float2 a1 = f1[o], a2 = f2[o], a3 = f3[o];
    if (i1 == 0 && i2 == 0 && i3 == 0) {
      a1 = a2 = a3 = make_float2(0.f, 0.f);
    } else {
      elastic_point(C, (float)i1 * C.ik1, wavenumber23f(i2, n2, C.ik2), wavenumber23f(i3, n3, C.ik3), a1, a2, a3);
    }
    f1[o] = a1; f2[o] = a2; f3[o] = a3;
  }
}

// one thread per (k1,k2) column, k3 summed in the reference's order
__global__ void k_surface_padded(ColumnConsts C, int n1, int n2, int n3, const float2* u1, const float2* u2,
                                 const float2* u3, const float2* t1, const float2* t2, const float2* t3,
                                 float2* p1, float2* p2, float2* p3) {
  const int h1 = n1 / 2 + 1;
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)h1 * n2) return;
  const int i1 = o % h1, i2 = o / h1;
  const float k1 = (float)i1 * C.ik1, k2 = wavenumber23f(i2, n2, C.ik2);
  float2 a1 = make_float2(0.f, 0.f), a2 = a1, a3 = a1;
  for (int i3 = 0; i3 < n3; ++i3) {
    const long q = o + (long)i3 * h1 * n2;
    float2 s1, s2, s3;
    traction_point(C, k1, k2, wavenumber23f(i3, n3, C.ik3), u1[q], u2[q], u3[q], s1, s2, s3);
    a1 = cadd(a1, s1); a2 = cadd(a2, s2); a3 = cadd(a3, s3);
  }
  a1 = make_float2(a1.x / C.inv_n3dx3, a1.y / C.inv_n3dx3);
  a2 = make_float2(a2.x / C.inv_n3dx3, a2.y / C.inv_n3dx3);
  a3 = make_float2(a3.x / C.inv_n3dx3, a3.y / C.inv_n3dx3);
  if (t1) {  // greenfunctioncowling: p = t - p
    a1 = csub(t1[o], a1); a2 = csub(t2[o], a2); a3 = csub(t3[o], a3);
  }
  p1[o] = a1; p2[o] = a2; p3[o] = a3;
}

__global__ void k_cerruti_padded(ColumnConsts C, double mu, double alpha, double gamma, int n1, int n2, int n3,
                                 const float2* p1, const float2* p2, const float2* p3, float2* u1, float2* u2,
                                 float2* u3) {
  const int h1 = n1 / 2 + 1;
  const long n = (long)h1 * n2 * n3;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % h1, i2 = (o / h1) % n2, i3 = o / ((long)h1 * n2);
    const long q = o % ((long)h1 * n2);
    const float k1 = (float)i1 * C.ik1, k2 = wavenumber23f(i2, n2, C.ik2);
    const CerrutiCol cc = cerruti_column(C, mu, alpha, gamma, k1, k2, p1[q], p2[q], p3[q]);
    float2 a1 = u1[o], a2 = u2[o], a3 = u3[o];
    cerruti_point(C, cc, p1[q], p2[q], wavenumber23f(i3, n3, C.ik3), a1, a2, a3);
    u1[o] = a1; u2[o] = a2; u3[o] = a3;
  }
}

static dim3 grid_for(long n) {
  long b = (n + 255) / 256;
  return dim3((unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b)));
}

int launch_elastic_padded(const GreenConst& gc, float2* f1, float2* f2, float2* f3, hipStream_t st) {
  const ColumnConsts C = make_consts(gc);
  k_elastic_padded<<<grid_for((long)(gc.n1 / 2 + 1) * gc.n2 * gc.n3), 256, 0, st>>>(C, gc.n1, gc.n2, gc.n3, f1, f2,
                                                                                      f3);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_surface_padded(const GreenConst& gc, const float2* u1, const float2* u2, const float2* u3,
                          const float2* t1, const float2* t2, const float2* t3, float2* p1, float2* p2,
                          float2* p3, hipStream_t st) {
  const ColumnConsts C = make_consts(gc);
  const long n = (long)(gc.n1 / 2 + 1) * gc.n2;
  k_surface_padded<<<(unsigned)((n + 127) / 128), 128, 0, st>>>(C, gc.n1, gc.n2, gc.n3, u1, u2, u3, t1, t2, t3, p1,
                                                                 p2, p3);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_cerruti_padded(const GreenConst& gc, const float2* p1, const float2* p2, const float2* p3,
                          float2* u1, float2* u2, float2* u3, hipStream_t st) {
  const ColumnConsts C = make_consts(gc);
  const double alpha = (gc.lambda + gc.mu) / (gc.lambda + 2.0 * gc.mu);
  k_cerruti_padded<<<grid_for((long)(gc.n1 / 2 + 1) * gc.n2 * gc.n3), 256, 0, st>>>(
      C, gc.mu, alpha, gc.gamma, gc.n1, gc.n2, gc.n3, p1, p2, p3, u1, u2, u3);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
