// fft.hip -- pass kernels of the 3-D real FFTs (fourier.f90:177-355 and the
// cuFFT calls of cugreen.cu:1219-1266 / :1696-1789 replaced by hand-written
// Stockham transforms).
//
//   x pass:  R2C / C2R of rows of n1 reals via a complex FFT of n1/2 points
//            and the split-radix pre/post twiddle; the Nyquist bin k1 = n1/2
//            goes to a separate array so the spectrum keeps the real row's
//            footprint (in place).
//   columns: complex FFT along a strided axis for a tile of TX contiguous
//            columns (x2 axis, and x3 in the unpruned per-routine path).
//   rows:    complex FFT of contiguous rows (the Nyquist planes along x2).
#include <map>
#include <mutex>
#include <vector>
#include <cmath>

#include "fft_lds.h"
#include "kernels.h"

// ------------------------------------------------------------- twiddles ----
static std::mutex g_tw_mu;
static std::map<std::pair<int, int>, float2*> g_tw;  // (device, L) -> table

const float2* twiddle_table(int L) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, L);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  std::vector<float2> h(L);
  for (int j = 0; j < L; ++j) {
    // exp(-2 pi i j / L), evaluated in double through the exact symmetric
    // octant reduction of sin/cos so that the table is correctly rounded
    long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)L;
    h[j] = make_float2((float)cosl(a), (float)sinl(a));
  }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * L) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * L, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  g_tw[key] = d;
  return d;
}

// ---------------------------------------------------------------- x R2C ----
template <int N1, int NT>
__global__ __launch_bounds__(NT) void k_fft_x_r2c(const float* __restrict__ in, long in_pitch, float2* out,
                                                  long out_pitch, float2* nyq, long nyq_stride, long nrows,
                                                  float scale, const float2* __restrict__ tw) {
  constexpr int M = N1 / 2;
  constexpr int RPB = (2048 / M) > 0 ? 2048 / M : 1;
  constexpr int PM = M + 1;  // row pitch in LDS (breaks bank aliasing between rows)
  __shared__ float2 s[RPB * PM];
  __shared__ float2 stw[N1];  // twiddles in LDS: off the global-latency path of every stage
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * RPB;
  for (int e = tid; e < N1; e += NT) stw[e] = tw[e];
  for (int e = tid; e < RPB * M; e += NT) {
    const int b = e / M, j = e % M;
    const long row = row0 + b;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) v = *reinterpret_cast<const float2*>(in + row * in_pitch + 2 * j);
    s[b * PM + j] = v;
  }
  __syncthreads();
  lds_fft<M, RPB, NT, RPB, PM, 0, 1, -1, 2, false>(s, stw, tid);
  for (int e = tid; e < RPB * M; e += NT) {
    const int b = e / M, k = e % M;
    const long row = row0 + b;
    if (row >= nrows) continue;
    const float2 zk = s[b * PM + k];
    const float2 zr = cconj(s[b * PM + ((M - k) & (M - 1))]);
    const float2 w = stw[k];
    const float2 sum = cadd(zk, zr), dif = csub(zk, zr);
    // X[k] = (Z[k] + conj Z[M-k])/2 - i w^k (Z[k] - conj Z[M-k])/2
    const float2 wd = mul_mi(cmul(w, dif));
    const float2 X = make_float2(0.5f * (sum.x + wd.x) * scale, 0.5f * (sum.y + wd.y) * scale);
    out[row * out_pitch + k] = X;
    if (k == 0) {
      const float2 z0 = s[b * PM];
      nyq[row * nyq_stride] = make_float2((z0.x - z0.y) * scale, 0.f);
    }
  }
}

// ---------------------------------------------------------------- x C2R ----
template <int N1, int NT, int EPI>
__global__ __launch_bounds__(NT) void k_fft_x_c2r(const float2* in, long in_pitch, const float2* nyq,
                                                  long nyq_stride, float* out, long out_pitch, long nrows,
                                                  float scale, const float2* __restrict__ tw, const float* base,
                                                  long base_pitch, const StepScalars* __restrict__ sc) {
  constexpr int M = N1 / 2;
  constexpr int RPB = (2048 / M) > 0 ? 2048 / M : 1;
  constexpr int PM = M + 1;
  constexpr int EPT = (RPB * M + NT - 1) / NT;
  __shared__ float2 s[RPB * PM];
  __shared__ float2 xm[RPB];
  __shared__ float2 stw[N1];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * RPB;
  for (int e = tid; e < N1; e += NT) stw[e] = tw[e];
  for (int e = tid; e < RPB * M; e += NT) {
    const int b = e / M, j = e % M;
    const long row = row0 + b;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) v = in[row * in_pitch + j];
    if (j == 0) v.y = 0.f;  // a real output takes the real part of the k1=0 bin
    s[b * PM + j] = v;
  }
  for (int b = tid; b < RPB; b += NT) {
    const long row = row0 + b;
    xm[b] = row < nrows ? make_float2(nyq[row * nyq_stride].x, 0.f) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2 z[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    if (e >= RPB * M) break;
    const int b = e / M, k = e % M;
    const float2 xk = s[b * PM + k];
    const float2 xr = cconj(k == 0 ? xm[b] : s[b * PM + (M - k)]);
    const float2 w = cconj(stw[k]);  // exp(+2 pi i k / N1)
    // Z'[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k])
    z[i] = cadd(cadd(xk, xr), mul_i(cmul(w, csub(xk, xr))));
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    if (e >= RPB * M) break;
    s[(e / M) * PM + (e % M)] = z[i];
  }
  __syncthreads();
  lds_fft<M, RPB, NT, RPB, PM, 0, 1, +1, 2, false>(s, stw, tid);
  float coef = 1.f;
  if (EPI == EPI_PRED) coef = (float)(sc->Dt / 2);
  if (EPI == EPI_CORR) coef = (float)sc->Dt;
  for (int e = tid; e < RPB * M; e += NT) {
    const int b = e / M, j = e % M;
    const long row = row0 + b;
    if (row >= nrows) continue;
    const float2 v = s[b * PM + j];
    float a0 = v.x * scale, a1 = v.y * scale;
    if (EPI != EPI_PLAIN) {
      const float2 u = *reinterpret_cast<const float2*>(base + row * base_pitch + 2 * j);
      a0 = coef * a0 + u.x;
      a1 = coef * a1 + u.y;
    }
    *reinterpret_cast<float2*>(out + row * out_pitch + 2 * j) = make_float2(a0, a1);
  }
}

// --------------------------------------------------------- strided axis ----
// Columns per workgroup of the strided transforms (compile-time tile choice;
// RELAX_COLS_TX512 overrides the 512-point tile).
#ifndef RELAX_COLS_TX512
#define RELAX_COLS_TX512 8
#endif
template <int L>
struct ColTile {
  static constexpr int TX = L >= 1024 ? 8 : (L == 512 ? RELAX_COLS_TX512 : (L >= 128 ? 16 : (2048 / L > 64 ? 64 : 2048 / L)));
};

template <int L, int NT, int SIGN>
__global__ __launch_bounds__(NT) void k_fft_cols(float2* data, long jstride, long pstride, int ncols,
                                                 const float2* __restrict__ tw) {
  constexpr int TX = ColTile<L>::TX;
  __shared__ float2 s[L * TX];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  for (int e = tid; e < L; e += NT) stw[e] = tw[e];
  const int nchunk = (ncols + TX - 1) / TX;
  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const int col0 = (bid % nchunk) * TX;
  float2* base = data + (long)(bid / nchunk) * pstride;
  for (int e = tid; e < L * TX; e += NT) {
    const int b = e % TX, j = e / TX;
    const int col = col0 + b;
    s[e] = col < ncols ? base[j * jstride + col] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<L, TX, NT, TX, 1, 0, TX, SIGN, 1, true>(s, stw, tid);
  for (int e = tid; e < L * TX; e += NT) {
    const int b = e % TX, j = e / TX;
    const int col = col0 + b;
    if (col < ncols) base[j * jstride + col] = s[e];
  }
}

// ---------------------------------------------------------- contiguous ----
template <int L, int NT, int SIGN>
__global__ __launch_bounds__(NT) void k_fft_rows(float2* data, long pitch, long nrows,
                                                 const float2* __restrict__ tw) {
  constexpr int RPB = (2048 / L) > 0 ? 2048 / L : 1;
  constexpr int PL = L + 1;
  __shared__ float2 s[RPB * PL];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  for (int e = tid; e < L; e += NT) stw[e] = tw[e];
  const long row0 = (long)blockIdx.x * RPB;
  for (int e = tid; e < RPB * L; e += NT) {
    const int b = e / L, j = e % L;
    const long row = row0 + b;
    s[b * PL + j] = row < nrows ? data[row * pitch + j] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<L, RPB, NT, RPB, PL, 0, 1, SIGN, 1, false>(s, stw, tid);
  for (int e = tid; e < RPB * L; e += NT) {
    const int b = e / L, j = e % L;
    const long row = row0 + b;
    if (row < nrows) data[row * pitch + j] = s[b * PL + j];
  }
}

// ----------------------------------------------------------- launchers ----
#define NT_FFT 256

int launch_fft_x_r2c(int n1, const float* in, long in_pitch, float2* out, long out_pitch, float2* nyq,
                     long nyq_stride, long nrows, float scale, hipStream_t st) {
  const float2* tw = twiddle_table(n1);
  if (!tw) return RELAX_ERR_NOMEM;
  const int M = n1 / 2;
  const int rpb = 2048 / M > 0 ? 2048 / M : 1;
  const dim3 grid((unsigned)((nrows + rpb - 1) / rpb));
  switch (n1) {
#define X(N) case N: k_fft_x_r2c<N, NT_FFT><<<grid, NT_FFT, 0, st>>>(in, in_pitch, out, out_pitch, nyq, nyq_stride, nrows, scale, tw); break;
    X(16) X(32) X(64) X(128) X(256) X(512) X(1024)
#undef X
    default: return RELAX_ERR_SIZE;
  }
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_fft_x_c2r(int n1, int epi, const float2* in, long in_pitch, const float2* nyq, long nyq_stride,
                     float* out, long out_pitch, long nrows, float scale, const float* base, long base_pitch,
                     const StepScalars* sc, hipStream_t st) {
  const float2* tw = twiddle_table(n1);
  if (!tw) return RELAX_ERR_NOMEM;
  const int M = n1 / 2;
  const int rpb = 2048 / M > 0 ? 2048 / M : 1;
  const dim3 grid((unsigned)((nrows + rpb - 1) / rpb));
#define X(N, E) case N: k_fft_x_c2r<N, NT_FFT, E><<<grid, NT_FFT, 0, st>>>(in, in_pitch, nyq, nyq_stride, out, out_pitch, nrows, scale, tw, base, base_pitch, sc); break;
#define SW(E) switch (n1) { X(16, E) X(32, E) X(64, E) X(128, E) X(256, E) X(512, E) X(1024, E) default: return RELAX_ERR_SIZE; }
  if (epi == EPI_PLAIN) { SW(EPI_PLAIN) }
  else if (epi == EPI_PRED) { SW(EPI_PRED) }
  else { SW(EPI_CORR) }
#undef SW
#undef X
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_fft_cols(int L, int sign, float2* data, long jstride, long pstride, int ncols, int nplanes,
                    hipStream_t st) {
  const float2* tw = twiddle_table(L);
  if (!tw) return RELAX_ERR_NOMEM;
  int tx = 0;
  switch (L) {
#define T(N) case N: tx = ColTile<N>::TX; break;
    T(16) T(32) T(64) T(128) T(256) T(512) T(1024)
#undef T
    default: return RELAX_ERR_SIZE;
  }
  const dim3 grid((unsigned)(((ncols + tx - 1) / tx) * (long)nplanes));
#define X(N, S) case N: k_fft_cols<N, NT_FFT, S><<<grid, NT_FFT, 0, st>>>(data, jstride, pstride, ncols, tw); break;
  if (sign < 0) {
    switch (L) { X(16, -1) X(32, -1) X(64, -1) X(128, -1) X(256, -1) X(512, -1) X(1024, -1) default: return RELAX_ERR_SIZE; }
  } else {
    switch (L) { X(16, 1) X(32, 1) X(64, 1) X(128, 1) X(256, 1) X(512, 1) X(1024, 1) default: return RELAX_ERR_SIZE; }
  }
#undef X
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_fft_rows(int L, int sign, float2* data, long pitch, long nrows, hipStream_t st) {
  const float2* tw = twiddle_table(L);
  if (!tw) return RELAX_ERR_NOMEM;
  const int rpb = 2048 / L > 0 ? 2048 / L : 1;
  const dim3 grid((unsigned)((nrows + rpb - 1) / rpb));
#define X(N, S) case N: k_fft_rows<N, NT_FFT, S><<<grid, NT_FFT, 0, st>>>(data, pitch, nrows, tw); break;
  if (sign < 0) {
    switch (L) { X(16, -1) X(32, -1) X(64, -1) X(128, -1) X(256, -1) X(512, -1) X(1024, -1) default: return RELAX_ERR_SIZE; }
  } else {
    switch (L) { X(16, 1) X(32, 1) X(64, 1) X(128, 1) X(256, 1) X(512, 1) X(1024, 1) default: return RELAX_ERR_SIZE; }
  }
#undef X
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
