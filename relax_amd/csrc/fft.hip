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
// The tables exp(-2 pi i j / L), j < L, of every power of two L = 2 ..
// 4096, back to back (table L at offset L - 2), in one static device array.
// They are filled once per device, at the first transform on it, so no
// transform ever allocates device memory.  The values are evaluated in long
// double and rounded once, so each table is correctly rounded.
static const int TW_MAX = 4096;
static __device__ __align__(256) float2 twiddles[2 * TW_MAX - 2];
static std::mutex g_tw_mu;
static std::map<int, float2*> g_tw;  // device -> address of its twiddle array

const float2* twiddle_table(int L) {
  if (L < 2 || L > TW_MAX || (L & (L - 1))) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw.find(dev);
  if (it == g_tw.end()) {
    float2* base = nullptr;
    if (hipGetSymbolAddress((void**)&base, HIP_SYMBOL(twiddles)) != hipSuccess || !base) return nullptr;
    std::vector<float2> h(2 * TW_MAX - 2);
    for (int n = 2; n <= TW_MAX; n *= 2)
      for (int j = 0; j < n; ++j) {
        const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)n;
        h[n - 2 + j] = make_float2((float)cosl(a), (float)sinl(a));
      }
    if (hipMemcpy(base, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    it = g_tw.emplace(dev, base).first;
  }
  return it->second + (L - 2);
}

// ---------------------------------------------------------------- x R2C ----
// largest radix of the row transforms (compile-time choice)
#ifndef RELAX_X_RMAX
#define RELAX_X_RMAX 8
#endif

// With YS (the split y transform of the model path), the rows of a plane of
// n2 rows are taken in groups of 8, rows j1 + L2*j2 (L2 = n2/8), and after
// the x transform each column of a group gets the first stage of the y
// transform: an 8-point DFT over j2 and the twiddle w_n2^(j1*k2), written
// to row j1 + L2*k2.  The L2-point transforms of the blocks of L2 rows then
// finish the y transform (k_fft_cols), leaving y frequency 8*k1 + k2 at row
// L2*k2 + k1 of the plane.  The Nyquist column keeps natural row order.
// complex samples of the rows a workgroup transforms together
#ifndef RELAX_X_ELEMS
#define RELAX_X_ELEMS 2048
#endif
// threads of the x transforms
#ifndef RELAX_X_NT
#define RELAX_X_NT 256
#endif
// LDS padding of the x transforms: one spare element after every
// 2^RELAX_X_PAD (0: none; lpad of fft_lds.h)
#ifndef RELAX_X_PAD
#define RELAX_X_PAD 3
#endif

template <int N1, bool YS>
struct XRows {
  static constexpr int M = N1 / 2;
  static constexpr int RPB0 = (RELAX_X_ELEMS / M) > 0 ? RELAX_X_ELEMS / M : 1;
  static constexpr int RPB = YS ? (RPB0 < 8 ? 8 : RPB0) : RPB0;  // rows per workgroup
  static constexpr int G = YS ? RPB / 8 : 1;                       // groups of 8 rows
};

// Rows of one workgroup: bid is the workgroup's index among those of its
// array; s, stw and ytw are the kernel's LDS (sized for YS).
template <int N1, int NT, bool YS>
__device__ __forceinline__ void x_r2c_block(float2* s, const float2* stw, float2 (*ytw)[8], long bid,
                                            const float* in, long in_pitch, float2* out, long out_pitch,
                                            float2* __restrict__ nyq, long nyq_stride, long nrows, float scale,
                                            int n2, const float2* __restrict__ twy) {
  constexpr int M = N1 / 2;
  constexpr int RPB = XRows<N1, YS>::RPB, G = XRows<N1, YS>::G;
  constexpr int PM = M + 1;  // row pitch in LDS (breaks bank aliasing between rows)
  constexpr int EPT = (RPB * M + NT - 1) / NT;
  constexpr int XP = RELAX_X_PAD;
  const int tid = threadIdx.x;
  // n2 is a power of two (64 or more with YS): L2 = n2/8 as shift and mask
  const int L2 = n2 / 8, ls2 = YS ? __builtin_ctz((unsigned)n2) - 3 : 0;
  // global row of LDS row b
  auto rowof = [&](int b) -> long {
    if (!YS) return bid * RPB + b;
    const long q = bid * G + b / 8;
    return ((q >> ls2) << (ls2 + 3)) + (q & (L2 - 1)) + (long)L2 * (b % 8);
  };
  if (YS && tid < G * 8) {
    const int j1 = (int)((bid * G + tid / 8) & (L2 - 1));
    ytw[tid / 8][tid % 8] = twy[(j1 * (tid % 8)) & (n2 - 1)];
  }
  for (int e = tid; e < RPB * M; e += NT) {
    const int b = e / M, j = e % M;
    const long row = rowof(b);
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) v = gload(reinterpret_cast<const float2*>(in + row * in_pitch + 2 * j));
    s[lpad<XP>(b * PM + j)] = v;
  }
  __syncthreads();
  lds_fft<M, RPB, NT, RPB, PM, 0, 1, -1, 2, false, RELAX_X_RMAX, XP>(s, stw, tid);
  float2 X[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    if (e >= RPB * M) break;
    const int b = e / M, k = e % M;
    const float2 zk = s[lpad<XP>(b * PM + k)];
    const float2 zr = cconj(s[lpad<XP>(b * PM + ((M - k) & (M - 1)))]);
    const float2 w = stw[k];
    const float2 sum = cadd(zk, zr), dif = csub(zk, zr);
    // X[k] = (Z[k] + conj Z[M-k])/2 - i w^k (Z[k] - conj Z[M-k])/2
    const float2 wd = mul_mi(cmul(w, dif));
    X[i] = make_float2(0.5f * (sum.x + wd.x) * scale, 0.5f * (sum.y + wd.y) * scale);
    const long row = rowof(b);
    if (k == 0 && row < nrows) gstore(nyq + row * nyq_stride, make_float2((zk.x - zk.y) * scale, 0.f));
    if (!YS && row < nrows) gstore(out + row * out_pitch + k, X[i]);
  }
  if (!YS) return;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    if (e >= RPB * M) break;
    s[lpad<XP>((e / M) * PM + (e % M))] = X[i];
  }
  __syncthreads();
  // first stage of the y transform, one column of one group per item
  for (int e = tid; e < G * M; e += NT) {
    const int gi = e / M, k = e % M;
    const long q = bid * G + gi;
    if (q * 8 >= nrows) continue;
    float2 a[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) a[t] = s[lpad<XP>((8 * gi + t) * PM + k)];
    dft8<-1>(a);
    const long base_row = ((q >> ls2) << (ls2 + 3)) + (q & (L2 - 1));
#pragma unroll
    for (int t = 0; t < 8; ++t) gstore(out + (base_row + (long)L2 * t) * out_pitch + k, t ? cmul(a[t], ytw[gi][t]) : a[t]);
  }
}

// The first side.nb workgroups transform the rows of the side array (plain
// rows, no split y stage); the others the main array's.
template <int N1, int NT, bool YS>
__global__ __launch_bounds__(NT) void k_fft_x_r2c(MPtr<const float> in_m, long in_pitch, MPtr<float2> out_m,
                                                  long out_pitch, MPtr<float2> nyq_m, long nyq_stride, long nrows,
                                                  float scale, const float2* __restrict__ tw, int n2,
                                                  const float2* __restrict__ twy, XSide side) {
  // in and out may be the same buffer (rows are staged in LDS before any write)
  constexpr int M = N1 / 2;
  constexpr int RPB = XRows<N1, YS>::RPB, G = XRows<N1, YS>::G;
  static_assert(XRows<N1, false>::RPB <= RPB, "the side rows fit the buffer");
  __shared__ float2 s[lpad<RELAX_X_PAD>(RPB * (M + 1)) + 1];
  __shared__ float2 stw[N1];  // twiddles in LDS: off the global-latency path of every stage
  __shared__ float2 ytw[G][8];
  for (int e = threadIdx.x; e < N1; e += NT) stw[e] = tw[e];
  const int y = blockIdx.y;
  if ((int)blockIdx.x < side.nb) {
    x_r2c_block<N1, NT, false>(s, stw, nullptr, blockIdx.x, side.in.at(y), side.in_pitch, side.out.at(y),
                               side.out_pitch, side.nyq.at(y), side.nyq_stride, side.nrows, side.scale, 0, nullptr);
    return;
  }
  x_r2c_block<N1, NT, YS>(s, stw, ytw, (long)blockIdx.x - side.nb, in_m.at(y), in_pitch, out_m.at(y), out_pitch,
                          nyq_m.at(y), nyq_stride, nrows, scale, n2, twy);
}

// ---------------------------------------------------------------- x C2R ----
// With YS the workgroup first undoes the 8-point y stage of its row groups
// (the blocks' L2-point inverse transforms ran before, see k_fft_x_r2c).
// Waves per SIMD the 512-point inverse is held to: 7 (72 VGPRs, no spills;
// its LDS admits 7 workgroups per CU) runs it in 0.230 ms at 512x512x256
// against 0.247 ms for the compiler's own choice of 73 VGPRs and 6 waves.
// The shorter rows would spill under the same bound, and the 1024-point
// rows are held to 3 workgroups per CU by their LDS.
#ifndef RELAX_C2R_MINW
#define RELAX_C2R_MINW 7
#endif
template <int N1, int NT, int EPI, bool YS>
__global__ __launch_bounds__(NT, N1 == 512 ? RELAX_C2R_MINW : 1) void k_fft_x_c2r(MPtr<const float2> in_m, long in_pitch, MPtr<const float2> nyq_m,
                                                  long nyq_stride, MPtr<float> out_m, long out_pitch, long nrows,
                                                  float scale, const float2* __restrict__ tw, MPtr<const float> base_m,
                                                  long base_pitch, MPtr<const StepScalars> sc_m, int n2,
                                                  const float2* __restrict__ twy) {
  // in and out may be the same buffer (rows are staged in LDS before any write)
  const float2* in = in_m.at(blockIdx.y);
  const float2* __restrict__ nyq = nyq_m.at(blockIdx.y);
  float* out = out_m.at(blockIdx.y);
  const float* base = base_m.at(blockIdx.y);
  const StepScalars* __restrict__ sc = sc_m.at(blockIdx.y);
  constexpr int M = N1 / 2;
  constexpr int RPB = XRows<N1, YS>::RPB, G = XRows<N1, YS>::G;
  constexpr int PM = M + 1;
  constexpr int EPT = (RPB * M + NT - 1) / NT;
  constexpr int XP = RELAX_X_PAD;
  __shared__ float2 s[lpad<XP>(RPB * PM) + 1];
  __shared__ float2 xm[RPB];
  __shared__ float2 stw[N1];
  __shared__ float2 ytw[G][8];
  const int tid = threadIdx.x;
  // n2 is a power of two (64 or more with YS): L2 = n2/8 as shift and mask
  const int L2 = n2 / 8, ls2 = YS ? __builtin_ctz((unsigned)n2) - 3 : 0;
  auto rowof = [&](int b) -> long {
    if (!YS) return (long)blockIdx.x * RPB + b;
    const long q = (long)blockIdx.x * G + b / 8;
    return ((q >> ls2) << (ls2 + 3)) + (q & (L2 - 1)) + (long)L2 * (b % 8);
  };
  for (int e = tid; e < N1; e += NT) stw[e] = tw[e];
  if (YS && tid < G * 8) {
    const int j1 = (int)(((long)blockIdx.x * G + tid / 8) & (L2 - 1));
    ytw[tid / 8][tid % 8] = twy[(j1 * (tid % 8)) & (n2 - 1)];
  }
  for (int e = tid; e < RPB * M; e += NT) {
    const int b = e / M, j = e % M;
    const long row = rowof(b);
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) v = gload(in + row * in_pitch + j);
    if (!YS && j == 0) v.y = 0.f;  // a real output takes the real part of the k1=0 bin
    s[lpad<XP>(b * PM + j)] = v;
  }
  for (int b = tid; b < RPB; b += NT) {
    const long row = rowof(b);
    xm[b] = row < nrows ? make_float2(gload(nyq + row * nyq_stride).x, 0.f) : make_float2(0.f, 0.f);
  }
  // the epilogue's base values (the same elements this thread writes), loaded
  // now so that their latency overlaps the transform
  float2 ub[EPI != EPI_PLAIN ? EPT : 1];
  if (EPI != EPI_PLAIN) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT;
      ub[i] = make_float2(0.f, 0.f);
      if (e < RPB * M) {
        const long row = rowof(e / M);
        if (row < nrows) ub[i] = gload(reinterpret_cast<const float2*>(base + row * base_pitch + 2 * (e % M)));
      }
    }
  }
  __syncthreads();
  if (YS) {
    // y frequencies k2 of a group sit in LDS rows 8*gi + k2; the inverse
    // 8-point stage leaves rows j1 + L2*j2 in LDS rows 8*gi + j2
    for (int e = tid; e < G * M; e += NT) {
      const int gi = e / M, j = e % M;
      float2 a[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float2 v = s[lpad<XP>((8 * gi + t) * PM + j)];
        a[t] = t ? cmul(v, cconj(ytw[gi][t])) : v;
      }
      dft8<+1>(a);
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        float2 v = a[t];
        if (j == 0) v.y = 0.f;  // a real output takes the real part of the k1=0 bin
        s[lpad<XP>((8 * gi + t) * PM + j)] = v;
      }
    }
    __syncthreads();
  }
  float2 z[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    if (e >= RPB * M) break;
    const int b = e / M, k = e % M;
    const float2 xk = s[lpad<XP>(b * PM + k)];
    const float2 xr = cconj(k == 0 ? xm[b] : s[lpad<XP>(b * PM + (M - k))]);
    const float2 w = cconj(stw[k]);  // exp(+2 pi i k / N1)
    // Z'[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k])
    z[i] = cadd(cadd(xk, xr), mul_i(cmul(w, csub(xk, xr))));
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    if (e >= RPB * M) break;
    s[lpad<XP>((e / M) * PM + (e % M))] = z[i];
  }
  __syncthreads();
  lds_fft<M, RPB, NT, RPB, PM, 0, 1, +1, 2, false, RELAX_X_RMAX, XP>(s, stw, tid);
  float coef = 1.f;
  if (EPI == EPI_PRED) coef = (float)(sc->Dt / 2);
  if (EPI == EPI_CORR) coef = (float)sc->Dt;
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    if (e >= RPB * M) break;
    const int b = e / M, j = e % M;
    const long row = rowof(b);
    if (row >= nrows) continue;
    const float2 v = s[lpad<XP>(b * PM + j)];
    float a0 = v.x * scale, a1 = v.y * scale;
    if (EPI != EPI_PLAIN) {
      a0 = coef * a0 + ub[i].x;
      a1 = coef * a1 + ub[i].y;
    }
    gstore(reinterpret_cast<float2*>(out + row * out_pitch + 2 * j), make_float2(a0, a1));
  }
}

// --------------------------------------------------------- strided axis ----
// Columns per workgroup of the strided transforms: variant 0 is the
// default, 1 half as many, 2 twice as many (within 64 columns and 64 KB of
// LDS); the autotuner picks one per grid.
template <int L, int V>
struct ColTile {
  static constexpr int T0 = L >= 512 ? 8 : (L >= 128 ? 16 : (2048 / L > 64 ? 64 : 2048 / L));
  static constexpr int T1 = T0 / 2 < 4 ? 4 : T0 / 2;
  static constexpr int T2 = (T0 * 2 > 64 || L * T0 * 2 * 8 > 65536) ? T0 : T0 * 2;
  static constexpr int TX = V == 0 ? T0 : (V == 1 ? T1 : T2);
};

// TX columns of one workgroup: tile bid (plane bid / chunks, column chunk
// bid % chunks); s holds L * TX samples.
template <int L, int NT, int SIGN, int TX>
__device__ __forceinline__ void cols_block(float2* s, const float2* stw, float2* data, long jstride, long pstride,
                                           int ncols, int bid) {
  const int tid = threadIdx.x;
  const int nchunk = (ncols + TX - 1) / TX;
  const int col0 = (bid % nchunk) * TX;
  float2* base = data + (long)(bid / nchunk) * pstride;
  for (int e = tid; e < L * TX; e += NT) {
    const int b = e % TX, j = e / TX;
    const int col = col0 + b;
    s[e] = col < ncols ? gload(base + j * jstride + col) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<L, TX, NT, TX, 1, 0, TX, SIGN, 1, true>(s, stw, tid);
  for (int e = tid; e < L * TX; e += NT) {
    const int b = e % TX, j = e / TX;
    const int col = col0 + b;
    if (col < ncols) gstore(base + j * jstride + col, s[e]);
  }
}

template <int L, int NT, int SIGN, int TX>
__global__ __launch_bounds__(NT) void k_fft_cols(MPtr<float2> data_m, long jstride, long pstride, int ncols,
                                                 const float2* __restrict__ tw) {
  __shared__ float2 s[L * TX];
  __shared__ float2 stw[L];
  for (int e = threadIdx.x; e < L; e += NT) stw[e] = tw[e];
  cols_block<L, NT, SIGN, TX>(s, stw, data_m.at(blockIdx.y), jstride, pstride, ncols,
                              xcd_tile(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------- contiguous ----
// RPB rows of one workgroup, row pitch PL in LDS.
template <int L, int NT, int SIGN, int RPB, int PL>
__device__ __forceinline__ void rows_block(float2* s, const float2* stw, float2* data, long pitch, long nrows,
                                           long bid) {
  const int tid = threadIdx.x;
  const long row0 = bid * RPB;
  for (int e = tid; e < RPB * L; e += NT) {
    const int b = e / L, j = e % L;
    const long row = row0 + b;
    s[b * PL + j] = row < nrows ? gload(data + row * pitch + j) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<L, RPB, NT, RPB, PL, 0, 1, SIGN, 1, false>(s, stw, tid);
  for (int e = tid; e < RPB * L; e += NT) {
    const int b = e / L, j = e % L;
    const long row = row0 + b;
    if (row < nrows) gstore(data + row * pitch + j, s[b * PL + j]);
  }
}

template <int L, int NT, int SIGN>
__global__ __launch_bounds__(NT) void k_fft_rows(MPtr<float2> data_m, long pitch, long nrows,
                                                 const float2* __restrict__ tw) {
  constexpr int RPB = (2048 / L) > 0 ? 2048 / L : 1;
  constexpr int PL = L + 1;
  __shared__ float2 s[RPB * PL];
  __shared__ float2 stw[L];
  for (int e = threadIdx.x; e < L; e += NT) stw[e] = tw[e];
  rows_block<L, NT, SIGN, RPB, PL>(s, stw, data_m.at(blockIdx.y), pitch, nrows, blockIdx.x);
}

// ---------------------------------------------------------------- y pass ----
// The y transform of a Green solve's work spectrum (L-point columns, TX per
// workgroup) with the small LS-point transforms of the same solve riding
// along (YRide): the first nside workgroups take the traction planes'
// columns and the Nyquist rows, L*TX/LS columns or rows each, with their
// twiddles read from the global table.  One launch in place of four: the
// small transforms are a few microseconds of work each, but as launches of
// their own each costs a launch.
template <int L, int LS, int NT, int SIGN, int TX>
__global__ __launch_bounds__(NT) void k_fft_ypass(MPtr<float2> data_m, long jstride, long pstride, int ncols,
                                                  const float2* __restrict__ tw, YRideDev rd) {
  constexpr int TS = L * TX / LS;  // columns or rows of a side workgroup
  static_assert(TS >= 1, "a side workgroup holds one LS-point transform at least");
  __shared__ float2 s[L * TX];
  __shared__ float2 stw[L];
  const int y = blockIdx.y;
  int b = blockIdx.x;
  if (b < rd.nside) {
    if (b < rd.nbt) {
      cols_block<LS, NT, SIGN, TS>(s, rd.tw, rd.r.tc.at(y), rd.r.tjs, rd.r.tps, rd.r.tnc, b);
      return;
    }
    b -= rd.nbt;
    if (b < rd.nb0) {
      rows_block<LS, NT, SIGN, TS, LS>(s, rd.tw, rd.r.r0.at(y), rd.r.rp0, rd.r.rn0, b);
      return;
    }
    b -= rd.nb0;
    if (b < rd.nb1) rows_block<LS, NT, SIGN, TS, LS>(s, rd.tw, rd.r.r1.at(y), rd.r.rp1, rd.r.rn1, b);
    return;
  }
  for (int e = threadIdx.x; e < L; e += NT) stw[e] = tw[e];
  cols_block<L, NT, SIGN, TX>(s, stw, data_m.at(y), jstride, pstride, ncols,
                              xcd_tile(b - rd.nside, gridDim.x - rd.nside));
}

// ----------------------------------------------------------- launchers ----
#define NT_FFT 256

// ys_n2 > 0: the split y transform over planes of ys_n2 rows (see
// k_fft_x_r2c); nrows is then a multiple of ys_n2.
int launch_fft_x_r2c(int n1, MPtr<const float> in, long in_pitch, MPtr<float2> out, long out_pitch,
                     MPtr<float2> nyq, long nyq_stride, long nrows, float scale, int nm, hipStream_t st, int ys_n2,
                     const XSide* side) {
  const float2* tw = twiddle_table(n1);
  if (!tw) return RELAX_ERR_NOMEM;
  const float2* twy = nullptr;
  if (ys_n2) {
    if (ys_n2 < 64 || (ys_n2 & (ys_n2 - 1)) || nrows % ys_n2) return RELAX_ERR_SIZE;
    if (!(twy = twiddle_table(ys_n2))) return RELAX_ERR_NOMEM;
  }
  XSide sd;
  if (side && side->nrows > 0) sd = *side;
  long nb = 0;
  switch (n1) {
#define X(N)                                                                                                \
  case N:                                                                                                   \
    sd.nb = (int)((sd.nrows + XRows<N, false>::RPB - 1) / XRows<N, false>::RPB);                            \
    if (ys_n2) {                                                                                            \
      constexpr int G = XRows<N, true>::G;                                                                  \
      nb = (nrows / 8 + G - 1) / G;                                                                         \
      k_fft_x_r2c<N, RELAX_X_NT, true><<<dim3((unsigned)(nb + sd.nb), (unsigned)nm), RELAX_X_NT, 0, st>>>(          \
          in, in_pitch, out, out_pitch, nyq, nyq_stride, nrows, scale, tw, ys_n2, twy, sd);                 \
    } else {                                                                                                \
      constexpr int RPB = XRows<N, false>::RPB;                                                             \
      nb = (nrows + RPB - 1) / RPB;                                                                         \
      k_fft_x_r2c<N, RELAX_X_NT, false><<<dim3((unsigned)(nb + sd.nb), (unsigned)nm), RELAX_X_NT, 0, st>>>(         \
          in, in_pitch, out, out_pitch, nyq, nyq_stride, nrows, scale, tw, 0, nullptr, sd);                 \
    }                                                                                                       \
    break;
    X(16) X(32) X(64) X(128) X(256) X(512) X(1024)
#undef X
    default: return RELAX_ERR_SIZE;
  }
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_fft_x_c2r(int n1, int epi, MPtr<const float2> in, long in_pitch, MPtr<const float2> nyq,
                     long nyq_stride, MPtr<float> out, long out_pitch, long nrows, float scale,
                     MPtr<const float> base, long base_pitch, MPtr<const StepScalars> sc, int nm, hipStream_t st,
                     int ys_n2) {
  const float2* tw = twiddle_table(n1);
  if (!tw) return RELAX_ERR_NOMEM;
  const float2* twy = nullptr;
  if (ys_n2) {
    if (ys_n2 < 64 || (ys_n2 & (ys_n2 - 1)) || nrows % ys_n2) return RELAX_ERR_SIZE;
    if (!(twy = twiddle_table(ys_n2))) return RELAX_ERR_NOMEM;
  }
#define LAUNCH(N, E, YS)                                                                                    \
  {                                                                                                         \
    constexpr int RPB = XRows<N, YS>::RPB, G = XRows<N, YS>::G;                                             \
    const long nb = YS ? (nrows / 8 + G - 1) / G : (nrows + RPB - 1) / RPB;                                 \
    k_fft_x_c2r<N, RELAX_X_NT, E, YS><<<dim3((unsigned)nb, (unsigned)nm), RELAX_X_NT, 0, st>>>(                     \
        in, in_pitch, nyq, nyq_stride, out, out_pitch, nrows, scale, tw, base, base_pitch, sc, ys_n2, twy); \
  }
#define X(N, E) case N: if (ys_n2) LAUNCH(N, E, true) else LAUNCH(N, E, false) break;
#define SW(E) switch (n1) { X(16, E) X(32, E) X(64, E) X(128, E) X(256, E) X(512, E) X(1024, E) default: return RELAX_ERR_SIZE; }
  if (epi == EPI_PLAIN) { SW(EPI_PLAIN) }
  else if (epi == EPI_PRED) { SW(EPI_PRED) }
  else if (epi == EPI_ADD) { SW(EPI_ADD) }
  else { SW(EPI_CORR) }
#undef SW
#undef X
#undef LAUNCH
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

template <int L, int SIGN, int V>
static void cols_launch(MPtr<float2> data, long jstride, long pstride, int ncols, int nplanes, int nm,
                        const float2* tw, hipStream_t st) {
  constexpr int TX = ColTile<L, V>::TX;
  const dim3 grid((unsigned)(((ncols + TX - 1) / TX) * (long)nplanes), (unsigned)nm);
  k_fft_cols<L, NT_FFT, SIGN, TX><<<grid, NT_FFT, 0, st>>>(data, jstride, pstride, ncols, tw);
}

int launch_fft_cols(int L, int sign, MPtr<float2> data, long jstride, long pstride, int ncols, int nplanes, int nm,
                    hipStream_t st, int variant) {
  const float2* tw = twiddle_table(L);
  if (!tw) return RELAX_ERR_NOMEM;
  if (variant < 0 || variant >= COLS_VARIANTS) return RELAX_ERR_ARG;
#define V(N, S)                                                                                 \
  case N:                                                                                       \
    if (variant == 0) cols_launch<N, S, 0>(data, jstride, pstride, ncols, nplanes, nm, tw, st); \
    else if (variant == 1) cols_launch<N, S, 1>(data, jstride, pstride, ncols, nplanes, nm, tw, st); \
    else cols_launch<N, S, 2>(data, jstride, pstride, ncols, nplanes, nm, tw, st);               \
    break;
  if (sign < 0) {
    switch (L) { V(16, -1) V(32, -1) V(64, -1) V(128, -1) V(256, -1) V(512, -1) V(1024, -1) default: return RELAX_ERR_SIZE; }
  } else {
    switch (L) { V(16, 1) V(32, 1) V(64, 1) V(128, 1) V(256, 1) V(512, 1) V(1024, 1) default: return RELAX_ERR_SIZE; }
  }
#undef V
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_fft_rows(int L, int sign, MPtr<float2> data, long pitch, long nrows, int nm, hipStream_t st) {
  const float2* tw = twiddle_table(L);
  if (!tw) return RELAX_ERR_NOMEM;
  const int rpb = 2048 / L > 0 ? 2048 / L : 1;
  const dim3 grid((unsigned)((nrows + rpb - 1) / rpb), (unsigned)nm);
#define X(N, S) case N: k_fft_rows<N, NT_FFT, S><<<grid, NT_FFT, 0, st>>>(data, pitch, nrows, tw); break;
  if (sign < 0) {
    switch (L) { X(16, -1) X(32, -1) X(64, -1) X(128, -1) X(256, -1) X(512, -1) X(1024, -1) default: return RELAX_ERR_SIZE; }
  } else {
    switch (L) { X(16, 1) X(32, 1) X(64, 1) X(128, 1) X(256, 1) X(512, 1) X(1024, 1) default: return RELAX_ERR_SIZE; }
  }
#undef X
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

template <int L, int LS, int SIGN, int V>
static void ypass_launch(MPtr<float2> data, long jstride, long pstride, int ncols, int nplanes, YRideDev rd, int nm,
                         const float2* tw, hipStream_t st) {
  constexpr int TX = ColTile<L, V>::TX, TS = L * TX / LS;
  rd.nbt = rd.r.tnp > 0 ? ((rd.r.tnc + TS - 1) / TS) * rd.r.tnp : 0;
  rd.nb0 = (int)((rd.r.rn0 + TS - 1) / TS);
  rd.nb1 = (int)((rd.r.rn1 + TS - 1) / TS);
  // a multiple of 8 workgroups, so that the main tiles keep their XCD order
  rd.nside = (rd.nbt + rd.nb0 + rd.nb1 + 7) / 8 * 8;
  const long nmain = ((ncols + TX - 1) / TX) * (long)nplanes;
  const dim3 grid((unsigned)(rd.nside + nmain), (unsigned)nm);
  k_fft_ypass<L, LS, NT_FFT, SIGN, TX><<<grid, NT_FFT, 0, st>>>(data, jstride, pstride, ncols, tw, rd);
}

int launch_fft_ypass(int L, int LS, int sign, MPtr<float2> data, long jstride, long pstride, int ncols, int nplanes,
                     const YRide& ride, int nm, hipStream_t st, int variant) {
  // the work spectrum's columns are LS/8 points from LS = 128 up (the
  // split y transform), LS points below
  if (L != (LS >= 128 ? LS / 8 : LS)) return RELAX_ERR_SIZE;
  if (variant < 0 || variant >= COLS_VARIANTS) return RELAX_ERR_ARG;
  if (ride.tnc < 0 || ride.tnp < 0 || ride.rn0 < 0 || ride.rn1 < 0) return RELAX_ERR_ARG;
  const float2* tw = twiddle_table(L);
  YRideDev rd{};
  rd.r = ride;
  rd.tw = twiddle_table(LS);
  if (!tw || !rd.tw) return RELAX_ERR_NOMEM;
#define V(N, S)                                                                                                     \
  case N:                                                                                                           \
    if (variant == 0) ypass_launch<(N >= 128 ? N / 8 : N), N, S, 0>(data, jstride, pstride, ncols, nplanes, rd, nm, tw, st);  \
    else if (variant == 1) ypass_launch<(N >= 128 ? N / 8 : N), N, S, 1>(data, jstride, pstride, ncols, nplanes, rd, nm, tw, st); \
    else ypass_launch<(N >= 128 ? N / 8 : N), N, S, 2>(data, jstride, pstride, ncols, nplanes, rd, nm, tw, st);  \
    break;
  if (sign < 0) {
    switch (LS) { V(32, -1) V(64, -1) V(128, -1) V(256, -1) V(512, -1) V(1024, -1) default: return RELAX_ERR_SIZE; }
  } else {
    switch (LS) { V(32, 1) V(64, 1) V(128, 1) V(256, 1) V(512, 1) V(1024, 1) default: return RELAX_ERR_SIZE; }
  }
#undef V
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
