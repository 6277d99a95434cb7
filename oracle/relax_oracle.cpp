// relax_oracle.cpp -- CPU restatement of RELAX's per-time-step elastic +
// relaxation solve.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP
// library is compared against (tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py).  Nothing in the product links or calls it.
//
// Every routine follows one reference routine; the file:line of the routine it
// restates is given above each function (paths relative to the reference's
// src/ directory).  The layout is the reference's ALIGN_DATA layout (FFTW/MKL
// builds, src/include.f90:45-56):
//   vector fields   REAL*4 (sx1+2, sx2, sx3)    i1 fastest, row pitch sx1+2
//   tensor fields   TYPE(TENSOR) (sx1, sx2, sx3) 6 floats per point, AoS
//   surface planes  REAL*4 (sx1+2, sx2)
// The template parameter R is the storage type: R=float reproduces the
// reference's FP32 storage with its FP64 intermediates at the same places;
// R=double is the all-FP64 build of the same pipeline.
//
// Parity status of each function (how it is pinned) is listed in the header
// of tests/test_oracle.py.  Functions whose only pin is internal consistency
// say so there.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

typedef std::complex<double> cd;

// Planted-mistake hook for the tests that prove each pin can fail: 0 = the
// faithful restatement; any other value switches on one deliberate error.
static int g_mut = 0;
#define MUT(k) (g_mut == (k))

static const double PI = 3.141592653589793115997963468544185161;
static const double PI2 = 6.28318530717958623199592693708837032318;
static const double DEG2RAD = 0.01745329251994329547437168059786927;

// ---------------------------------------------------------------------------
// ABI structures (types.f90:26-97, SEQUENCE layouts)
// ---------------------------------------------------------------------------
extern "C" {
struct orc_layer { double z, gammadot0, stressexponent, cohesion, friction, Gk; };
struct orc_tensor_layer { float z, dum; float t[6]; };
struct orc_patch {
  double x1, x2, x3, lx, lz, slip, ss, ds, v, vss, vds, taus;
  float sig[6];
};
struct orc_plane { double x, y, z, width, length, strike, dip, rake; int px2, px3; };
struct orc_source { double slip, x, y, z, width, length, strike, dip, rake, period, phase, beta, opening; };
}

template <class R> struct T6 { R s11, s12, s13, s22, s23, s33; };

// ---------------------------------------------------------------------------
// FFT (power-of-two sizes).  Numerically an exact DFT in FP64; stores R.
// fourier.f90:177-201 (FFTW: R2C forward, in place, padded), :330-355 (ctfft),
// sign -1 forward (fourier.f90:36), forward scaled by dx1*dx2*dx3, inverse by
// 1/(sx1*dx1*sx2*dx2*sx3*dx3); fft2 :372-396 / :510-535 likewise in 2-D.
// ---------------------------------------------------------------------------
static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Twiddles w[k] = exp(-2 pi i k / n), k < n/2, evaluated in long double.
static std::vector<cd> twiddle_table(int n) {
  std::vector<cd> w(std::max(1, n / 2));
  for (int k = 0; k < n / 2; ++k) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * k / n;
    w[k] = cd((double)cosl(a), (double)sinl(a));
  }
  return w;
}

// In-place iterative radix-2 DFT of a contiguous sequence: bit-reversed
// reorder, then butterflies with the table twiddles (conjugated for sign +1).
static void fft1d(cd* a, int n, const cd* w, int sign) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  // butterflies in real arithmetic (std::complex's operator* takes the
  // library's NaN-recovery path)
  double* d = reinterpret_cast<double*>(a);
  const double* wd = reinterpret_cast<const double*>(w);
  const double sg = sign < 0 ? 1.0 : -1.0;
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len / 2, step = n / len;
    for (int i = 0; i < n; i += len) {
      for (int k = 0; k < half; ++k) {
        const double wr = wd[2 * (k * step)], wi = sg * wd[2 * (k * step) + 1];
        double* p = d + 2 * (i + k);
        double* q = d + 2 * (i + k + half);
        const double vr = q[0] * wr - q[1] * wi, vi = q[0] * wi + q[1] * wr;
        const double ur = p[0], ui = p[1];
        p[0] = ur + vr; p[1] = ui + vi;
        q[0] = ur - vr; q[1] = ui - vi;
      }
    }
  }
}

// DFT along one strided axis of a complex array: ncol sequences of length n,
// element j of sequence c at base(c) + j*stride, where the sequences of
// neighbouring c are contiguous in memory (columns c0..c0+B-1 share rows).
// Blocks of B neighbouring sequences are gathered into a scratch buffer, so
// each row read is B consecutive values.
static void fft_axis(cd* data, long ncol, long ncol_inner, long plane_stride, int n, long stride, int sign) {
  const std::vector<cd> w = twiddle_table(n);
  const int B = 16;
  const long nblk_inner = (ncol_inner + B - 1) / B;
  const long nouter = ncol / ncol_inner;
#pragma omp parallel
  {
    std::vector<cd> buf((size_t)B * n);
#pragma omp for schedule(static)
    for (long blk = 0; blk < nouter * nblk_inner; ++blk) {
      const long o = blk / nblk_inner, c0 = (blk % nblk_inner) * B;
      const int nb = (int)std::min<long>(B, ncol_inner - c0);
      cd* base = data + o * plane_stride + c0;
      for (int j = 0; j < n; ++j)
        for (int b = 0; b < nb; ++b) buf[(size_t)b * n + j] = base[j * stride + b];
      for (int b = 0; b < nb; ++b) fft1d(&buf[(size_t)b * n], n, w.data(), sign);
      for (int j = 0; j < n; ++j)
        for (int b = 0; b < nb; ++b) base[j * stride + b] = buf[(size_t)b * n + j];
    }
  }
}

// forward R2C along x1 then C2C along x2 (and x3): data is R[(sx1+2)*sx2*sx3]
template <class R>
static int fftn_forward(R* data, int sx1, int sx2, int sx3, double scale) {
  if (!is_pow2(sx1) || !is_pow2(sx2) || !is_pow2(sx3)) return 11;
  const int h1 = sx1 / 2 + 1;
  const long pitch = sx1 + 2;
  std::vector<cd> spec((size_t)h1 * sx2 * sx3);
  const std::vector<cd> w1 = twiddle_table(sx1);
#pragma omp parallel
  {
    std::vector<cd> row(sx1);
#pragma omp for
    for (long r = 0; r < (long)sx2 * sx3; ++r) {
      for (int i = 0; i < sx1; ++i) row[i] = cd((double)data[r * pitch + i], 0.0);
      fft1d(row.data(), sx1, w1.data(), -1);
      for (int k = 0; k < h1; ++k) spec[r * h1 + k] = row[k];
    }
  }
  // x2: in each x3 plane, h1 columns of sx2 at stride h1
  if (sx2 > 1) fft_axis(spec.data(), (long)h1 * sx3, h1, (long)h1 * sx2, sx2, h1, -1);
  // x3: h1*sx2 columns of sx3 at stride h1*sx2
  if (sx3 > 1) fft_axis(spec.data(), (long)h1 * sx2, (long)h1 * sx2, 0, sx3, (long)h1 * sx2, -1);
#pragma omp parallel for
  for (long r = 0; r < (long)sx2 * sx3; ++r)
    for (int k = 0; k < h1; ++k) {
      cd z = spec[r * h1 + k] * scale;
      data[r * pitch + 2 * k] = (R)z.real();
      data[r * pitch + 2 * k + 1] = (R)z.imag();
    }
  return 0;
}

template <class R>
static int fftn_inverse(R* data, int sx1, int sx2, int sx3, double scale) {
  if (!is_pow2(sx1) || !is_pow2(sx2) || !is_pow2(sx3)) return 11;
  const int h1 = sx1 / 2 + 1;
  const long pitch = sx1 + 2;
  std::vector<cd> spec((size_t)h1 * sx2 * sx3);
  const std::vector<cd> w1 = twiddle_table(sx1);
#pragma omp parallel for
  for (long r = 0; r < (long)sx2 * sx3; ++r)
    for (int k = 0; k < h1; ++k)
      spec[r * h1 + k] = cd((double)data[r * pitch + 2 * k], (double)data[r * pitch + 2 * k + 1]);
  if (sx3 > 1) fft_axis(spec.data(), (long)h1 * sx2, (long)h1 * sx2, 0, sx3, (long)h1 * sx2, +1);
  if (sx2 > 1) fft_axis(spec.data(), (long)h1 * sx3, h1, (long)h1 * sx2, sx2, h1, +1);
#pragma omp parallel
  {
    std::vector<cd> row(sx1);
#pragma omp for
    for (long r = 0; r < (long)sx2 * sx3; ++r) {
      // Hermitian extension; the imaginary parts of the k1=0 and Nyquist
      // bins do not contribute to a real output
      for (int k = 0; k < h1; ++k) row[k] = spec[r * h1 + k];
      row[0] = cd(row[0].real(), 0.0);
      row[sx1 / 2] = cd(row[sx1 / 2].real(), 0.0);
      for (int k = 1; k < sx1 / 2; ++k) row[sx1 - k] = std::conj(row[k]);
      fft1d(row.data(), sx1, w1.data(), +1);
      for (int i = 0; i < sx1; ++i) data[r * pitch + i] = (R)(row[i].real() * scale);
      data[r * pitch + sx1] = 0;
      data[r * pitch + sx1 + 1] = 0;
    }
  }
  return 0;
}

// fourier.f90:177-201 / :330-355
template <class R>
static int fft3(R* data, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, int direction) {
  if (direction == -1) return fftn_forward(data, sx1, sx2, sx3, MUT(12) ? dx1 * dx2 : dx1 * dx2 * dx3);
  return fftn_inverse(data, sx1, sx2, sx3, 1.0 / (sx1 * dx1 * sx2 * dx2 * sx3 * dx3));
}
// fourier.f90:372-396 / :510-535
template <class R>
static int fft2(R* data, int sx1, int sx2, double dx1, double dx2, int direction) {
  if (direction == -1) return fftn_forward(data, sx1, sx2, 1, dx1 * dx2);
  return fftn_inverse(data, sx1, sx2, 1, 1.0 / (sx1 * dx1 * sx2 * dx2));
}

// ---------------------------------------------------------------------------
// wavenumbers, fourier.f90:64-81 (cycles per unit length, negative above
// Nyquist in x2 and x3, k1 >= 0).  Indices here are 0-based.
// ---------------------------------------------------------------------------
static inline double wavenumber1(int i1, int sx1, double dx1) { return (double)i1 / (sx1 * dx1); }
static inline double wavenumber23(int i, int sx, double dx) {
  // 1-based: i < sx/2+1 -> (i-1)/(sx dx), else -(sx-i+1)/(sx dx)
  int i1b = i + 1;
  if (i1b < sx / 2 + 1) return ((double)i1b - 1.0) / (sx * dx);
  return -((double)(sx - i1b) + 1.0) / (sx * dx);
}

// ---------------------------------------------------------------------------
// elasticresponse, green.f90:51-101
// ---------------------------------------------------------------------------
template <class R>
static void elasticresponse(double lambda, double mu, R* f1, R* f2, R* f3, int sx1, int sx2, int sx3,
                            double dx1, double dx2, double dx3) {
  const double ratio1 = (lambda + mu) / (lambda + 2.0 * mu) / mu / (PI2 * PI2);
  const double ratio2 = mu / (lambda + mu);
  const long pitch = sx1 + 2;
#pragma omp parallel for
  for (int i3 = 0; i3 < sx3; ++i3) {
    double k3 = wavenumber23(i3, sx3, dx3);
    for (int i2 = 0; i2 < sx2; ++i2) {
      double k2 = wavenumber23(i2, sx2, dx2);
      for (int i1 = 0; i1 < sx1 / 2 + 1; ++i1) {
        double k1 = wavenumber1(i1, sx1, dx1);
        double r2 = k1 * k1 + k2 * k2 + k3 * k3;
        double denom = ratio1 / (r2 * r2);
        long o = (i2 + (long)sx2 * i3) * pitch + 2 * i1;
        cd c1(f1[o], f1[o + 1]), c2(f2[o], f2[o + 1]), c3(f3[o], f3[o + 1]);
        cd b1 = ((k2 * k2 + k3 * k3 + (MUT(2) ? 0.0 : ratio2) * r2) * c1 - (MUT(1) ? k2 : k1) * (k2 * c2 + k3 * c3)) * denom;
        cd b2 = ((k1 * k1 + k3 * k3 + ratio2 * r2) * c2 - k2 * (k1 * c1 + k3 * c3)) * denom;
        cd b3 = ((k1 * k1 + k2 * k2 + ratio2 * r2) * c3 - k3 * (k1 * c1 + k2 * c2)) * denom;
        f1[o] = (R)b1.real(); f1[o + 1] = (R)b1.imag();
        f2[o] = (R)b2.real(); f2[o + 1] = (R)b2.imag();
        f3[o] = (R)b3.real(); f3[o + 1] = (R)b3.imag();
      }
    }
  }
  // zero wavenumber, no net body force (green.f90:96-99)
  f1[0] = f1[1] = 0; f2[0] = f2[1] = 0; f3[0] = f3[1] = 0;
}

// ---------------------------------------------------------------------------
// surfacetractioncowling, green.f90:327-377.  p accumulates in R in i3 order.
// ---------------------------------------------------------------------------
template <class R>
static void surfacetractioncowling(double lambda, double mu, double gamma, const R* u1, const R* u2, const R* u3,
                                   int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,
                                   R* p1, R* p2, R* p3) {
  const double modulus = lambda + 2.0 * mu;
  const double alpha = (lambda + mu) / (lambda + 2.0 * mu);
  const double grav = 2.0 * mu * alpha * gamma;
  const long pitch = sx1 + 2;
  const cd I(0.0, MUT(3) ? -PI2 : PI2);
  std::fill(p1, p1 + pitch * sx2, (R)0);
  std::fill(p2, p2 + pitch * sx2, (R)0);
  std::fill(p3, p3 + pitch * sx2, (R)0);
#pragma omp parallel for
  for (int i2 = 0; i2 < sx2; ++i2) {
    double k2 = wavenumber23(i2, sx2, dx2);
    for (int i3 = 0; i3 < sx3; ++i3) {
      double k3 = wavenumber23(i3, sx3, dx3);
      for (int i1 = 0; i1 < sx1 / 2 + 1; ++i1) {
        double k1 = wavenumber1(i1, sx1, dx1);
        long o = (i2 + (long)sx2 * i3) * pitch + 2 * i1;
        cd c1(u1[o], u1[o + 1]), c2(u2[o], u2[o + 1]), c3(u3[o], u3[o + 1]);
        cd s1 = I * mu * (k3 * c1 + k1 * c3);
        cd s2 = I * mu * (k3 * c2 + k2 * c3);
        cd s3 = I * (modulus * k3 * c3 + lambda * (k1 * c1 + k2 * c2)) - (MUT(4) ? 0.0 : grav) * c3;
        long q = i2 * pitch + 2 * i1;
        p1[q] = p1[q] + (R)s1.real(); p1[q + 1] = p1[q + 1] + (R)s1.imag();
        p2[q] = p2[q] + (R)s2.real(); p2[q + 1] = p2[q + 1] + (R)s2.imag();
        p3[q] = p3[q] + (R)s3.real(); p3[q + 1] = p3[q + 1] + (R)s3.imag();
      }
    }
  }
  const R d = (R)(sx3 * dx3);
  for (long q = 0; q < pitch * sx2; ++q) { p1[q] /= d; p2[q] /= d; p3[q] /= d; }
}

// ---------------------------------------------------------------------------
// cerrutimodified, green.f90:520-612 (with cerrutisolmodified :571-610)
// ---------------------------------------------------------------------------
template <class R> struct CX { typedef std::complex<R> type; };

template <class R>
static void cerrutisolmodified(double mu, std::complex<R> p1, std::complex<R> p2, std::complex<R> p3,
                               double alpha, double gamma, std::complex<R>& u1, std::complex<R>& u2,
                               std::complex<R>& u3, double k1, double k2, double k3) {
  const cd i(0.0, 1.0);
  double beta = PI2 * std::sqrt(k1 * k1 + k2 * k2);
  double fi = (2 * beta) / ((PI2 * PI2) * (k3 * k3) + beta * beta);
  double h = gamma / beta;
  if (0 == k1 && 0 == k2) {
    u1 = u2 = u3 = std::complex<R>(0, 0);
    return;
  }
  cd P1((double)p1.real(), (double)p1.imag()), P2((double)p2.real(), (double)p2.imag()),
      P3((double)p3.real(), (double)p3.imag());
  double temp = (1.0 / (2.0 * mu * beta * beta * beta)) * fi;
  cd b1 = temp * P1, b2 = temp * P2;
  cd b3 = (beta * P3 - (MUT(6) ? -i : i) * (1.0 - alpha) * (PI2 * k1 * P1 + PI2 * k2 * P2)) /
          (2.0 * alpha * mu * beta * beta * beta * beta * (MUT(5) ? 1.0 : 1 + h));
  cd tmp = alpha * i * beta * PI2 * b3 * (1.0 - 1.0 / alpha - i * PI2 * k3 * fi);
  cd v1 = MUT(7) ? cd(0.0) : tmp * k1 * fi, v2 = tmp * k2 * fi;
  cd v3 = -alpha * beta * beta * b3 * (1.0 / alpha - i * fi * PI2 * k3);
  cd w1 = v1 + (-(2.0 * beta * beta * b1) + (alpha * (PI2 * PI2) * k1 * (b1 * k1 + b2 * k2) * (1.0 - (i * PI2 * k3 * fi))));
  cd w2 = v2 + (-(2.0 * beta * beta * b2) + (alpha * (PI2 * PI2) * k2 * (b1 * k1 + b2 * k2) * (1.0 - (i * PI2 * k3 * fi))));
  cd w3 = v3 * fi + alpha * beta * PI2 * (k1 * b1 + k2 * b2) * PI2 * k3 * fi;
  u1 = std::complex<R>((R)w1.real(), (R)w1.imag());
  u2 = std::complex<R>((R)w2.real(), (R)w2.imag());
  u3 = std::complex<R>((R)w3.real(), (R)w3.imag());
}

template <class R>
static void cerrutimodified(const R* p1, const R* p2, const R* p3, double lambda, double mu, double gamma,
                            R* u1, R* u2, R* u3, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3) {
  const double alpha = (lambda + mu) / (lambda + 2 * mu);
  const long pitch = sx1 + 2;
#pragma omp parallel for
  for (int i2 = 0; i2 < sx2; ++i2) {
    double k2 = wavenumber23(i2, sx2, dx2);
    for (int i1 = 0; i1 < sx1 / 2 + 1; ++i1) {
      double k1 = wavenumber1(i1, sx1, dx1);
      long q = i2 * pitch + 2 * i1;
      std::complex<R> t1(p1[q], p1[q + 1]), t2(p2[q], p2[q + 1]), t3(p3[q], p3[q + 1]);
      for (int i3 = 0; i3 < sx3; ++i3) {
        double k3 = wavenumber23(i3, sx3, dx3);
        std::complex<R> b1, b2, b3;
        cerrutisolmodified<R>(mu, t1, t2, t3, alpha, gamma, b1, b2, b3, k1, k2, k3);
        long o = (i2 + (long)sx2 * i3) * pitch + 2 * i1;
        u1[o] += b1.real(); u1[o + 1] += b1.imag();
        u2[o] += b2.real(); u2[o + 1] += b2.imag();
        u3[o] += b3.real(); u3[o + 1] += b3.imag();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// greenfunctioncowling, green.f90:1008-1129 (non-CUDA, cerrutimodified path)
// ---------------------------------------------------------------------------
template <class R>
static int greenfunctioncowling(R* c1, R* c2, R* c3, R* t1, R* t2, R* t3, int sx1, int sx2, int sx3,
                                double dx1, double dx2, double dx3, double lambda, double mu, double gamma) {
  const long pitch = sx1 + 2;
  std::vector<R> p1(pitch * sx2, 0), p2(pitch * sx2, 0), p3(pitch * sx2, 0);
  int rc = 0;
  rc |= fft3(c1, sx1, sx2, sx3, dx1, dx2, dx3, -1);
  rc |= fft3(c2, sx1, sx2, sx3, dx1, dx2, dx3, -1);
  rc |= fft3(c3, sx1, sx2, sx3, dx1, dx2, dx3, -1);
  rc |= fft2(t1, sx1, sx2, dx1, dx2, -1);
  rc |= fft2(t2, sx1, sx2, dx1, dx2, -1);
  rc |= fft2(t3, sx1, sx2, dx1, dx2, -1);
  if (rc) return rc;
  elasticresponse(lambda, mu, c1, c2, c3, sx1, sx2, sx3, dx1, dx2, dx3);
  surfacetractioncowling(lambda, mu, gamma, c1, c2, c3, sx1, sx2, sx3, dx1, dx2, dx3, p1.data(), p2.data(), p3.data());
  for (long q = 0; q < pitch * sx2; ++q) {
    p1[q] = t1[q] - p1[q]; p2[q] = t2[q] - p2[q]; p3[q] = t3[q] - p3[q];
  }
  cerrutimodified(p1.data(), p2.data(), p3.data(), lambda, mu, gamma, c1, c2, c3, sx1, sx2, sx3, dx1, dx2, dx3);
  fft3(c1, sx1, sx2, sx3, dx1, dx2, dx3, 1);
  fft3(c2, sx1, sx2, sx3, dx1, dx2, dx3, 1);
  fft3(c3, sx1, sx2, sx3, dx1, dx2, dx3, 1);
  fft2(t1, sx1, sx2, dx1, dx2, 1);
  fft2(t2, sx1, sx2, dx1, dx2, 1);
  fft2(t3, sx1, sx2, dx1, dx2, 1);
  return 0;
}

// ---------------------------------------------------------------------------
// FIR differentiators: optimalfilter, elastic3d.f90:846-907; taps kernel1.inc
// (fir1 = 0.5) and kernel7.inc (single-precision literals stored in REAL*8).
// ---------------------------------------------------------------------------
static void optimalfilter1(int sx, bool vertical, double* ker, int* len, int* err) {
  static const float fir7f[7] = {8.77856e-01f, -2.81913e-01f, +6.22696e-02f, +2.82441e-02f,
                                 -5.09029e-02f, +4.20471e-02f, -1.59409e-02f};
  if (!vertical && sx >= 2 && sx <= 4) { *len = 1; ker[0] = 0.5; return; }
  if (sx >= 5 && sx <= 14) { *len = 7; for (int l = 0; l < 7; ++l) ker[l] = (double)fir7f[l]; return; }
  if (sx >= 15) { *len = 1; ker[0] = 0.5; return; }
  *err = 2;
}

template <class R> static inline R rnd(double x) { return (R)x; }

// isotropicstressstrain, elastic3d.f90:129-142; tensortrace :287 sums in the
// storage precision before widening.
template <class R>
static inline void isotropicstressstrain(T6<R>& t, double lambda, double mu) {
  double epskk = (double)(t.s11 + t.s22 + t.s33);
  R m2 = (R)(2.0 * mu);
  t.s11 *= m2; t.s12 *= m2; t.s13 *= m2; t.s22 *= m2; t.s23 *= m2; t.s33 *= m2;
  t.s11 = (R)(t.s11 + lambda * epskk);
  t.s22 = (R)(t.s22 + lambda * epskk);
  t.s33 = (R)(t.s33 + lambda * epskk);
}

// ---------------------------------------------------------------------------
// stressupdate, elastic3d.f90:956-1180 (localstrain_fir2 :1032,
// localstrain_ani :1140).  v is padded (sx1+2, sx2, sx3) with sx3 the depth
// extent of the tensor field (half the FFT grid in the RK2 loop).
// ---------------------------------------------------------------------------
template <class R>
static int stressupdate(const R* v1, const R* v2, const R* v3, double lambda, double mu, double dx1, double dx2,
                        double dx3, int sx1, int sx2, int sx3, R* sig, long vpitch) {
  double ker1[16], ker2[16], ker3[16];
  int len1 = 0, len2 = 0, len3 = 0, err = 0;
  optimalfilter1(sx1, false, ker1, &len1, &err);
  optimalfilter1(sx2, false, ker2, &len2, &err);
  optimalfilter1(sx3, true, ker3, &len3, &err);
  if (err) return err;
  for (int l = 0; l < 16; ++l) { ker1[l] /= dx1; ker2[l] /= dx2; ker3[l] /= dx3; }
  auto V = [&](const R* v, int i1, int i2, int i3) { return v[i1 + vpitch * (i2 + (long)sx2 * i3)]; };
  T6<R>* S = reinterpret_cast<T6<R>*>(sig);
#pragma omp parallel for
  for (int i3 = 0; i3 < sx3; ++i3) {
    bool boundary = !((i3 + 1 > len3) && (i3 + 1 < sx3 - len3 + 1));
    int i3p = 0, i3m = 0;
    double px3 = 0;
    if (boundary) {
      if (i3 == 0) { px3 = MUT(8) ? 2.0 * dx3 : dx3; i3p = 1; i3m = 0; }
      else if (i3 == sx3 - 1) { px3 = dx3; i3p = sx3 - 1; i3m = sx3 - 2; }
      else { px3 = dx3 * 2.0; i3m = i3 - 1; i3p = i3 + 1; }
    }
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        T6<R> e = {0, 0, 0, 0, 0, 0};
        for (int l = 1; l <= len1; ++l) {
          int m = (sx1 + i1 - l) % sx1, p = (i1 + l) % sx1;
          e.s11 = rnd<R>(e.s11 + (double)(V(v1, p, i2, i3) - V(v1, m, i2, i3)) * ker1[l - 1]);
          e.s12 = rnd<R>(e.s12 + (double)(V(v2, p, i2, i3) - V(v2, m, i2, i3)) * ker1[l - 1]);
          e.s13 = rnd<R>(e.s13 + (double)(V(v3, p, i2, i3) - V(v3, m, i2, i3)) * ker1[l - 1]);
        }
        for (int l = 1; l <= len2; ++l) {
          int m = (sx2 + i2 - l) % sx2, p = (i2 + l) % sx2;
          e.s12 = rnd<R>(e.s12 + (double)(V(v1, i1, p, i3) - V(v1, i1, m, i3)) * ker2[l - 1]);
          e.s22 = rnd<R>(e.s22 + (double)(V(v2, i1, p, i3) - V(v2, i1, m, i3)) * ker2[l - 1]);
          e.s23 = rnd<R>(e.s23 + (double)(V(v3, i1, p, i3) - V(v3, i1, m, i3)) * ker2[l - 1]);
        }
        if (boundary) {
          e.s13 = rnd<R>(e.s13 + (double)(V(v1, i1, i2, i3p) - V(v1, i1, i2, i3m)) / px3);
          e.s23 = rnd<R>(e.s23 + (double)(V(v2, i1, i2, i3p) - V(v2, i1, i2, i3m)) / px3);
          e.s33 = rnd<R>((double)(V(v3, i1, i2, i3p) - V(v3, i1, i2, i3m)) / px3);
        } else {
          for (int l = 1; l <= len3; ++l) {
            int m = MUT(25) ? i3 : i3 - l, p = i3 + l;
            e.s13 = rnd<R>(e.s13 + (double)(V(v1, i1, i2, p) - V(v1, i1, i2, m)) * ker3[l - 1]);
            e.s23 = rnd<R>(e.s23 + (double)(V(v2, i1, i2, p) - V(v2, i1, i2, m)) * ker3[l - 1]);
            e.s33 = rnd<R>(e.s33 + (double)(V(v3, i1, i2, p) - V(v3, i1, i2, m)) * ker3[l - 1]);
          }
        }
        e.s12 = rnd<R>(e.s12 / 2.0);
        e.s13 = rnd<R>(e.s13 / 2.0);
        e.s23 = rnd<R>(e.s23 / 2.0);
        isotropicstressstrain(e, lambda, mu);
        T6<R>& s = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
        s.s11 += e.s11; s.s12 += e.s12; s.s13 += e.s13;
        s.s22 += e.s22; s.s23 += e.s23; s.s33 += e.s33;
      }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// equivalentbodyforce, elastic3d.f90:1214-1506 (localdivergence_fir2 :1324,
// localdivergence_ani :1431); FP64 accumulators, c -= REAL(f).
// ---------------------------------------------------------------------------
template <class R>
static int equivalentbodyforce(const R* sig, double dx1, double dx2, double dx3, int sx1, int sx2, int sx3,
                               R* c1, R* c2, R* c3, R* t1, R* t2, R* t3, long cpitch, const R* mask) {
  double ker1[16], ker2[16], ker3[16];
  int len1 = 0, len2 = 0, len3 = 0, err = 0;
  optimalfilter1(sx1, false, ker1, &len1, &err);
  optimalfilter1(sx2, false, ker2, &len2, &err);
  optimalfilter1(sx3, true, ker3, &len3, &err);
  if (err) return err;
  for (int l = 0; l < 16; ++l) { ker1[l] /= dx1; ker2[l] /= dx2; ker3[l] /= dx3; }
  const T6<R>* S = reinterpret_cast<const T6<R>*>(sig);
  auto at = [&](int i1, int i2, int i3) -> const T6<R>& { return S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)]; };
  for (int i2 = 0; i2 < sx2; ++i2)
    for (int i1 = 0; i1 < sx1; ++i1) {
      t1[i1 + cpitch * i2] += at(i1, i2, 0).s13;
      t2[i1 + cpitch * i2] += at(i1, i2, 0).s23;
      t3[i1 + cpitch * i2] += at(i1, i2, 0).s33;
    }
#pragma omp parallel for
  for (int i3 = 0; i3 < sx3; ++i3) {
    if (mask && mask[i3] == 0) continue;
    bool boundary = !((i3 + 1 > len3) && (i3 + 1 < sx3 - len3 + 1));
    int i3p = 0, i3m = 0;
    double px3 = 0;
    if (boundary) {
      if (i3 == 0) { px3 = dx3; i3p = 1; i3m = 0; }
      else if (i3 == sx3 - 1) { px3 = dx3; i3p = sx3 - 1; i3m = sx3 - 2; }
      else { px3 = dx3 * 2.0; i3m = i3 - 1; i3p = i3 + 1; }
    }
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        double f1 = 0, f2 = 0, f3 = 0;
        for (int l = 1; l <= len1; ++l) {
          int m = (sx1 + i1 - l) % sx1, p = (i1 + l) % sx1;
          f1 += (double)(at(p, i2, i3).s11 - at(m, i2, i3).s11) * ker1[l - 1];
          f2 += (double)(at(p, i2, i3).s12 - at(m, i2, i3).s12) * ker1[l - 1];
          f3 += (double)(at(p, i2, i3).s13 - at(m, i2, i3).s13) * ker1[l - 1];
        }
        for (int l = 1; l <= len2; ++l) {
          int m = (sx2 + i2 - l) % sx2, p = (i2 + l) % sx2;
          f1 += (double)(at(i1, p, i3).s12 - at(i1, m, i3).s12) * ker2[l - 1];
          f2 += (double)(at(i1, p, i3).s22 - at(i1, m, i3).s22) * (MUT(9) ? -ker2[l - 1] : ker2[l - 1]);
          f3 += (double)(at(i1, p, i3).s23 - at(i1, m, i3).s23) * ker2[l - 1];
        }
        if (boundary) {
          f1 += (double)(at(i1, i2, i3p).s13 - at(i1, i2, i3m).s13) / px3;
          f2 += (double)(at(i1, i2, i3p).s23 - at(i1, i2, i3m).s23) / px3;
          f3 += (double)(at(i1, i2, i3p).s33 - at(i1, i2, i3m).s33) / px3;
        } else {
          for (int l = 1; l <= len3; ++l) {
            int m = i3 - l, p = MUT(26) ? i3 : i3 + l;
            f1 += (double)(at(i1, i2, p).s13 - at(i1, i2, m).s13) * ker3[l - 1];
            f2 += (double)(at(i1, i2, p).s23 - at(i1, i2, m).s23) * ker3[l - 1];
            f3 += (double)(at(i1, i2, p).s33 - at(i1, i2, m).s33) * ker3[l - 1];
          }
        }
        long o = i1 + cpitch * (i2 + (long)sx2 * i3);
        c1[o] = c1[o] - (R)f1;
        c2[o] = c2[o] - (R)f2;
        c3[o] = c3[o] - (R)f3;
      }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// tensor helpers, elastic3d.f90:150-330
// ---------------------------------------------------------------------------
template <class R> static inline T6<R> tensordeviatoric(const T6<R>& t) {
  R diag = (R)((double)(t.s11 + t.s22 + t.s33) / 3.0);
  return T6<R>{t.s11 - diag, t.s12, t.s13, t.s22 - diag, t.s23, t.s33 - diag};
}
template <class R> static inline double tensornorm(const T6<R>& t) {
  return std::sqrt(((double)(t.s11 * t.s11) + 2.0 * (double)(t.s12 * t.s12) + 2.0 * (double)(t.s13 * t.s13) +
                    (double)(t.s22 * t.s22) + 2.0 * (double)(t.s23 * t.s23) + (double)(t.s33 * t.s33)) / 2.0);
}

// ---------------------------------------------------------------------------
// viscouseigenstress, viscoelastic3d.f90:217-314
// ---------------------------------------------------------------------------
template <class R>
static int viscouseigenstress(double mu, const orc_layer* structure, const orc_tensor_layer* prestress,
                              const R* sig, int sx1, int sx2, int sx3, R* moment, double* maxwelltime,
                              const R* dgammadot0, R* gamma, long gpitch) {
  const T6<R>* S = reinterpret_cast<const T6<R>*>(sig);
  T6<R>* M = reinterpret_cast<T6<R>*>(moment);
  for (int i3 = 0; i3 < sx3; ++i3)
    if (structure[i3].stressexponent < 0.999999) return 1;
  float tm = maxwelltime ? (float)*maxwelltime : 1e30f;
#pragma omp parallel for reduction(min : tm)
  for (int i3 = 0; i3 < sx3; ++i3) {
    double power = structure[i3].stressexponent;
    double cohesion = structure[i3].cohesion;
    T6<R> pt = {(R)prestress[i3].t[0], (R)prestress[i3].t[1], (R)prestress[i3].t[2],
                (R)prestress[i3].t[3], (R)prestress[i3].t[4], (R)prestress[i3].t[5]};
    T6<R> sp = tensordeviatoric(pt);
    double taup = tensornorm(sp);
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        long o = i1 + (long)sx1 * (i2 + (long)sx2 * i3);
        double gammadot0 = structure[i3].gammadot0;
        if (dgammadot0) gammadot0 = gammadot0 + dgammadot0[o];
        if (1.0e-20 > gammadot0) continue;
        T6<R> s = tensordeviatoric(S[o]);
        T6<R> ss = {s.s11 + sp.s11, s.s12 + sp.s12, s.s13 + sp.s13, s.s22 + sp.s22, s.s23 + sp.s23, s.s33 + sp.s33};
        double tau = tensornorm(ss);
        double tauc = std::max(0.0, tau - cohesion);
        if (tauc <= 1.0e-20) continue;
        double gammadot = gammadot0 * std::pow(tauc / mu, MUT(10) ? power : power - 1);
        double gammadotp = gammadot0 * std::pow(taup / mu, power - 1);
        R a = (R)(2.0 * gammadot), b = (R)(2.0 * gammadotp);
        T6<R>& m = M[o];
        m.s11 = m.s11 + (a * s.s11 - b * sp.s11);
        m.s12 = m.s12 + (a * s.s12 - b * sp.s12);
        m.s13 = m.s13 + (a * s.s13 - b * sp.s13);
        m.s22 = m.s22 + (a * s.s22 - b * sp.s22);
        m.s23 = m.s23 + (a * s.s23 - b * sp.s23);
        m.s33 = m.s33 + (a * s.s33 - b * sp.s33);
        tm = std::min(tm, (float)(1.0 / gammadot));
        if (gamma) gamma[i1 + gpitch * (i2 + (long)sx2 * i3)] = (R)(gammadot * tauc / mu);
      }
  }
  if (maxwelltime) *maxwelltime = std::min((double)tm, *maxwelltime);
  return 0;
}

// ---------------------------------------------------------------------------
// gauss, omega, my_sinh, shiftedcoordinates, fix
// (elastic3d.f90:525-531, :549-562, :89-97, :4027-4048, :70-83)
// ---------------------------------------------------------------------------
static inline double gauss(double x, double sigma) {
  return std::exp(-0.5 * (x / sigma) * (x / sigma)) / std::sqrt(PI2) / sigma;
}
static inline double omega(double x, double beta) {
  double ax = std::fabs(x);
  if (ax <= (1.0 - 2.0 * beta) / (1.0 - beta) / 2.0) return 1.0;
  if (ax < 1.0 / (1 - beta) / 2.0) {
    double c = std::cos(PI * ((1.0 - beta) * ax - 0.5 + beta) / 2.0 / beta);
    return c * c;
  }
  return 0.0;
}
static inline double my_sinh(double x) {
  if (std::fabs(x) > 11.0) return std::copysign(1.0, x) * std::sinh(11.0);
  return std::sinh(x);
}
static inline double shifted1(int i /*0-based*/, int sx, double dx) {
  int ib = i + 1;
  if (ib <= sx / 2) return (double)(ib - 1) * dx;
  return (double)(ib - sx - 1) * dx;
}
static inline int fix(double number) {
  double f = std::floor(number), c = std::ceil(number);
  return (number - f > 0.5) ? (int)c : (int)f;
}

// ---------------------------------------------------------------------------
// transienteigenstress, viscoelastic3d.f90:123-202 (Burgers transient creep):
//   Q = s' - 2 Gk E,  gammadot = gammadot0 (|Q|/mu)^(n-1),
//   Edot = gammadot/mu Q,  moment += 2 gammadot Q,  tm = min 1/gammadot.
// epsilonik and epsilonikdot may be the same array (the corrector passes
// epsilonikdot for both, relax.f90:768-790): each point reads its E before
// it writes its Edot.
// ---------------------------------------------------------------------------
template <class R>
static int transienteigenstress(double mu, const orc_layer* structure, const R* sig, const R* epsilonik, int sx1,
                                int sx2, int sx3, R* moment, R* epsilonikdot, double* maxwelltime,
                                const R* dgammadot0) {
  const T6<R>* S = reinterpret_cast<const T6<R>*>(sig);
  const T6<R>* E = reinterpret_cast<const T6<R>*>(epsilonik);
  T6<R>* M = reinterpret_cast<T6<R>*>(moment);
  T6<R>* Ed = reinterpret_cast<T6<R>*>(epsilonikdot);
  for (int i3 = 0; i3 < sx3; ++i3)
    if (structure[i3].stressexponent < 0.999999) return 1;
  float tm = maxwelltime ? (float)*maxwelltime : 1e30f;
#pragma omp parallel for reduction(min : tm)
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double power = structure[i3].stressexponent;
    const double muk = structure[i3].Gk;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const long o = i1 + (long)sx1 * (i2 + (long)sx2 * i3);
        double gammadot0 = structure[i3].gammadot0;
        if (dgammadot0) gammadot0 = gammadot0 + dgammadot0[o];
        if (1.0e-20 > gammadot0) {
          Ed[o] = T6<R>{0, 0, 0, 0, 0, 0};
          continue;
        }
        const T6<R> s = tensordeviatoric(S[o]);
        const R k2 = (R)(MUT(14) ? muk : 2.0 * muk);
        const T6<R> e = E[o];
        const T6<R> Q = {s.s11 - k2 * e.s11, s.s12 - k2 * e.s12, s.s13 - k2 * e.s13,
                         s.s22 - k2 * e.s22, s.s23 - k2 * e.s23, s.s33 - k2 * e.s33};
        const double nq = tensornorm(Q);
        const double gammadot = gammadot0 * std::pow(nq / mu, MUT(15) ? power : power - 1);
        const R a = (R)(gammadot / mu);
        Ed[o] = T6<R>{a * Q.s11, a * Q.s12, a * Q.s13, a * Q.s22, a * Q.s23, a * Q.s33};
        const R b = (R)(2.0 * gammadot);
        T6<R>& m = M[o];
        m.s11 = m.s11 + b * Q.s11; m.s12 = m.s12 + b * Q.s12; m.s13 = m.s13 + b * Q.s13;
        m.s22 = m.s22 + b * Q.s22; m.s23 = m.s23 + b * Q.s23; m.s33 = m.s33 + b * Q.s33;
        tm = std::min(tm, (float)(1.0 / gammadot));
      }
  }
  if (maxwelltime) *maxwelltime = std::min((double)tm, *maxwelltime);
  return 0;
}

// ---------------------------------------------------------------------------
// builddgammadot0, viscoelastic3d.f90:320-435 (fdgammadot0 :351-433): the
// fluidity perturbation of weak zones, tapered by omega; boxes, spheres and
// the three cylinder orientations chosen by the signs of D, W, L.
// ---------------------------------------------------------------------------
extern "C" {
struct orc_weak {
  double dgammadot0, x, y, z, width, length, thickness, strike, dip;
  float e[6];
};
}

static double fdgammadot0(const orc_weak* zones, int n, double x1, double x2, double x3, double beta) {
  double f = 0.0;
  for (int i = 0; i < n; ++i) {
    const orc_weak& Z = zones[i];
    const double dg = Z.dgammadot0, x = Z.x, y = Z.y, z = Z.z;
    // the zone's "length" is the x2 extent W and its "width" the x3 extent L
    const double W = Z.length, L = Z.width, D = Z.thickness;
    const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0,
                 Dp = D * (1.0 + 2.0 * beta) / 2.0;
    const double LM = std::max(Wp, std::max(Lp, Dp));
    if ((std::fabs(x3 - z) > LM) || (std::fabs(x1 - x) > LM) || (std::fabs(x2 - y) > LM)) continue;
    const double cs = std::cos(Z.strike), ss = std::sin(Z.strike), cd_ = std::cos(Z.dip), sd = std::sin(Z.dip);
    double x2r = cs * x - ss * y;
    const double xr = cd_ * x2r - sd * z, yr = ss * x + cs * y, zr = sd * x2r + cd_ * z;
    x2r = cs * x1 - ss * x2;
    const double x1s = cd_ * x2r - sd * x3, x2s = ss * x1 + cs * x2, x3s = sd * x2r + cd_ * x3;
    const double a = (x1s - xr) / D, b = (x2s - yr) / W, c = (x3s - zr) / L;
    if (D >= 0 && W >= 0 && L >= 0) {
      f += omega(a, beta) * omega(b, beta) * omega(MUT(16) ? (x3s + zr) / L : c, beta) * dg;
    } else if (D < 0 && W < 0 && L < 0) {
      f += omega(std::sqrt(a * a + b * b + c * c), beta) * dg;
    } else if (D < 0 && W < 0) {
      f += omega(std::sqrt(a * a + b * b), beta) * omega(c, beta) * dg;
    } else if (D < 0 && L < 0) {
      f += omega(std::sqrt(a * a + c * c), beta) * omega(b, beta) * dg;
    } else if (W < 0 && L < 0) {
      f += omega(std::sqrt(b * b + c * c), beta) * omega(a, beta) * dg;
    }
  }
  return f;
}

template <class R>
static void builddgammadot0(int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, double beta, int nz,
                            const orc_weak* zones, R* dg) {
#pragma omp parallel for
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = (double)i3 * dx3;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1)
        dg[i1 + (long)sx1 * (i2 + (long)sx2 * i3)] =
            (R)fdgammadot0(zones, nz, shifted1(i1, sx1, dx1), shifted1(i2, sx2, dx2), x3, beta);
  }
}

// ---------------------------------------------------------------------------
// frictioneigenstress, friction3d.f90:82-241 (CPU path)
// ---------------------------------------------------------------------------
template <class R>
static int frictioneigenstress(double x, double y, double z, double L, double W, double strike, double dip,
                               double rake, double beta, const R* sig, double mu, const orc_layer* structure,
                               int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, R* moment,
                               double* maxwelltime, R* vel, long vpitch) {
  const T6<R>* S = reinterpret_cast<const T6<R>*>(sig);
  T6<R>* M = reinterpret_cast<T6<R>*>(moment);
  float tm = maxwelltime ? (float)*maxwelltime : 1e30f;
  const double scaling = std::sqrt(PI2) * dx1;
  const double cstrike = std::cos(strike), sstrike = std::sin(strike);
  const double cdip = std::cos(dip), sdip = std::sin(dip);
  const double cr = std::cos(rake), sr = std::sin(rake);
  const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0;
  double x2r = cstrike * x - sstrike * y;
  const double xr = cdip * x2r - sdip * z;
  const double yr = sstrike * x + cstrike * y;
  const double zr = sdip * x2r + cdip * z;
  const double n[3] = {+cdip * cstrike, -cdip * sstrike, -sdip};
  const double r[3] = {sstrike * cr + cstrike * sdip * sr, cstrike * cr - sstrike * sdip * sr, cdip * sr};
  int status = 0;
#pragma omp parallel for reduction(min : tm) reduction(max : status)
  for (int i3 = 0; i3 < sx3; ++i3) {
    double x3 = (double)i3 * dx3;
    if ((std::fabs(x3 - z) > Lp) && (std::fabs(x3 + z) > Lp)) continue;
    double vo = structure[i3].gammadot0, tauc = structure[i3].stressexponent;
    double friction = structure[i3].friction, cohesion = structure[i3].cohesion;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - x) > std::max(Wp, Lp)) || (std::fabs(x2 - y) > std::max(Wp, Lp))) continue;
        double x2rr = cstrike * x1 - sstrike * x2;
        double x1s = cdip * x2rr - sdip * x3;
        double x1i = cdip * x2rr + sdip * x3;
        if ((std::fabs(x1s - xr) > 7.01 * dx1) && (std::fabs(x1i - xr) > 7.01 * dx1)) continue;
        double x2s = sstrike * x1 + cstrike * x2;
        double x3s = sdip * x2rr + cdip * x3;
        double x3i = -sdip * x2rr + cdip * x3;
        double temp1 = gauss(x1s - xr, dx1);
        double temp2 = omega((x2s - yr) / W, beta);
        double temp3 = omega((x3s - zr) / L, beta);
        double sourc = temp1 * temp2 * temp3;
        temp1 = gauss(x1i - xr, dx1);
        temp3 = omega((x3i + zr) / L, beta);
        double image = temp1 * temp2 * temp3;
        double impulse = MUT(11) ? sourc : sourc + image;
        long o = i1 + (long)sx1 * (i2 + (long)sx2 * i3);
        const T6<R>& s = S[o];
        double t[3] = {s.s11 * n[0] + s.s12 * n[1] + s.s13 * n[2], s.s12 * n[0] + s.s22 * n[1] + s.s23 * n[2],
                       s.s13 * n[0] + s.s23 * n[1] + s.s33 * n[2]};
        double taun = t[0] * n[0] + t[1] * n[1] + t[2] * n[2];
        double ts[3] = {t[0] - taun * n[0], t[1] - taun * n[1], t[2] - taun * n[2]};
        double taus = std::sqrt(ts[0] * ts[0] + ts[1] * ts[1] + ts[2] * ts[2]);
        double tau = std::max(0.0, taus + friction * taun - cohesion);
        if ((ts[0] * r[0] + ts[1] * r[1] + ts[2] * r[2]) < 0.0 && std::fabs(rake) < PI2 * 1.5) continue;
        if ((tau / tauc) > 20) { status = 5; continue; }
        ts[0] /= taus; ts[1] /= taus; ts[2] /= taus;
        double gammadot = vo * 2.0 * my_sinh(tau / tauc);
        float cand = (float)(tau / mu / gammadot * (std::min(L, W) / std::sqrt(dx1 * (MUT(27) ? dx2 : dx3))));
        if (cand < tm) tm = cand;  // MIN ignores a NaN candidate
        if (vel) {
          long vo2 = i1 + vpitch * (i2 + (long)sx2 * i3);
          vel[vo2] = (R)(vel[vo2] + gammadot * impulse * scaling);
        }
        double a = 2.0 * mu * impulse * gammadot;
        double b[3] = {a * n[0], a * n[1], a * n[2]};
        T6<R>& m = M[o];
        m.s11 += (R)(ts[0] * b[0]);
        m.s12 += (R)((ts[0] * b[1] + ts[1] * b[0]) / 2.0);
        m.s13 += (R)((ts[0] * b[2] + ts[2] * b[0]) / 2.0);
        m.s22 += (R)(ts[1] * b[1]);
        m.s23 += (R)((ts[1] * b[2] + ts[2] * b[1]) / 2.0);
        m.s33 += (R)(ts[2] * b[2]);
      }
  }
  if (status) return status;
  if (maxwelltime) *maxwelltime = std::min((double)tm, *maxwelltime);
  return 0;
}

// ---------------------------------------------------------------------------
// shiftedindex, elastic3d.f90:4063-4101 (0-based indices; status 15 where the
// reference stops on a coordinate out of range)
// ---------------------------------------------------------------------------
static int shiftedindex(double x1, double x2, double x3, int sx1, int sx2, int sx3, double dx1, double dx2,
                        double dx3, int* i1, int* i2, int* i3) {
  if (x1 > (double)(sx1 / 2 - 1) * dx1 || x1 < -(double)(sx1 / 2) * dx1) return 15;
  if (x2 > (double)(sx2 / 2 - 1) * dx2 || x2 < -(double)(sx2 / 2) * dx2) return 15;
  if (x3 > (double)(sx3 - 1) * dx3 || x3 < 0.0) return 15;
  *i1 = (sx1 + fix(x1 / dx1)) % sx1;
  *i2 = (sx2 + fix(x2 / dx2)) % sx2;
  *i3 = MUT(17) ? (int)std::lround(x3 / dx3) : fix(x3 / dx3);
  return 0;
}

// exportpoints, util.f90:59-139 (the values it writes): u and sig at the
// shiftedindex of each observation point, u turned into the user frame by
// rotation(v1, v2, -rot) (input.f90:1809-1822).  out: n x 9 (v1 v2 v3 s11
// s12 s13 s22 s23 s33).
template <class R>
static int exportpoints(const R* u1, const R* u2, const R* u3, const R* sig, int sx1, int sx2, int sx3, double dx1,
                        double dx2, double dx3, const double* pts, int n, double rot, double* out) {
  const long pitch = sx1 + 2;
  const T6<R>* S = reinterpret_cast<const T6<R>*>(sig);
  for (int k = 0; k < n; ++k) {
    int i1, i2, i3;
    if (int rc = shiftedindex(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2], sx1, sx2, sx3, dx1, dx2, dx3, &i1, &i2,
                              &i3))
      return rc;
    const long o = i1 + pitch * (i2 + (long)sx2 * i3);
    double v1 = (double)u1[o], v2 = (double)u2[o];
    const double a = -rot * DEG2RAD, c = std::cos(a), s_ = std::sin(a);
    const double w1 = +v1 * c + v2 * s_, w2 = -v1 * s_ + v2 * c;
    const T6<R>& t = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
    double* r = out + 9 * k;
    r[0] = w1; r[1] = w2; r[2] = (double)u3[o];
    r[3] = t.s11; r[4] = t.s12; r[5] = t.s13; r[6] = t.s22; r[7] = t.s23; r[8] = t.s33;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// monitorfriction, friction3d.f90:263-416; shiftedindex elastic3d.f90:4063-4101
// ---------------------------------------------------------------------------
template <class R>
static int monitorfriction(double x, double y, double z, double L, double W, double strike, double dip,
                           double rake, double beta, int sx1, int sx2, int sx3, int sx3_alloc, double dx1,
                           double dx2, double dx3, const R* sig, const orc_layer* structure, orc_patch* patch,
                           int px2, int px3) {
  (void)L; (void)W; (void)beta;
  const T6<R>* S = reinterpret_cast<const T6<R>*>(sig);
  const double cstrike = std::cos(strike), sstrike = std::sin(strike);
  const double cdip = std::cos(dip), sdip = std::sin(dip);
  const double cr = std::cos(rake), sr = std::sin(rake);
  const double sv[3] = {sstrike, cstrike, 0.0};
  const double dv[3] = {-cstrike * sdip, +sstrike * sdip, -cdip};
  const double n[3] = {+cdip * cstrike, -cdip * sstrike, -sdip};
  const double r[3] = {sstrike * cr + cstrike * sdip * sr, cstrike * cr - sstrike * sdip * sr, cdip * sr};
  for (int j3 = 1; j3 <= px3; ++j3)
    for (int j2 = 1; j2 <= px2; ++j2) {
      orc_patch& P = patch[(j2 - 1) + (long)px2 * (j3 - 1)];
      double x2 = cstrike * x - sstrike * y;
      double xr = cdip * x2 - sdip * z;
      double yr = sstrike * x + cstrike * y;
      double zr = sdip * x2 + cdip * z;
      // yr=REAL(yr)+REAL(...): a single-precision sum stored in REAL*8 (friction3d.f90:335-336)
      yr = (double)((float)yr + (float)(((double)j2 - (double)px2 / 2.0 - 1.0) * (MUT(28) ? dx3 : dx2)));
      zr = (double)((float)zr + (float)(((double)j3 - (double)px3 / 2.0 - 1.0) * (MUT(28) ? dx2 : dx3)));
      double x2p = cdip * xr + sdip * zr;
      double X1 = cstrike * x2p + sstrike * yr;
      double X2 = -sstrike * x2p + cstrike * yr;
      double X3 = (MUT(30) ? sdip : -sdip) * xr + cdip * zr;
      P.x1 = X1; P.x2 = X2; P.x3 = X3; P.lx = yr; P.lz = zr;
      if ((X1 > (double)(sx1 / 2 - 1) * dx1) || (X1 < -(double)(sx1 / 2) * dx1) ||
          (X2 > (double)(sx2 / 2 - 1) * dx2) || (X2 < -(double)(sx2 / 2) * dx2) ||
          (X3 > (double)(sx3 - 1) * dx3) || (X3 < 0.0))
        continue;
      int i1 = (sx1 + fix(X1 / dx1)) % sx1;
      int i2 = (sx2 + fix(X2 / dx2)) % sx2;
      int i3 = fix(X3 / dx3);
      if (i3 >= sx3_alloc) return 15;  // beyond the allocated stress field
      double vo = structure[i3].gammadot0, tauc = structure[i3].stressexponent;
      double friction = structure[i3].friction, cohesion = structure[i3].cohesion;
      const T6<R>& s = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
      P.sig[0] = (float)s.s11; P.sig[1] = (float)s.s12; P.sig[2] = (float)s.s13;
      P.sig[3] = (float)s.s22; P.sig[4] = (float)s.s23; P.sig[5] = (float)s.s33;
      double t[3] = {s.s11 * n[0] + s.s12 * n[1] + s.s13 * n[2], s.s12 * n[0] + s.s22 * n[1] + s.s23 * n[2],
                     s.s13 * n[0] + s.s23 * n[1] + s.s33 * n[2]};
      double taun = t[0] * n[0] + t[1] * n[1] + t[2] * n[2];
      double ts[3] = {t[0] - taun * n[0], t[1] - taun * n[1], t[2] - taun * n[2]};
      double taus = std::sqrt(ts[0] * ts[0] + ts[1] * ts[1] + ts[2] * ts[2]);
      double tau = MUT(29) ? taus : std::max(0.0, taus + friction * taun - cohesion);
      if ((ts[0] * r[0] + ts[1] * r[1] + ts[2] * r[2]) < 0.0 && std::fabs(rake) < PI2 * 1.5) continue;
      P.taus = taus;
      P.v = vo * 2.0 * my_sinh(tau / tauc);
      ts[0] /= taus; ts[1] /= taus; ts[2] /= taus;
      P.vss = P.v * (ts[0] * sv[0] + ts[1] * sv[1] + ts[2] * sv[2]);
      P.vds = P.v * (ts[0] * dv[0] + ts[1] * dv[1] + ts[2] * dv[2]);
    }
  return 0;
}

// frictionadd, friction3d.f90:424-453
static void frictionadd(orc_patch* patch, long npatch, double dt) {
  for (long k = 0; k < npatch; ++k) {
    orc_patch& P = patch[k];
    P.ss = P.ss + dt * P.vss;
    P.ds = P.ds + dt * P.vds;
    P.slip = std::pow(P.ds * P.ds + P.ss * P.ss, 0.5);
  }
}

// ---------------------------------------------------------------------------
// Okada (1992) displacement of a rectangular dislocation in a half-space:
// DC3D of okada/dc3d.f (displacement only), restated from the paper
// (Y. Okada, BSSA 82:1018-1040, Table 6: the A, B and C parts of the
// strike-slip, dip-slip and tensile sources, Chinnery's sum over the four
// corners) with dc3d.f's conventions:
//  - arguments in single precision, arithmetic in double (dc3d.f:1-5);
//  - dip constants: |cos(dip)| < 1e-6 is taken as 0, sin as +-1 (DCCON0,
//    dc3d.f:509-545);
//  - |xi|, |eta|, |q| < 1e-6 are taken as 0 (:52-67, :578-580);
//  - a point on a fault edge (q = 0 and on an edge segment) is singular:
//    the whole result is zero (:70-74, :160-176);
//  - on the negative extension of an edge, where R + xi (R + eta) vanishes,
//    ln(R + xi) is taken as -ln(R - xi) and X11 = X32 = 0 (KXI, KET,
//    :75-86, :597-617); atan(xi eta / (q R)) is 0 for q = 0 (:591-595);
//  - the real source contributes -A, its image A + B + z C with the vertical
//    part of C reversed (:89-107, :133-158).
// Checked by tests/test_oracle.py against the Navier equation, the
// traction-free surface, the displacement jump across the fault and
// Okada (1985)'s surface formulas.
// ---------------------------------------------------------------------------
namespace okada {
static const double EPS = 1.0e-6;

struct Dip { double sd, cd; };

static Dip dip_constants(double dipdeg) {
  const double p18 = 6.283185307179586 / 360.0;
  Dip D{std::sin(dipdeg * p18), std::cos(dipdeg * p18)};
  if (std::fabs(D.cd) < EPS) {
    D.cd = 0.0;
    if (D.sd > 0.0) D.sd = 1.0;
    if (D.sd < 0.0) D.sd = -1.0;
  }
  return D;
}

// f[part][component] of one corner (xi, eta, q) for the dislocation u =
// (U1, U2, U3), each part (A, B, C) already combined over the three source
// kinds and divided by 2 pi.
static void corner(double xi, double et, double q, double z, const Dip& D, double alpha, const double U[3],
                   bool kxi, bool ket, bool want_bc, double fA[3], double fB[3], double fC[3]) {
  const double sd = D.sd, cd = D.cd;
  const double R = std::sqrt(xi * xi + et * et + q * q);
  const double yt = et * cd + q * sd, dt = et * sd - q * cd;
  const double th = (q == 0.0) ? 0.0 : std::atan(xi * et / (q * R));
  double lx, X11, X32, le, Y11, Y32;
  if (kxi) { lx = -std::log(R - xi); X11 = 0.0; X32 = 0.0; }
  else { lx = std::log(R + xi); X11 = 1.0 / (R * (R + xi)); X32 = (2.0 * R + xi) / (R * R * R * (R + xi) * (R + xi)); }
  if (ket) { le = -std::log(R - et); Y11 = 0.0; Y32 = 0.0; }
  else { le = std::log(R + et); Y11 = 1.0 / (R * (R + et)); Y32 = (2.0 * R + et) / (R * R * R * (R + et) * (R + et)); }
  const double a = alpha, b = 1.0 - alpha, k = (1.0 - alpha) / alpha;
  const double pi2 = 6.283185307179586;
  const double q2 = q * q;
  // A: the infinite-medium part
  const double A[3][3] = {
      {th / 2 + a / 2 * xi * q * Y11, a / 2 * q / R, b / 2 * le - a / 2 * q2 * Y11},
      {a / 2 * q / R, th / 2 + a / 2 * et * q * X11, b / 2 * lx - a / 2 * q2 * X11},
      {-b / 2 * le - a / 2 * q2 * Y11, -b / 2 * lx - a / 2 * q2 * X11, th / 2 - a / 2 * q * (et * X11 + xi * Y11)}};
  for (int c = 0; c < 3; ++c) fA[c] = (U[0] * A[0][c] + U[1] * A[1][c] + U[2] * A[2][c]) / pi2;
  if (!want_bc) return;
  const double Rd = R + dt;
  double I3, I4;
  if (cd != 0.0) {
    if (xi == 0.0) {
      I4 = 0.0;
    } else {
      const double X = std::sqrt(xi * xi + q2);
      I4 = (xi / Rd * sd * cd + 2.0 * std::atan((et * (X + q * cd) + X * (R + X) * sd) / (xi * (R + X) * cd))) /
           (cd * cd);
    }
    I3 = (yt * cd / Rd - le + sd * std::log(Rd)) / (cd * cd);
  } else {
    I3 = (et / Rd + yt * q / (Rd * Rd) - le) / 2.0;
    I4 = xi * yt / (Rd * Rd) / 2.0;
  }
  const double I1 = -xi / Rd * cd - I4 * sd;
  const double I2 = std::log(Rd) + I3 * sd;
  const double B[3][3] = {
      {-xi * q * Y11 - th - k * I1 * sd, -q / R + k * yt / Rd * sd, q2 * Y11 - k * I2 * sd},
      {-q / R + k * I3 * sd * cd, -et * q * X11 - th - k * xi / Rd * sd * cd, q2 * X11 + k * I4 * sd * cd},
      {q2 * Y11 - k * I3 * sd * sd, q2 * X11 + k * xi / Rd * sd * sd, q * (et * X11 + xi * Y11) - th - k * I4 * sd * sd}};
  const double R3 = R * R * R;
  const double ct = dt + z, h = q * cd - z;
  const double Z32 = sd / R3 - h * Y32;
  const double C[3][3] = {
      {b * xi * Y11 * cd - a * xi * q * Z32, b * (cd / R + 2.0 * q * Y11 * sd) - a * ct * q / R3,
       b * q * Y11 * cd - a * (ct * et / R3 - z * Y11 + xi * xi * Z32)},
      {b * cd / R - q * Y11 * sd - a * ct * q / R3, b * yt * X11 - a * ct * et * q * X32,
       -dt * X11 - xi * Y11 * sd - a * ct * (X11 - q2 * X32)},
      {-b * (sd / R + q * Y11 * cd) - a * (z * Y11 - q2 * Z32),
       b * 2.0 * xi * Y11 * sd + dt * X11 - a * ct * (X11 - q2 * X32),
       b * (yt * X11 + xi * Y11 * cd) + a * q * (ct * et * X32 + xi * Z32)}};
  for (int c = 0; c < 3; ++c) {
    fB[c] = (U[0] * B[0][c] + U[1] * B[1][c] + U[2] * B[2][c]) / pi2;
    fC[c] = (U[0] * C[0][c] + U[1] * C[1][c] + U[2] * C[2][c]) / pi2;
  }
}

// DC3D's displacement in double precision; returns 1 (and zero
// displacement) at a singular point, 0 otherwise.
static int dc3d_core(double alpha, double x, double y, double z, double c, double dipdeg, double al1, double al2,
                     double aw1, double aw2, const double U[3], double u[3]) {
  const Dip D = dip_constants(dipdeg);
  double xi[2] = {x - al1, x - al2};
  for (int j = 0; j < 2; ++j)
    if (std::fabs(xi[j]) < EPS) xi[j] = 0.0;
  u[0] = u[1] = u[2] = 0.0;
  for (int section = 0; section < 2; ++section) {  // 0: real source, 1: image
    const double d = section == 0 ? c + z : c - z;
    const double p = y * D.cd + d * D.sd;
    double q = y * D.sd - d * D.cd;
    double et[2] = {p - aw1, p - aw2};
    if (std::fabs(q) < EPS) q = 0.0;
    for (int k = 0; k < 2; ++k)
      if (std::fabs(et[k]) < EPS) et[k] = 0.0;
    // on a fault edge
    if (q == 0.0 && ((xi[0] * xi[1] <= 0.0 && et[0] * et[1] == 0.0) ||
                     (et[0] * et[1] <= 0.0 && xi[0] * xi[1] == 0.0))) {
      u[0] = u[1] = u[2] = 0.0;
      return 1;
    }
    // on the negative extension of an edge: kxi per eta corner, ket per xi
    // corner (the tests dc3d.f makes, :78-86)
    const double r12 = std::sqrt(xi[0] * xi[0] + et[1] * et[1] + q * q);
    const double r21 = std::sqrt(xi[1] * xi[1] + et[0] * et[0] + q * q);
    const double r22 = std::sqrt(xi[1] * xi[1] + et[1] * et[1] + q * q);
    bool kxi[2] = {xi[0] < 0.0 && r21 + xi[1] < EPS, xi[0] < 0.0 && r22 + xi[1] < EPS};
    bool ket[2] = {et[0] < 0.0 && r12 + et[1] < EPS, et[0] < 0.0 && r22 + et[1] < EPS};
    if (MUT(33)) kxi[0] = kxi[1] = ket[0] = ket[1] = false;
    for (int kk = 0; kk < 2; ++kk)
      for (int j = 0; j < 2; ++j) {
        double fA[3], fB[3], fC[3];
        corner(xi[j], et[kk], q, z, D, alpha, U, kxi[kk], ket[j], section == 1, fA, fB, fC);
        double du[3];
        if (section == 0) {
          du[0] = -fA[0];
          du[1] = -fA[1] * D.cd + fA[2] * D.sd;
          du[2] = -fA[1] * D.sd - fA[2] * D.cd;
        } else {
          const double zc = MUT(32) ? -z : z;
          du[0] = fA[0] + fB[0] + z * fC[0];
          du[1] = (fA[1] + fB[1] + z * fC[1]) * D.cd - (fA[2] + fB[2] + z * fC[2]) * D.sd;
          du[2] = (fA[1] + fB[1] - zc * fC[1]) * D.sd + (fA[2] + fB[2] - zc * fC[2]) * D.cd;
        }
        const double sgn = (j + kk == 1) ? -1.0 : 1.0;  // Chinnery's sign
        for (int m = 0; m < 3; ++m) u[m] += sgn * du[m];
      }
  }
  return 0;
}

// DC3D's interface: single-precision arguments and results.
static int dc3d(float ALPHA, float X, float Y, float Z, float DEPTH, float DIP, float AL1, float AL2, float AW1,
                float AW2, float DISL1, float DISL2, float DISL3, float* UX, float* UY, float* UZ) {
  const double U[3] = {DISL1, DISL2, DISL3};
  double u[3];
  const int iret = dc3d_core(ALPHA, X, Y, Z, DEPTH, DIP, AL1, AL2, AW1, AW2, U, u);
  *UX = (float)u[0]; *UY = (float)u[1]; *UZ = (float)u[2];
  return iret;
}

// okada, okada/green_space.f90:29-105: RELAX's fault (upper corner x, y, z
// at the start of the strike, north-east-down, angles in degrees, deck slip
// and rake) to DC3D's arguments and back.
static void point(double lambda, double mu, double slip, double opening, double x1, double x2, double x3,
                  double width, double length, double strike, double dipd, double rake, double xrec, double yrec,
                  double zrec, double disp[3]) {
  const double degtorad = (double)1.745329252e-2f;  // a single-precision literal (green_space.f90:47)
  const float ALPHA = (float)((lambda + mu) / (lambda + 2.0 * mu));
  const float Z = (float)(-zrec);
  const double st = strike * degtorad, csst = std::cos(st), ssst = std::sin(st);
  const double ra = rake * degtorad, csra = std::cos(ra), ssra = std::sin(ra);
  const float X = (float)((xrec - x1) * csst + (yrec - x2) * ssst);
  const float Yv = (float)((xrec - x1) * ssst - (yrec - x2) * csst);
  float UX, UY, UZ;
  const int iret = dc3d(ALPHA, X, Yv, Z, (float)x3, (float)dipd, 0.0f, (float)length, -(float)width, 0.0f,
                        (float)(slip * csra), (float)(slip * ssra), (float)opening, &UX, &UY, &UZ);
  if (iret == 0) {
    disp[0] = (double)UX * csst + (double)UY * ssst;
    disp[1] = (double)UX * ssst - (double)UY * csst;
    disp[2] = -(double)UZ;
  } else {
    disp[0] = disp[1] = disp[2] = 0.0;
  }
}
}  // namespace okada

// grnfct_okada, okada/green_space.f90:119-146: the displacement of one
// fault added to u1..u3 (sx3 planes from the surface, padded rows).
template <class R>
static void grnfct_okada(double lambda, double mu, double slip, double opening, double x, double y, double z,
                         double width, double length, double strike, double dip, double rake, int sx1, int sx2,
                         int sx3, double dx1, double dx2, double dx3, R* u1, R* u2, R* u3) {
  const long pitch = sx1 + 2;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = (double)i3 * dx3;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        double d[3];
        okada::point(lambda, mu, slip, opening, x, y, z, width, length, strike, dip, rake, shifted1(i1, sx1, dx1),
                     shifted1(i2, sx2, dx2), x3, d);
        const long o = i1 + pitch * (i2 + (long)sx2 * i3);
        u1[o] = u1[o] + (R)d[0];
        u2[o] = u2[o] + (R)d[1];
        u3[o] = u3[o] + (R)d[2];
      }
  }
}

// dislocations_disp, relax.f90:1300-1349: the Okada displacement of the
// shear patches and tensile cracks at or below the nyquist size, from the
// deck's copies of the sources (sc, tsc: upper corner, degrees, deck slip);
// the size gate reads the sources' length and width.  u is the half grid.
template <class R>
static void dislocations_disp(const orc_source* sc, int ns, const orc_source* tsc, int nt, double lambda, double mu,
                              double nyquist, int sx1, int sx2, int h3, double dx1, double dx2, double dx3, R* u1,
                              R* u2, R* u3) {
  const double dmin = nyquist * std::min(dx1, std::min(dx2, dx3));
  for (int k = 0; k < ns; ++k)
    if (dmin >= sc[k].length || dmin >= sc[k].width)
      grnfct_okada(lambda, mu, sc[k].slip, 0.0, sc[k].x, sc[k].y, sc[k].z, sc[k].width, sc[k].length, sc[k].strike,
                   sc[k].dip, sc[k].rake, sx1, sx2, h3, dx1, dx2, dx3, u1, u2, u3);
  for (int k = 0; k < nt; ++k)
    if (dmin >= tsc[k].length || dmin >= tsc[k].width)
      grnfct_okada(lambda, mu, 0.0, tsc[k].opening, tsc[k].x, tsc[k].y, tsc[k].z, tsc[k].width, tsc[k].length,
                   tsc[k].strike, tsc[k].dip, 0.0, sx1, sx2, h3, dx1, dx2, dx3, u1, u2, u3);
}

// tensoramplitude, elastic3d.f90:446-466
template <class R>
static double tensoramplitude(const R* t, long npts, double dx1, double dx2, double dx3) {
  const T6<R>* T = reinterpret_cast<const T6<R>*>(t);
  double a = 0;
  for (long o = 0; o < npts; ++o) a += tensornorm(T[o]);
  return a * (dx1 * dx2 * dx3);
}

// momentdensityshear, elastic3d.f90:2661-2750 (coseismic set-up of tau)
template <class R>
static void momentdensityshear(double mu, double slip, double x, double y, double z, double L, double W,
                               double strike, double dip, double rake, double beta, int sx1, int sx2, int sx3,
                               double dx1, double dx2, double dx3, R* sig) {
  T6<R>* S = reinterpret_cast<T6<R>*>(sig);
  const float rmu = 2.0f * (float)mu;
  const double cstrike = std::cos(strike), sstrike = std::sin(strike);
  const double cdip = std::cos(dip), sdip = std::sin(dip);
  const double cr = std::cos(rake), sr = std::sin(rake);
  const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0;
  double x2r = cstrike * x - sstrike * y;
  const double xr = cdip * x2r - sdip * z, yr = sstrike * x + cstrike * y, zr = sdip * x2r + cdip * z;
#pragma omp parallel for
  for (int i3 = 0; i3 < sx3; ++i3) {
    double x3 = (double)i3 * dx3;
    if (std::fabs(x3 - z) > Lp) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - x) > Wp) || (std::fabs(x2 - y) > Wp)) continue;
        double x2rr = cstrike * x1 - sstrike * x2;
        double x1s = cdip * x2rr - sdip * x3, x1i = cdip * x2rr + sdip * x3;
        if ((std::fabs(x1s - xr) > 7.01 * dx1) && (std::fabs(x1i - xr) > 7.01 * dx1)) continue;
        double x2s = sstrike * x1 + cstrike * x2;
        double x3s = sdip * x2rr + cdip * x3, x3i = -sdip * x2rr + cdip * x3;
        double temp1 = gauss(x1s - xr, dx1), temp2 = omega((x2s - yr) / W, beta), temp3 = omega((x3s - zr) / L, beta);
        double aperture = temp1 * temp2 * temp3;
        temp1 = gauss(x1i - xr, dx1);
        temp3 = omega((x3i + zr) / L, beta);
        aperture = aperture + temp1 * temp2 * temp3;
        double n[3] = {+cdip * cstrike * aperture, -cdip * sstrike * aperture, -sdip * aperture};
        double s[3] = {sstrike * cr + cstrike * sdip * sr, cstrike * cr - sstrike * sdip * sr, cdip * sr};
        for (int k = 0; k < 3; ++k) s[k] *= slip;
        R e[6] = {(R)(n[0] * s[0]), (R)((n[0] * s[1] + n[1] * s[0]) / 2.0), (R)((n[0] * s[2] + n[2] * s[0]) / 2.0),
                  (R)(n[1] * s[1]), (R)((n[1] * s[2] + n[2] * s[1]) / 2.0), (R)(n[2] * s[2])};
        T6<R>& o = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
        o.s11 += (R)rmu * e[0]; o.s12 += (R)rmu * e[1]; o.s13 += (R)rmu * e[2];
        o.s22 += (R)rmu * e[3]; o.s23 += (R)rmu * e[4]; o.s33 += (R)rmu * e[5];
      }
  }
}

// ---------------------------------------------------------------------------
// source, elastic3d.f90:2064-2252: the equivalent body force of a tapered
// shear dislocation in the space domain (x1 derivative of the Gaussian
// aperture, x2 and x3 derivatives of the raised-cosine tapers), with the
// surface image for near-vertical planes, and its surface traction.  L is
// the down-dip extent and W the along-strike one (the driver passes width,
// length).  f1..f3 are padded (sx1+2, sx2, sx3) fields of which the top half
// is written; t1..t3 padded planes.
// ---------------------------------------------------------------------------
static inline double gaussp(double x, double sigma) {
  return -x * std::exp(-0.5 * (x / sigma) * (x / sigma)) / std::sqrt(PI2) / (sigma * sigma * sigma);
}
static inline double omegap(double x, double beta) {
  const double ax = std::fabs(x);
  if (ax > (1.0 - 2.0 * beta) / (1.0 - beta) / 2.0 && ax < 1.0 / (1 - beta) / 2.0)
    return -std::copysign(1.0, x) * PI * (1.0 - beta) / 2.0 / beta *
           std::sin(PI * ((1.0 - beta) * ax - 0.5 + beta) / beta);
  return 0.0;
}

template <class R>
static void source(double mu, double s, double x, double y, double z, double L, double W, double strike, double dip,
                   double rake, double beta, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, R* f1,
                   R* f2, R* f3, R* t1, R* t2, R* t3) {
  const long pitch = sx1 + 2;
  const double cs = std::cos(strike), ss = std::sin(strike), cd = std::cos(dip), sd = std::sin(dip);
  const double cr = std::cos(rake), sr = std::sin(rake);
  const double scale = -mu * s;
  const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0;
  const double wl = std::max(Lp, Wp);
  double x2r = cs * x - ss * y;
  const double xr = cd * x2r - sd * z, yr = ss * x + cs * y, zr = sd * x2r + cd * z;
  // surface traction of the moment density at x3 = 0
  for (int i2 = 0; i2 < sx2; ++i2)
    for (int i1 = 0; i1 < sx1; ++i1) {
      const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2), x3 = 0.0;
      if ((std::fabs(x1 - x) > wl) || (std::fabs(x2 - y) > wl)) continue;
      const double xx = cs * x1 - ss * x2;
      const double x1s = cd * xx - sd * x3, x1i = cd * xx + sd * x3;
      if ((std::fabs(x1s - xr) > 7.01 * dx1) && (std::fabs(x1i - xr) > 7.01 * dx1)) continue;
      const double x2s = ss * x1 + cs * x2, x3s = sd * xx + cd * x3, x3i = -sd * xx + cd * x3;
      const double om2 = omega((x2s - yr) / W, beta);
      const double ap = gauss(x1s - xr, dx1) * om2 * omega((x3s - zr) / L, beta) +
                        gauss(x1i - xr, dx1) * om2 * omega((x3i + zr) / L, beta);
      const double n[3] = {cd * cs * ap, -cd * ss * ap, -sd * ap};
      const double b[3] = {mu * s * (ss * cr + cs * sd * sr), mu * s * (cs * cr - ss * sd * sr), mu * s * (cd * sr)};
      const long q = i1 + pitch * i2;
      t1[q] = t1[q] + (R)((n[0] * b[2] + n[2] * b[0]) / 2.0);
      t2[q] = t2[q] + (R)((n[1] * b[2] + n[2] * b[1]) / 2.0);
      t3[q] = t3[q] + (R)(n[2] * b[2]);
    }
  // body force in the top half
  const bool vertical = 2.01 * DEG2RAD > dip;
  for (int i3 = 0; i3 < sx3 / 2; ++i3) {
    const double x3 = (double)i3 * dx3;
    if ((std::fabs(x3 - z) > Lp) && (std::fabs(x3 + z) > Lp)) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - x) > wl) || (std::fabs(x2 - y) > wl)) continue;
        const double xx = cs * x1 - ss * x2;
        const double x1s = cd * xx - sd * x3, x1i = cd * xx + sd * x3;
        if ((std::fabs(x1s - xr) > 7.01 * dx1) && (std::fabs(x1i - xr) > 7.01 * dx1)) continue;
        const double x2s = ss * x1 + cs * x2, x3s = sd * xx + cd * x3, x3i = -sd * xx + cd * x3;
        double g = gauss(x1s - xr, dx1);
        const double om2 = omega((x2s - yr) / W, beta), om2p = omegap((x2s - yr) / W, beta);
        double om3 = omega((x3s - zr) / L, beta);
        const double sourc = scale * gaussp(x1s - xr, dx1) * om2 * om3;
        const double dblcp = scale / W * g * om2p * om3;
        const double dipcs = scale / L * g * om2 * omegap((x3s - zr) / L, beta);
        g = gauss(x1i - xr, dx1);
        om3 = omega((x3i + zr) / L, beta);
        const double image = scale * gaussp(x1i - xr, dx1) * om2 * om3;
        const double cplei = scale / W * g * om2p * om3;
        const long o = i1 + pitch * (i2 + (long)sx2 * i3);
        if (vertical && !MUT(19)) {
          f1[o] = (R)((double)f1[o] + (cr * ss * (sourc + image) + cr * cd * cs * (dblcp + cplei)));
          f2[o] = (R)((double)f2[o] + (cr * cs * (sourc + image) - cr * cd * ss * (dblcp + cplei)));
          f3[o] = (R)((double)f3[o] + (-cr * sd * (dblcp - cplei)));
        } else {
          f1[o] = (R)((double)f1[o] + (cr * ss * sourc + cr * cd * cs * dblcp));
          f2[o] = (R)((double)f2[o] + (cr * cs * sourc - cr * cd * ss * dblcp));
          f3[o] = (R)((double)f3[o] + (-cr * sd * dblcp));
        }
        // dip-slip component
        f1[o] = (R)((double)f1[o] + (cd * sr * cs * dipcs + sd * sr * cs * sourc));
        f2[o] = (R)((double)f2[o] + (-cd * sr * ss * dipcs - sd * sr * ss * sourc));
        f3[o] = (R)((double)f3[o] + (cd * sr * sourc - sd * sr * dipcs));
      }
  }
}

// ---------------------------------------------------------------------------
// traction, elastic3d.f90:2517-2590: the surface loads l of an event
// subtracted from t3 (padded plane) at time t, or their rate over [t, t+Dt]
// ---------------------------------------------------------------------------
template <class R>
static void traction(const orc_source* l, int nl, int sx1, int sx2, double dx1, double dx2, double t, double Dt,
                     R* t3, bool rate) {
  const long pitch = sx1 + 2;
  for (int k = 0; k < nl; ++k) {
    const orc_source& L = l[k];
    const double Lp = L.length * (1.0 + 2.0 * L.beta) / 2.0, Wp = L.width * (1.0 + 2.0 * L.beta) / 2.0;
    if (0 == Lp || 0 == Wp) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - L.x) > std::max(Lp, Wp)) || (std::fabs(x2 - L.y) > std::max(Lp, Wp))) continue;
        const double amp = omega((x1 - L.x) / L.length, L.beta) * omega((x2 - L.y) / L.width, L.beta) * L.slip;
        R& v = t3[i1 + pitch * i2];
        const double sg = MUT(21) ? -1.0 : 1.0;
        if (rate) {
          if (0 != L.period)
            v = (R)((double)v - sg * amp * (std::sin(2 * PI * (t + Dt) / L.period + L.phase) -
                                            std::sin(2 * PI * t / L.period + L.phase)) / Dt);
        } else if (L.period == 0) {
          v = (R)((double)v - sg * amp);
        } else {
          v = (R)((double)v - sg * amp * std::sin(2 * PI * t / L.period + L.phase));
        }
      }
  }
}

// ---------------------------------------------------------------------------
// tensilesource, elastic3d.f90:2315-2411: the body force of a tapered
// opening crack over the full depth of f (x3 from shiftedcoordinates, so the
// image above the surface lands in the lower half), one rounding per point.
// ---------------------------------------------------------------------------
template <class R>
static void tensilesource(double lambda, double mu, double s, double x, double y, double z, double L, double W,
                          double strike, double dip, double beta, int sx1, int sx2, int sx3, double dx1, double dx2,
                          double dx3, R* f1, R* f2, R* f3) {
  const long pitch = sx1 + 2;
  const double cs = std::cos(strike), ss = std::sin(strike), cd = std::cos(dip), sd = std::sin(dip);
  const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0;
  const double x2r = cs * x - ss * y;
  const double xr = cd * x2r - sd * z, yr = ss * x + cs * y, zr = sd * x2r + cd * z;
  const double scale1 = -s * (lambda + 2.0 * mu), scale2 = -s * lambda;
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = shifted1(i3, sx3, dx3);
    if ((std::fabs(x3 - z) > Lp) && (std::fabs(x3 + z) > Lp)) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - x) > Wp) || (std::fabs(x2 - y) > Wp)) continue;
        const double xx = cs * x1 - ss * x2;
        const double x1s = cd * xx - sd * x3, x1i = cd * xx + sd * x3;
        if ((std::fabs(x1s - xr) > 7.01 * dx1) && (std::fabs(x1i - xr) > 7.01 * dx1)) continue;
        const double x2s = ss * x1 + cs * x2, x3s = sd * xx + cd * x3, x3i = -sd * xx + cd * x3;
        double t1 = gauss(x1s - xr, dx1);
        const double t2 = omega((x2s - yr) / W, beta);
        double t3 = omega((x3s - zr) / L, beta);
        const double sourc = scale1 * gaussp(x1s - xr, dx1) * t2 * t3;
        const double dblcp = scale2 / W * t1 * omegap((x2s - yr) / W, beta) * t3;
        const double dipcs = scale2 / L * t1 * t2 * omegap((x3s - zr) / L, beta);
        t1 = gauss(x1i - xr, dx1);
        t3 = omega((x3i + zr) / L, beta);
        const double image = scale1 * gaussp(x1i - xr, dx1) * t2 * t3;
        const double cplei = scale2 / W * t1 * omegap((x2s - yr) / W, beta) * t3;
        const double dipci = scale2 / L * t1 * t2 * omegap((x3i + zr) / L, beta);
        const long o = i1 + pitch * (i2 + (long)sx2 * i3);
        f1[o] = (R)((double)f1[o] + cs * cd * (sourc + image) + ss * (dblcp + cplei) + cs * sd * (dipcs + dipci));
        f2[o] = (R)((double)f2[o] - ss * cd * (sourc + image) + cs * (dblcp + cplei) - ss * sd * (dipcs + dipci));
        f3[o] = (R)((double)f3[o] - sd * (sourc - image) + cd * (MUT(22) ? dipcs + dipci : dipcs - dipci));
      }
  }
}

// mogisource, elastic3d.f90:2458-2509: the body force of a point dilatation
// source (volume change o) and its image, over the full depth of f
template <class R>
static void mogisource(double lambda, double mu, double o, double xs, double ys, double zs, int sx1, int sx2, int sx3,
                       double dx1, double dx2, double dx3, R* f1, R* f2, R* f3) {
  const long pitch = sx1 + 2;
  const double scale = -(lambda + 2.0 * mu / 3.0) * o;
  const double Wp = 6.0 * std::max(dx1, std::max(dx2, dx3)), Lp = Wp;
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = shifted1(i3, sx3, dx3);
    if ((std::fabs(x3 - zs) > Lp) && (std::fabs(x3 + zs) > Lp)) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - xs) > Wp) || (std::fabs(x2 - ys) > Wp)) continue;
        const double t1 = gauss(x1 - xs, dx1), t2 = gauss(x2 - ys, dx2);
        double t3 = gauss(x3 - zs, dx3);
        const double s1 = scale * gaussp(x1 - xs, dx1) * t2 * t3, s2 = scale * t1 * gaussp(x2 - ys, dx2) * t3;
        const double s3 = scale * t1 * t2 * gaussp(x3 - zs, dx3);
        t3 = gauss(x3 + zs, dx3);
        const double m1 = scale * gaussp(x1 - xs, dx1) * t2 * t3, m2 = scale * t1 * gaussp(x2 - ys, dx2) * t3;
        const double m3 = scale * t1 * t2 * gaussp(x3 + zs, dx3);
        const long q = i1 + pitch * (i2 + (long)sx2 * i3);
        f1[q] = (R)((double)f1[q] + (s1 + m1));
        f2[q] = (R)((double)f2[q] + (s2 + m2));
        f3[q] = (R)((double)f3[q] + (MUT(24) ? s3 + m3 : s3 - m3));
      }
  }
}

// momentdensitytensile, elastic3d.f90:3137-3215: Hooke's law of the
// eigenstrain n sdyad (slip n) of an opening crack (n carries the aperture,
// so the eigenstrain goes with its square), added to sig (sx3 = the depth
// of sig, x3 unshifted)
template <class R>
static void momentdensitytensile(double lambda, double mu, double slip, double x, double y, double z, double L,
                                 double W, double strike, double dip, double beta, int sx1, int sx2, int sx3,
                                 double dx1, double dx2, double dx3, R* sig) {
  T6<R>* S = reinterpret_cast<T6<R>*>(sig);
  const double cs = std::cos(strike), ss = std::sin(strike), cd = std::cos(dip), sd = std::sin(dip);
  const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0;
  const double x2r = cs * x - ss * y;
  const double xr = cd * x2r - sd * z, yr = ss * x + cs * y, zr = sd * x2r + cd * z;
  const R m2 = (R)(2.0 * mu);
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = (double)i3 * dx3;
    if (std::fabs(x3 - z) > Lp) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - x) > Wp) || (std::fabs(x2 - y) > Wp)) continue;
        const double xx = cs * x1 - ss * x2;
        const double x1s = cd * xx - sd * x3, x1i = cd * xx + sd * x3;
        if ((std::fabs(x1s - xr) > 7.01 * dx1) && (std::fabs(x1i - xr) > 7.01 * dx1)) continue;
        const double x2s = ss * x1 + cs * x2, x3s = sd * xx + cd * x3, x3i = -sd * xx + cd * x3;
        const double t2 = omega((x2s - yr) / W, beta);
        double ap = gauss(x1s - xr, dx1) * t2 * omega((x3s - zr) / L, beta);
        ap = ap + gauss(x1i - xr, dx1) * t2 * omega((x3i + zr) / L, beta);
        const double n[3] = {+cd * cs * ap, -cd * ss * ap, -sd * ap};
        const double b[3] = {slip * n[0], slip * n[1], slip * n[2]};
        R e[6] = {(R)(n[0] * b[0]), (R)((n[0] * b[1] + n[1] * b[0]) / 2.0), (R)((n[0] * b[2] + n[2] * b[0]) / 2.0),
                  (R)(n[1] * b[1]), (R)((n[1] * b[2] + n[2] * b[1]) / 2.0), (R)(n[2] * b[2])};
        // isotropicstressstrain, elastic3d.f90:129-142
        const double tr = (double)(e[0] + e[3] + e[5]);
        for (R& v : e) v = m2 * v;
        e[0] = (R)((double)e[0] + lambda * tr);
        e[3] = (R)((double)e[3] + lambda * tr);
        e[5] = (R)((double)e[5] + (MUT(23) ? 0.0 : lambda * tr));
        T6<R>& q = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
        q.s11 += e[0]; q.s12 += e[1]; q.s13 += e[2]; q.s22 += e[3]; q.s23 += e[4]; q.s33 += e[5];
      }
  }
}

// momentdensitymogi, elastic3d.f90:3240-3286: the isotropic moment density
// kappa o G of a point dilatation source (without its image), x3 unshifted
template <class R>
static void momentdensitymogi(double lambda, double mu, double o, double xs, double ys, double zs, int sx1, int sx2,
                              int sx3, double dx1, double dx2, double dx3, R* sig) {
  T6<R>* S = reinterpret_cast<T6<R>*>(sig);
  const double kappa = lambda + 2.0 / 3.0 * mu;
  const double Wp = 6.0 * std::max(dx1, std::max(dx2, dx3)), Lp = Wp;
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = (double)i3 * dx3;
    if (std::fabs(x3 - zs) > Lp) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - xs) > Wp) || (std::fabs(x2 - ys) > Wp)) continue;
        double g = o * gauss(x1 - xs, dx1) * gauss(x2 - ys, dx2) * gauss(x3 - zs, dx3);
        g = kappa * g;
        const R m = (R)g;
        T6<R>& q = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
        q.s11 += m; q.s22 += m; q.s33 += m;
      }
  }
}

// ---------------------------------------------------------------------------
// Distributed eigenstrain (--with-eigenstrain): cuboids, elastic3d.f90
// cuboidSource :1680-1884 and momentDensityCuboid :2927-3029, and
// tetrahedra (--curvilinear), tetrahedronSource :1916-2039 and
// momentDensityTetrahedron :2788-2873.  A cuboid is a WEAK_STRUCT (its
// eigenstrain in e); the driver passes (width, length) as the routines'
// (L, W) (relax.f90:1153-1162, :1192-1198).  TENSOR arithmetic (the
// operators .times., .plus., .sdyad., elastic3d.f90:186-255) rounds to the
// storage type R after each operation, as REAL*4 does in the reference.
// ---------------------------------------------------------------------------
extern "C" {
struct orc_tetra {
  float e[6];
  int i1, i2, i3, i4;  // 1-based vertex indices
};
}

template <class R>
static T6<R> sdyad(const double a[3], const double b[3]) {
  return T6<R>{(R)(a[0] * b[0]), (R)((a[0] * b[1] + a[1] * b[0]) / 2.0), (R)((a[0] * b[2] + a[2] * b[0]) / 2.0),
               (R)(a[1] * b[1]), (R)((a[1] * b[2] + a[2] * b[1]) / 2.0), (R)(a[2] * b[2])};
}
template <class R>
static T6<R> tscale(R c, const T6<R>& t) {
  return T6<R>{(R)(c * t.s11), (R)(c * t.s12), (R)(c * t.s13), (R)(c * t.s22), (R)(c * t.s23), (R)(c * t.s33)};
}
template <class R>
static T6<R> tplus(const T6<R>& a, const T6<R>& b) {
  return T6<R>{(R)(a.s11 + b.s11), (R)(a.s12 + b.s12), (R)(a.s13 + b.s13),
               (R)(a.s22 + b.s22), (R)(a.s23 + b.s23), (R)(a.s33 + b.s33)};
}
// the cull of both cuboid routines: SQRT(2.) is the single-precision root
static const double SQRT2_SINGLE = (double)std::sqrt(2.0f);

template <class R>
static void cuboidsource(double lambda, double mu, const float* e, double x, double y, double z, double L, double W,
                         double T, double strike, double dip, double beta, int sx1, int sx2, int sx3, double dx1,
                         double dx2, double dx3, R* f1, R* f2, R* f3) {
  const long pitch = sx1 + 2;
  const double e11 = e[0], e12 = e[1], e13 = e[2], e22 = e[3], e23 = e[4], e33 = e[5];
  const double ekk = e11 + e22 + e33;
  const double cs = std::cos(strike), ss = std::sin(strike), cd = std::cos(dip), sd = std::sin(dip);
  const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0, Tp = T * (1.0 + 2.0 * beta) / 2.0;
  const double x2r0 = cs * x - ss * y;
  const double xr = cd * x2r0 - sd * z, yr = ss * x + cs * y, zr = sd * x2r0 + cd * z;
  const double lim = SQRT2_SINGLE * std::max(Wp, std::max(Lp, Tp));
  // the surface-traction loop of cuboidSource (:1743-1775) writes nothing
  const double le11 = lambda * ekk + 2.0 * mu * e11, le22 = lambda * ekk + 2.0 * mu * e22,
               le33 = lambda * ekk + 2.0 * mu * e33;
  const double m12 = 2.0 * mu * e12, m13 = 2.0 * mu * e13, m23 = 2.0 * mu * e23;
  for (int i3 = 0; i3 < sx3 / 2; ++i3) {
    const double x3 = (double)i3 * dx3;
    if (std::fabs(x3 - z) > lim) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - x) > lim) || (std::fabs(x2 - y) > lim) || (std::fabs(x3 - z) > lim)) continue;
        const double x2r = cs * x1 - ss * x2;
        const double x1s = cd * x2r - sd * x3, x1i = cd * x2r + sd * x3;
        const double x2s = ss * x1 + cs * x2, x3s = sd * x2r + cd * x3, x3i = -sd * x2r + cd * x3;
        double t1 = omega((x1s - xr) / T, beta);
        const double t2 = omega((x2s - yr) / W, beta);
        double t3 = omega((x3s - zr) / L, beta);
        const double sourc = 1.0 / T * omegap((x1s - xr) / T, beta) * t2 * t3;
        const double dblcp = 1.0 / W * t1 * omegap((x2s - yr) / W, beta) * t3;
        const double dipcs = 1.0 / L * t1 * t2 * omegap((x3s - zr) / L, beta);
        t1 = omega((x1i - xr) / T, beta);
        t3 = omega((x3i + zr) / L, beta);
        const double image = MUT(35) ? 0.0 : 1.0 / T * omegap((x1i - xr) / T, beta) * t2 * t3;
        const double cplei = 1.0 / W * t1 * omegap((x2s - yr) / W, beta) * t3;
        const long o = i1 + pitch * (i2 + (long)sx2 * i3);
        if (2.01 * DEG2RAD > dip) {
          // subvertical: method of images
          f1[o] = (R)((double)f1[o] + m12 * ss * (sourc + image) + m12 * cd * cs * (dblcp + cplei) -
                      le11 * ss * (dblcp + cplei) - le22 * cd * cs * (sourc + image));
          f2[o] = (R)((double)f2[o] + m12 * cs * (sourc + image) - m12 * cd * ss * (dblcp + cplei) -
                      le11 * cs * (dblcp + cplei) + le22 * cd * ss * (sourc + image));
          f3[o] = (R)((double)f3[o] - m12 * sd * (dblcp - cplei) + le22 * sd * (sourc - image));
        } else {
          f1[o] = (R)((double)f1[o] + m12 * ss * sourc + m12 * cd * cs * dblcp - le11 * ss * dblcp -
                      le22 * cd * cs * sourc);
          f2[o] = (R)((double)f2[o] + m12 * cs * sourc - m12 * cd * ss * dblcp - le11 * cs * dblcp +
                      le22 * cd * ss * sourc);
          f3[o] = (R)((double)f3[o] - m12 * sd * dblcp + le22 * sd * sourc);
        }
        // vertical shear components
        f1[o] = (R)((double)f1[o] + m13 * cd * ss * dipcs + m13 * sd * cs * dblcp - m23 * cd * cs * dipcs -
                    m23 * sd * cs * sourc - le33 * sd * dipcs * cs);
        f2[o] = (R)((double)f2[o] + m13 * cd * cs * dipcs - m13 * sd * ss * dblcp + m23 * cd * ss * dipcs +
                    m23 * sd * ss * sourc + le33 * sd * dipcs * ss);
        f3[o] = (R)((double)f3[o] + m13 * cd * dblcp - m23 * cd * sourc - le33 * cd * dipcs);
      }
  }
}

template <class R>
static void momentdensitycuboid(double mu, double lambda, const float* e, double x, double y, double z, double L,
                                double W, double T, double strike, double dip, double beta, int sx1, int sx2,
                                int sx3, double dx1, double dx2, double dx3, R* sig) {
  T6<R>* S = reinterpret_cast<T6<R>*>(sig);
  const T6<R> E{(R)e[0], (R)e[1], (R)e[2], (R)e[3], (R)e[4], (R)e[5]};
  const R lam = (R)lambda, rmu = (R)2 * (R)mu;
  const double cs = std::cos(strike), ss = std::sin(strike), cd = std::cos(dip), sd = std::sin(dip);
  const double Wp = W * (1.0 + 2.0 * beta) / 2.0, Lp = L * (1.0 + 2.0 * beta) / 2.0, Tp = T * (1.0 + 2.0 * beta) / 2.0;
  const double x2r0 = cs * x - ss * y;
  const double xr = cd * x2r0 - sd * z, yr = ss * x + cs * y, zr = sd * x2r0 + cd * z;
  const double lim = SQRT2_SINGLE * std::max(Wp, std::max(Lp, Tp));
  // normal, strike and dip directions
  const double n[3] = {cd * cs, -cd * ss, -sd}, sv[3] = {-ss, MUT(36) ? cs : -cs, 0.0}, d[3] = {cs * sd, -ss * sd, cd};
  const T6<R> SS = sdyad<R>(sv, sv), NS = sdyad<R>(n, sv), SD = sdyad<R>(sv, d), NN = sdyad<R>(n, n),
              ND = sdyad<R>(n, d), DD = sdyad<R>(d, d);
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = (double)i3 * dx3;
    if (std::fabs(x3 - z) > lim) continue;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x1 = shifted1(i1, sx1, dx1), x2 = shifted1(i2, sx2, dx2);
        if ((std::fabs(x1 - x) > lim) || (std::fabs(x2 - y) > lim) || (std::fabs(x3 - z) > lim)) continue;
        const double x2r = cs * x1 - ss * x2;
        const double x1s = cd * x2r - sd * x3, x2s = ss * x1 + cs * x2, x3s = sd * x2r + cd * x3;
        // the image aperture is formed but not added (:2993-2995)
        const R ap = (R)(omega((x1s - xr) / T, beta) * omega((x2s - yr) / W, beta) * omega((x3s - zr) / L, beta));
        const T6<R> ea = tscale(ap, E);
        T6<R> Ei = tscale(ea.s11, SS);
        Ei = tplus(Ei, tscale((R)((R)2 * ea.s12), NS));
        Ei = tplus(Ei, tscale((R)((R)2 * ea.s13), SD));
        Ei = tplus(Ei, tscale(ea.s22, NN));
        Ei = tplus(Ei, tscale((R)((R)2 * ea.s23), ND));
        Ei = tplus(Ei, tscale(ea.s33, DD));
        const R ekk = (R)((double)Ei.s11 + (double)Ei.s22 + (double)Ei.s33);
        T6<R>& q = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
        const T6<R> lk{(R)(ekk * lam), (R)0, (R)0, (R)(ekk * lam), (R)0, (R)(ekk * lam)};
        q = tplus(tplus(q, lk), tscale(rmu, Ei));
      }
  }
}

// the tetrahedron's outward unit normals and face centroids; false for a
// tetrahedron without volume (:1952-1983)
static bool tetra_faces(const double* A, const double* B, const double* C, const double* D, double n[4][3],
                        double c[4][3]) {
  const double* P[4] = {A, B, C, D};
  // face opposite vertex k: (B,C,D), (C,D,A), (D,A,B), (A,B,C)
  for (int k = 0; k < 4; ++k) {
    const double* p = P[(k + 1) % 4];
    const double* q = P[(k + 2) % 4];
    const double* r = P[(k + 3) % 4];
    // nA=(C-B)x(D-B), nB=(D-C)x(A-C), nC=(A-D)x(B-D), nD=(B-A)x(C-A)
    const double u[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]}, v[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]};
    n[k][0] = u[1] * v[2] - u[2] * v[1];
    n[k][1] = u[2] * v[0] - u[0] * v[2];
    n[k][2] = u[0] * v[1] - u[1] * v[0];
    for (int j = 0; j < 3; ++j) c[k][j] = (p[j] + q[j] + r[j]) / 3.0;
  }
  for (int k = 0; k < 4; ++k)
    if (0.0 == std::sqrt(n[k][0] * n[k][0] + n[k][1] * n[k][1] + n[k][2] * n[k][2])) return false;
  for (int k = 0; k < 4; ++k) {
    const double a = std::sqrt(n[k][0] * n[k][0] + n[k][1] * n[k][1] + n[k][2] * n[k][2]);
    for (int j = 0; j < 3; ++j) n[k][j] = n[k][j] / a;
  }
  for (int k = 0; k < 4; ++k) {
    // outward: (vertex - centroid of the opposite face) . n <= 0
    const double* w = P[k];
    const double dot = n[k][0] * (w[0] - c[k][0]) + n[k][1] * (w[1] - c[k][1]) + n[k][2] * (w[2] - c[k][2]);
    if (MUT(34) ? dot < 0 : dot > 0)
      for (int j = 0; j < 3; ++j) n[k][j] = -n[k][j];
  }
  return true;
}
static inline double gaussi(double x, double sigma) { return (1.0 + std::erf(x / std::sqrt(2.0) / sigma)) / 2.0; }

template <class R>
static void tetrahedronsource(double lambda, double mu, const float* e, const double* A, const double* B,
                              const double* C, const double* D, int sx1, int sx2, int sx3, double dx1, double dx2,
                              double dx3, R* f1, R* f2, R* f3) {
  const long pitch = sx1 + 2;
  // m = C:e (isotropicstressstrain, :129-142)
  const double epskk = (double)e[0] + (double)e[3] + (double)e[5];
  const R twomu = (R)(2.0 * mu);
  T6<R> m{(R)(twomu * e[0]), (R)(twomu * e[1]), (R)(twomu * e[2]), (R)(twomu * e[3]), (R)(twomu * e[4]),
          (R)(twomu * e[5])};
  m.s11 = (R)(m.s11 + lambda * epskk);
  m.s22 = (R)(m.s22 + lambda * epskk);
  m.s33 = (R)(m.s33 + lambda * epskk);
  double n[4][3], c[4][3];
  if (!tetra_faces(A, B, C, D, n, c)) return;
  double tr[4][3];
  for (int k = 0; k < 4; ++k) {
    tr[k][0] = m.s11 * n[k][0] + m.s12 * n[k][1] + m.s13 * n[k][2];
    tr[k][1] = m.s12 * n[k][0] + m.s22 * n[k][1] + m.s23 * n[k][2];
    tr[k][2] = m.s13 * n[k][0] + m.s23 * n[k][1] + m.s33 * n[k][2];
  }
  for (int i3 = 0; i3 < sx3 / 2; ++i3) {
    const double x3 = (double)i3 * dx3;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x[3] = {shifted1(i1, sx1, dx1), shifted1(i2, sx2, dx2), x3};
        double dist[4];
        bool far = true;
        for (int k = 0; k < 4; ++k) {
          dist[k] = (c[k][0] - x[0]) * n[k][0] + (c[k][1] - x[1]) * n[k][1] + (c[k][2] - x[2]) * n[k][2];
          far = far && std::fabs(dist[k]) > 7.0 * dx1;
        }
        if (far) continue;
        double H[4], dl[4];
        for (int k = 0; k < 4; ++k) {
          H[k] = gaussi(dist[k], dx1);
          dl[k] = gauss(dist[k], dx1);
        }
        const double w[4] = {dl[0] * H[1] * H[2] * H[3], H[0] * dl[1] * H[2] * H[3], H[0] * H[1] * dl[2] * H[3],
                             H[0] * H[1] * H[2] * dl[3]};
        const long o = i1 + pitch * (i2 + (long)sx2 * i3);
        R* f[3] = {f1, f2, f3};
        for (int j = 0; j < 3; ++j)
          f[j][o] = (R)((double)f[j][o] + tr[0][j] * w[0] + tr[1][j] * w[1] + tr[2][j] * w[2] + tr[3][j] * w[3]);
      }
  }
}

template <class R>
static void momentdensitytetrahedron(double mu, double lambda, const float* e, const double* A, const double* B,
                                     const double* C, const double* D, int sx1, int sx2, int sx3, double dx1,
                                     double dx2, double dx3, R* sig) {
  T6<R>* S = reinterpret_cast<T6<R>*>(sig);
  const T6<R> E{(R)e[0], (R)e[1], (R)e[2], (R)e[3], (R)e[4], (R)e[5]};
  const R lam = (R)lambda, rmu = (R)2 * (R)mu;
  double n[4][3], c[4][3];
  if (!tetra_faces(A, B, C, D, n, c)) return;
  for (int i3 = 0; i3 < sx3; ++i3) {
    const double x3 = (double)i3 * dx3;
    for (int i2 = 0; i2 < sx2; ++i2)
      for (int i1 = 0; i1 < sx1; ++i1) {
        const double x[3] = {shifted1(i1, sx1, dx1), shifted1(i2, sx2, dx2), x3};
        double dist[4];
        bool near = true;
        for (int k = 0; k < 4; ++k) {
          dist[k] = (c[k][0] - x[0]) * n[k][0] + (c[k][1] - x[1]) * n[k][1] + (c[k][2] - x[2]) * n[k][2];
          near = near && dist[k] < 7.0 * dx1;
        }
        // the reference skips the points within 7 dx1 of the inner side of
        // every face (:2838-2841), as written
        if (near) continue;
        const R h = (R)(gaussi(dist[0], dx1) * gaussi(dist[1], dx1) * gaussi(dist[2], dx1) * gaussi(dist[3], dx1));
        const T6<R> ea = tscale(h, E);
        const R ekk = (R)((double)ea.s11 + (double)ea.s22 + (double)ea.s33);
        T6<R>& q = S[i1 + (long)sx1 * (i2 + (long)sx2 * i3)];
        const T6<R> lk{(R)(ekk * lam), (R)0, (R)0, (R)(ekk * lam), (R)0, (R)(ekk * lam)};
        q = tplus(tplus(q, lk), tscale(rmu, ea));
      }
  }
}

// the eigenstrain sources of an event: body forces then moment densities
// (relax.f90:1151-1178, :1192-1223)
// (body forces into v1..v3 when given, moment densities into tau when given)
template <class R>
static void eigenstrain(const orc_weak* cb, int nc, const orc_tetra* th, int nth, const double* vtx, int nv,
                        double lambda, double mu, double beta, int sx1, int sx2, int sx3, double dx1, double dx2,
                        double dx3, R* v1, R* v2, R* v3, R* tau) {
  auto V = [&](int i) { return vtx + 3 * (long)(i - 1); };
  auto ok = [&](const orc_tetra& t) {
    return t.i1 >= 1 && t.i1 <= nv && t.i2 >= 1 && t.i2 <= nv && t.i3 >= 1 && t.i3 <= nv && t.i4 >= 1 && t.i4 <= nv;
  };
  if (v1) {
    for (int k = 0; k < nc; ++k)
      cuboidsource(lambda, mu, cb[k].e, cb[k].x, cb[k].y, cb[k].z, cb[k].width, cb[k].length, cb[k].thickness,
                   cb[k].strike, cb[k].dip, beta, sx1, sx2, sx3, dx1, dx2, dx3, v1, v2, v3);
    for (int k = 0; k < nth; ++k)
      if (ok(th[k]))
        tetrahedronsource(lambda, mu, th[k].e, V(th[k].i1), V(th[k].i2), V(th[k].i3), V(th[k].i4), sx1, sx2, sx3,
                          dx1, dx2, dx3, v1, v2, v3);
  }
  if (tau) {
    for (int k = 0; k < nc; ++k)
      momentdensitycuboid(mu, lambda, cb[k].e, cb[k].x, cb[k].y, cb[k].z, cb[k].width, cb[k].length,
                          cb[k].thickness, cb[k].strike, cb[k].dip, beta, sx1, sx2, sx3 / 2, dx1, dx2, dx3, tau);
    for (int k = 0; k < nth; ++k)
      if (ok(th[k]))
        momentdensitytetrahedron(mu, lambda, th[k].e, V(th[k].i1), V(th[k].i2), V(th[k].i3), V(th[k].i4), sx1, sx2,
                                 sx3 / 2, dx1, dx2, dx3, tau);
  }
}

// ---------------------------------------------------------------------------
// dislocations, relax.f90:1116-1289 (shear patches), then traction
// (relax.f90:319, :874): the source() body force of the patches longer or
// wider than nyquist*min(dx) into v (padded, sx3 deep) and t, the moment
// density of every patch into tau, and the surface loads into t3.
// ---------------------------------------------------------------------------
template <class R>
static void dislocations(const orc_source* s, int ns, const orc_source* ts, int nt, const orc_source* mg, int nm,
                         const orc_source* l, int nl, double lambda, double mu, double beta, double nyquist,
                         double t, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, R* v1, R* v2,
                         R* v3, R* t1, R* t2, R* t3, R* tau, const orc_weak* cb = nullptr, int nc = 0,
                         const orc_tetra* th = nullptr, int nth = 0, const double* vtx = nullptr, int nv = 0) {
  const double dmin = nyquist * std::min(dx1, std::min(dx2, dx3));
  for (int k = 0; k < ns; ++k) {
    const bool big = dmin < s[k].length || dmin < s[k].width;
    if (MUT(20) ? !big : big)
      source(mu, s[k].slip, s[k].x, s[k].y, s[k].z, s[k].width, s[k].length, s[k].strike, s[k].dip, s[k].rake,
             s[k].beta, sx1, sx2, sx3, dx1, dx2, dx3, v1, v2, v3, t1, t2, t3);
  }
  // the eigenstrain volumes' body forces follow the shear patches'
  // (:1151-1178); their moment densities follow the patches' (:1192-1223)
  eigenstrain(cb, nc, th, nth, vtx, nv, lambda, mu, beta, sx1, sx2, sx3, dx1, dx2, dx3, v1, v2, v3, (R*)nullptr);
  for (int k = 0; k < ns; ++k)
    momentdensityshear(mu, s[k].slip, s[k].x, s[k].y, s[k].z, s[k].width, s[k].length, s[k].strike, s[k].dip,
                       s[k].rake, s[k].beta, sx1, sx2, sx3 / 2, dx1, dx2, dx3, tau);
  eigenstrain(cb, nc, th, nth, vtx, nv, lambda, mu, beta, sx1, sx2, sx3, dx1, dx2, dx3, (R*)nullptr, (R*)nullptr,
              (R*)nullptr, tau);
  // tensile cracks (:1225-1254): the body force of the large ones with the
  // model's taper, every crack's moment density
  for (int k = 0; k < nt; ++k)
    if (dmin < ts[k].length || dmin < ts[k].width)
      tensilesource(lambda, mu, ts[k].opening, ts[k].x, ts[k].y, ts[k].z, ts[k].width, ts[k].length, ts[k].strike,
                    ts[k].dip, beta, sx1, sx2, sx3, dx1, dx2, dx3, v1, v2, v3);
  for (int k = 0; k < nt; ++k)
    momentdensitytensile(lambda, mu, ts[k].opening, ts[k].x, ts[k].y, ts[k].z, ts[k].width, ts[k].length,
                         ts[k].strike, ts[k].dip, beta, sx1, sx2, sx3 / 2, dx1, dx2, dx3, tau);
  // point dilatation sources (:1256-1287): slip is the volume change
  for (int k = 0; k < nm; ++k)
    mogisource(lambda, mu, mg[k].slip, mg[k].x, mg[k].y, mg[k].z, sx1, sx2, sx3, dx1, dx2, dx3, v1, v2, v3);
  for (int k = 0; k < nm; ++k)
    momentdensitymogi(lambda, mu, mg[k].slip, mg[k].x, mg[k].y, mg[k].z, sx1, sx2, sx3 / 2, dx1, dx2, dx3, tau);
  traction(l, nl, sx1, sx2, dx1, dx2, t, 0.0, t3, false);
}

// ---------------------------------------------------------------------------
// RK2 loop: relax.f90:555-890 with integrationstep :1420-1452 and the
// isoutput counter :1364-1376 (oi advances whenever isoutput holds).
// Mechanisms: linear and power-law viscosity (with weak zones), fault creep,
// and linear and power-law transient creep (with weak zones); later events
// through the event switch (:858-889, dislocations and traction); the rate of
// the current event's time-dependent surface loads in both stages (:684-685,
// :805-806); no interseismic loading.
// ---------------------------------------------------------------------------
extern "C" {
struct orc_model {
  int sx1, sx2, sx3;
  double dx1, dx2, dx3, lambda, mu, gam, beta;
  const orc_layer* linear;     // sx3/2 entries or NULL
  const orc_layer* nonlinear;  // sx3/2 entries or NULL
  const orc_layer* faultcreep; // sx3/2 entries or NULL
  const orc_tensor_layer* prestress;  // sx3/2 entries
  int np;
  const orc_plane* planes;
  orc_patch* patches;  // concatenated px2*px3 per plane
  double interval, odt, tscale;
  int skip;
  int ne;
  const double* event_times;
  // transient creep (relax.f90:614-650, :764-800): linear and power-law
  // Burgers structures, sx3/2 entries or NULL; either one needs the
  // epsilonik state of rk2_run
  const orc_layer* ltransient;
  const orc_layer* nltransient;
  // weak zones (builddgammadot0 fields, sx1*sx2*sx3/2 values of the storage
  // type, or NULL): linear, power-law, linear transient, power-law transient
  const void* dg0[4];
  // the later events (NULL: none): event k (1..ne-1) owns the shear patches
  // ev_src[ev_off[k] .. ev_off[k+1]) and the surface loads
  // ev_loads[ev_loff[k] .. ev_loff[k+1]); dislocations' nyquist.  Entry 0
  // of the loads is the first event's, whose rate the steps before the
  // second event add.
  const orc_source* ev_src;
  const int* ev_off;
  const orc_source* ev_loads;
  const int* ev_loff;
  double nyquist;
  // tensile cracks and point dilatation sources of the later events, listed
  // as the shear patches
  const orc_source* ev_ts;
  const int* ev_tsoff;
  const orc_source* ev_mg;
  const int* ev_mgoff;
  // inter-seismic loading (in%inter): shear patches and tensile cracks whose
  // slip or opening is a rate (input.f90:1114-1250); NULL/0: none
  const orc_source* inter_s;
  int inter_ns;
  const orc_source* inter_ts;
  int inter_nt;
  // eigenstrain volumes of the later events, listed as the shear patches:
  // cuboids, tetrahedra and their vertices (rows of x1, x2, x3)
  const orc_weak* ev_cb;
  const int* ev_cboff;
  const orc_tetra* ev_th;
  const int* ev_thoff;
  const double* ev_vx;
  const int* ev_vxoff;
};
}

static const double STEP_MAX = 1e7;

// dislocations(in%inter, ..., tau, eigenstress=moment), relax.f90:795-799
// with :1116-1289 in its eigenstress form: every loading source's moment
// density into the corrector's moment, then (the loops that run in either
// form, :1199-1206, :1247-1254) into tau as well, unscaled by the time step.
template <class R>
static void interseismic(const orc_model* m, int sx1, int sx2, int h3, R* moment, R* tau) {
  const orc_source* s = m->inter_s;
  const orc_source* ts = m->inter_ts;
  for (int k = 0; k < m->inter_ns; ++k)
    momentdensityshear(m->mu, s[k].slip, s[k].x, s[k].y, s[k].z, s[k].width, s[k].length, s[k].strike, s[k].dip,
                       s[k].rake, s[k].beta, sx1, sx2, h3, m->dx1, m->dx2, m->dx3, moment);
  if (!MUT(31))
    for (int k = 0; k < m->inter_ns; ++k)
      momentdensityshear(m->mu, s[k].slip, s[k].x, s[k].y, s[k].z, s[k].width, s[k].length, s[k].strike,
                         s[k].dip, s[k].rake, s[k].beta, sx1, sx2, h3, m->dx1, m->dx2, m->dx3, tau);
  for (int k = 0; k < m->inter_nt; ++k)
    momentdensitytensile(m->lambda, m->mu, ts[k].opening, ts[k].x, ts[k].y, ts[k].z, ts[k].width, ts[k].length,
                         ts[k].strike, ts[k].dip, m->beta, sx1, sx2, h3, m->dx1, m->dx2, m->dx3, moment);
  if (!MUT(31))
    for (int k = 0; k < m->inter_nt; ++k)
      momentdensitytensile(m->lambda, m->mu, ts[k].opening, ts[k].x, ts[k].y, ts[k].z, ts[k].width, ts[k].length,
                           ts[k].strike, ts[k].dip, m->beta, sx1, sx2, h3, m->dx1, m->dx2, m->dx3, tau);
}

template <class R>
static int rk2_run(const orc_model* m, R* u1, R* u2, R* u3, R* tau, R* sig, R* gamma, R* epsilonik, double* t_io,
                   int* oi_io, int* i_io, int* e_io, int nsteps, const double* dt_replay, double* info) {
  const int sx1 = m->sx1, sx2 = m->sx2, sx3 = m->sx3, h3 = sx3 / 2;
  const long pitch = sx1 + 2;
  const long nv = pitch * sx2 * sx3, nu = pitch * sx2 * h3, nt = pitch * sx2, npts = (long)sx1 * sx2 * h3;
  std::vector<R> v1(nv), v2(nv), v3(nv), t1(nt), t2(nt), t3(nt), moment(6 * npts, 0);
  const bool transient = m->ltransient || m->nltransient;
  if (transient && !epsilonik) return 10;
  std::vector<R> epsilonikdot(transient ? 6 * npts : 0, (R)0);
  const R* dg[4];
  for (int k = 0; k < 4; ++k) dg[k] = static_cast<const R*>(m->dg0[k]);
  const orc_layer* trans[2] = {m->ltransient, m->nltransient};
  double t = *t_io;
  int oi = *oi_io, it = *i_io;
  int e = *e_io;  // current event, 1-based as in relax.f90 (1: the set-up at t = 0)
  for (int step = 0; step < nsteps; ++step) {
    if (t >= m->interval) break;
    ++it;
    double maxwell[5] = {STEP_MAX, STEP_MAX, STEP_MAX, STEP_MAX, STEP_MAX};
    double mech[5] = {0, 0, 0, 0, 0};
    std::fill(moment.begin(), moment.end(), (R)0);
    int rc;
    if (m->linear) {
      rc = viscouseigenstress(m->mu, m->linear, m->prestress, sig, sx1, sx2, h3, moment.data(), &maxwell[0], dg[0], (R*)0, pitch);
      if (rc) return rc;
      mech[0] = 1;
    }
    if (m->nonlinear) {
      rc = viscouseigenstress(m->mu, m->nonlinear, m->prestress, sig, sx1, sx2, h3, moment.data(), &maxwell[1], dg[1], (R*)0, pitch);
      if (rc) return rc;
      mech[1] = 1;
    }
    if (m->faultcreep) {
      for (int k = 0; k < m->np; ++k) {
        const orc_plane& P = m->planes[k];
        rc = frictioneigenstress(P.x, P.y, P.z, P.width, P.length, P.strike, P.dip, P.rake, m->beta, sig, m->mu,
                                 m->faultcreep, sx1, sx2, h3, m->dx1, m->dx2, m->dx3, moment.data(), &maxwell[2],
                                 (R*)0, pitch);
        if (rc) return rc;
      }
      mech[2] = 1;
    }
    // transient creep reads epsilonik and writes epsilonikdot
    for (int k = 0; k < 2; ++k) {
      if (!trans[k]) continue;
      rc = transienteigenstress(m->mu, trans[k], sig, epsilonik, sx1, sx2, h3, moment.data(), epsilonikdot.data(),
                                &maxwell[3 + k], dg[2 + k]);
      if (rc) return rc;
      mech[3 + k] = 1;
    }
    double tm = 1.0 / (mech[0] / maxwell[0] + mech[1] / maxwell[1] + mech[2] / maxwell[2] + mech[3] / maxwell[3] +
                       mech[4] / maxwell[4]);
    tm = std::min(tm, STEP_MAX);
    // with inter-seismic loading and nothing relaxing, a small step
    // (relax.f90:669-674)
    const bool inter = m->inter_ns > 0 || m->inter_nt > 0;
    if (inter && tm == STEP_MAX) tm = m->interval / 20.0;
    // integrationstep, relax.f90:1420-1452
    double Dt = tm / 10.0;
    if (0 == m->skip) {
      if ((t + Dt) >= ((double)oi * m->odt) - Dt * 0.04) Dt = (double)oi * m->odt - t;
    } else {
      Dt = Dt * m->tscale;
      if (e < m->ne) {
        if ((t + Dt - m->event_times[e]) >= 0.0) Dt = m->event_times[e] - t;
      }
    }
    if (dt_replay) Dt = dt_replay[step];
    // sig <- moment; v = 0; t = 0; equivalentbodyforce; green
    for (long o = 0; o < 6 * npts; ++o) sig[o] = moment[o];
    std::fill(v1.begin(), v1.end(), (R)0); std::fill(v2.begin(), v2.end(), (R)0); std::fill(v3.begin(), v3.end(), (R)0);
    std::fill(t1.begin(), t1.end(), (R)0); std::fill(t2.begin(), t2.end(), (R)0); std::fill(t3.begin(), t3.end(), (R)0);
    rc = equivalentbodyforce(sig, m->dx1, m->dx2, m->dx3, sx1, sx2, h3, v1.data(), v2.data(), v3.data(), t1.data(),
                             t2.data(), t3.data(), pitch, (const R*)0);
    if (rc) return rc;
    // the rate of the current event's time-dependent surface loads
    // (relax.f90:684-685; the reference's step there is Dt/2e8)
    if (m->ev_loff && m->ev_loff[e] > m->ev_loff[e - 1])
      traction(m->ev_loads + m->ev_loff[e - 1], m->ev_loff[e] - m->ev_loff[e - 1], sx1, sx2, m->dx1, m->dx2, t,
               Dt / 2.e8, t3.data(), true);
    rc = greenfunctioncowling(v1.data(), v2.data(), v3.data(), t1.data(), t2.data(), t3.data(), sx1, sx2, sx3, m->dx1,
                              m->dx2, m->dx3, m->lambda, m->mu, m->gam);
    if (rc) return rc;
    const R hdt = (R)(MUT(13) ? Dt : Dt / 2);
    for (long o = 0; o < nu; ++o) {
      v1[o] = hdt * v1[o] + u1[o];
      v2[o] = hdt * v2[o] + u2[o];
      v3[o] = hdt * v3[o] + u3[o];
    }
    // epsilonikdot <- Dt/2 epsilonikdot + epsilonik (the mid-step strain)
    if (transient)
      for (long o = 0; o < 6 * npts; ++o) epsilonikdot[o] = hdt * epsilonikdot[o] + (R)1 * epsilonik[o];
    for (long o = 0; o < 6 * npts; ++o) sig[o] = (-hdt) * sig[o] + (R)(-1) * tau[o];
    // corrector
    stressupdate(v1.data(), v2.data(), v3.data(), m->lambda, m->mu, m->dx1, m->dx2, m->dx3, sx1, sx2, h3, sig, pitch);
    std::fill(moment.begin(), moment.end(), (R)0);
    const orc_layer* visc[2] = {m->linear, m->nonlinear};
    for (int k = 0; k < 2; ++k) {
      if (!visc[k]) continue;
      std::fill(v1.begin(), v1.end(), (R)0);
      rc = viscouseigenstress(m->mu, visc[k], m->prestress, sig, sx1, sx2, h3, moment.data(), (double*)0, dg[k], v1.data(), pitch);
      if (rc) return rc;
      const R cdt = (R)Dt;
      for (long o = 0; o < nu; ++o) gamma[o] = gamma[o] + cdt * v1[o];
    }
    if (m->faultcreep) {
      long off = 0;
      for (int k = 0; k < m->np; ++k) {
        const orc_plane& P = m->planes[k];
        rc = frictioneigenstress(P.x, P.y, P.z, P.width, P.length, P.strike, P.dip, P.rake, m->beta, sig, m->mu,
                                 m->faultcreep, sx1, sx2, h3, m->dx1, m->dx2, m->dx3, moment.data(), (double*)0,
                                 (R*)0, pitch);
        if (rc) return rc;
        if (m->patches) {
          rc = monitorfriction(P.x, P.y, P.z, P.width, P.length, P.strike, P.dip, P.rake, m->beta, sx1, sx2, sx3, h3,
                               m->dx1, m->dx2, m->dx3, sig, m->faultcreep, m->patches + off, P.px2, P.px3);
          if (rc) return rc;
          off += (long)P.px2 * P.px3;
        }
      }
    }
    if (transient) {
      // the corrector passes epsilonikdot as both the strain and the rate
      for (int k = 0; k < 2; ++k) {
        if (!trans[k]) continue;
        rc = transienteigenstress(m->mu, trans[k], sig, epsilonikdot.data(), sx1, sx2, h3, moment.data(),
                                  epsilonikdot.data(), (double*)0, dg[2 + k]);
        if (rc) return rc;
      }
      const R cdt = (R)Dt;
      for (long o = 0; o < 6 * npts; ++o) epsilonik[o] = epsilonik[o] + cdt * epsilonikdot[o];
      std::fill(epsilonikdot.begin(), epsilonikdot.end(), (R)0);
    }
    // inter-seismic loading (relax.f90:795-799)
    if (inter) interseismic(m, sx1, sx2, h3, moment.data(), tau);
    std::fill(v1.begin(), v1.end(), (R)0); std::fill(v2.begin(), v2.end(), (R)0); std::fill(v3.begin(), v3.end(), (R)0);
    std::fill(t1.begin(), t1.end(), (R)0); std::fill(t2.begin(), t2.end(), (R)0); std::fill(t3.begin(), t3.end(), (R)0);
    rc = equivalentbodyforce(moment.data(), m->dx1, m->dx2, m->dx3, sx1, sx2, h3, v1.data(), v2.data(), v3.data(),
                             t1.data(), t2.data(), t3.data(), pitch, (const R*)0);
    if (rc) return rc;
    // the rate of the surface loads over the step (relax.f90:805-806)
    if (m->ev_loff && m->ev_loff[e] > m->ev_loff[e - 1])
      traction(m->ev_loads + m->ev_loff[e - 1], m->ev_loff[e] - m->ev_loff[e - 1], sx1, sx2, m->dx1, m->dx2, t, Dt,
               t3.data(), true);
    rc = greenfunctioncowling(v1.data(), v2.data(), v3.data(), t1.data(), t2.data(), t3.data(), sx1, sx2, sx3, m->dx1,
                              m->dx2, m->dx3, m->lambda, m->mu, m->gam);
    if (rc) return rc;
    const R cdt = (R)Dt;
    for (long o = 0; o < nu; ++o) {
      u1[o] = u1[o] + cdt * v1[o];
      u2[o] = u2[o] + cdt * v2[o];
      u3[o] = u3[o] + cdt * v3[o];
    }
    for (long o = 0; o < 6 * npts; ++o) tau[o] = tau[o] + cdt * moment[o];
    if (m->patches) {
      long off = 0;
      for (int k = 0; k < m->np; ++k) {
        frictionadd(m->patches + off, (long)m->planes[k].px2 * m->planes[k].px3, Dt);
        off += (long)m->planes[k].px2 * m->planes[k].px3;
      }
    }
    t = t + Dt;
    // the event switch (relax.f90:858-886)
    if (e < m->ne && std::fabs(t - m->event_times[e]) < 1e-6) {
      e = e + 1;
      const int s0 = m->ev_off ? m->ev_off[e - 1] : 0, s1 = m->ev_off ? m->ev_off[e] : 0;
      const int l0 = m->ev_loff ? m->ev_loff[e - 1] : 0, l1 = m->ev_loff ? m->ev_loff[e] : 0;
      const int a0 = m->ev_tsoff ? m->ev_tsoff[e - 1] : 0, a1 = m->ev_tsoff ? m->ev_tsoff[e] : 0;
      const int b0 = m->ev_mgoff ? m->ev_mgoff[e - 1] : 0, b1 = m->ev_mgoff ? m->ev_mgoff[e] : 0;
      const int c0 = m->ev_cboff ? m->ev_cboff[e - 1] : 0, c1 = m->ev_cboff ? m->ev_cboff[e] : 0;
      const int h0 = m->ev_thoff ? m->ev_thoff[e - 1] : 0, h1 = m->ev_thoff ? m->ev_thoff[e] : 0;
      const int w0 = m->ev_vxoff ? m->ev_vxoff[e - 1] : 0, w1 = m->ev_vxoff ? m->ev_vxoff[e] : 0;
      if (s1 > s0 || l1 > l0 || a1 > a0 || b1 > b0 || c1 > c0 || h1 > h0) {
        std::fill(v1.begin(), v1.end(), (R)0); std::fill(v2.begin(), v2.end(), (R)0);
        std::fill(v3.begin(), v3.end(), (R)0); std::fill(t1.begin(), t1.end(), (R)0);
        std::fill(t2.begin(), t2.end(), (R)0); std::fill(t3.begin(), t3.end(), (R)0);
        std::vector<R> lost(MUT(18) ? 6 * npts : 0, (R)0);
        dislocations(m->ev_src ? m->ev_src + s0 : nullptr, s1 - s0, m->ev_ts ? m->ev_ts + a0 : nullptr, a1 - a0,
                     m->ev_mg ? m->ev_mg + b0 : nullptr, b1 - b0, m->ev_loads ? m->ev_loads + l0 : nullptr, l1 - l0,
                     m->lambda, m->mu, m->beta, m->nyquist, t, sx1, sx2, sx3, m->dx1, m->dx2, m->dx3, v1.data(),
                     v2.data(), v3.data(), t1.data(), t2.data(), t3.data(), MUT(18) ? lost.data() : tau,
                     m->ev_cb ? m->ev_cb + c0 : nullptr, c1 - c0, m->ev_th ? m->ev_th + h0 : nullptr, h1 - h0,
                     m->ev_vx ? m->ev_vx + 3L * w0 : nullptr, w1 - w0);
        rc = greenfunctioncowling(v1.data(), v2.data(), v3.data(), t1.data(), t2.data(), t3.data(), sx1, sx2, sx3,
                                  m->dx1, m->dx2, m->dx3, m->lambda, m->mu, m->gam);
        if (rc) return rc;
        for (long o = 0; o < nu; ++o) {
          u1[o] = u1[o] + v1[o];
          u2[o] = u2[o] + v2[o];
          u3[o] = u3[o] + v3[o];
        }
      }
    }
    for (long o = 0; o < 6 * npts; ++o) sig[o] = (R)(-1) * tau[o];
    stressupdate(u1, u2, u3, m->lambda, m->mu, m->dx1, m->dx2, m->dx3, sx1, sx2, h3, sig, pitch);
    // isoutput, relax.f90:1364-1376, advances the output counter
    const double etime = m->ne > 0 ? m->event_times[e - 1] : 0.0;
    bool out = ((0 == m->skip) && (std::fabs(t - oi * m->odt) < 1e-6 * m->odt)) ||
               ((0 < m->skip) && ((it - 1) % m->skip == 0)) || (std::fabs(t - etime) < 1e-6);
    if (info) {
      double* I = info + 10 * step;
      I[0] = Dt; I[1] = tm; I[2] = maxwell[0]; I[3] = maxwell[1]; I[4] = maxwell[2];
      I[5] = tensoramplitude(moment.data(), npts, m->dx1, m->dx2, m->dx3);
      I[6] = tensoramplitude(tau, npts, m->dx1, m->dx2, m->dx3);
      I[7] = t; I[8] = maxwell[3]; I[9] = maxwell[4];
    }
    if (out) ++oi;
  }
  *t_io = t; *oi_io = oi; *i_io = it; *e_io = e;
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI of the oracle (suffix _f: FP32 storage, _d: FP64 storage)
// ---------------------------------------------------------------------------
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))
#define ORC_BOTH(SUFFIX, R)                                                                                          \
  ORC_EXPORT int orc_fft3_##SUFFIX(R* d, int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, int dir) { \
    return fft3(d, sx1, sx2, sx3, dx1, dx2, dx3, dir);                                                               \
  }                                                                                                                  \
  ORC_EXPORT int orc_fft2_##SUFFIX(R* d, int sx1, int sx2, double dx1, double dx2, int dir) {                        \
    return fft2(d, sx1, sx2, dx1, dx2, dir);                                                                         \
  }                                                                                                                  \
  ORC_EXPORT void orc_elasticresponse_##SUFFIX(double lambda, double mu, R* f1, R* f2, R* f3, int sx1, int sx2,    \
                                               int sx3, double dx1, double dx2, double dx3) {                        \
    elasticresponse(lambda, mu, f1, f2, f3, sx1, sx2, sx3, dx1, dx2, dx3);                                           \
  }                                                                                                                  \
  ORC_EXPORT void orc_surfacetractioncowling_##SUFFIX(double lambda, double mu, double gamma, const R* u1,         \
                                                      const R* u2, const R* u3, int sx1, int sx2, int sx3,          \
                                                      double dx1, double dx2, double dx3, R* p1, R* p2, R* p3) {    \
    surfacetractioncowling(lambda, mu, gamma, u1, u2, u3, sx1, sx2, sx3, dx1, dx2, dx3, p1, p2, p3);                 \
  }                                                                                                                  \
  ORC_EXPORT void orc_cerrutimodified_##SUFFIX(const R* p1, const R* p2, const R* p3, double lambda, double mu,    \
                                               double gamma, R* u1, R* u2, R* u3, int sx1, int sx2, int sx3,        \
                                               double dx1, double dx2, double dx3) {                                 \
    cerrutimodified(p1, p2, p3, lambda, mu, gamma, u1, u2, u3, sx1, sx2, sx3, dx1, dx2, dx3);                        \
  }                                                                                                                  \
  ORC_EXPORT int orc_greenfunctioncowling_##SUFFIX(R* c1, R* c2, R* c3, R* t1, R* t2, R* t3, int sx1, int sx2,     \
                                                   int sx3, double dx1, double dx2, double dx3, double lambda,      \
                                                   double mu, double gamma) {                                        \
    return greenfunctioncowling(c1, c2, c3, t1, t2, t3, sx1, sx2, sx3, dx1, dx2, dx3, lambda, mu, gamma);            \
  }                                                                                                                  \
  ORC_EXPORT int orc_stressupdate_##SUFFIX(const R* v1, const R* v2, const R* v3, double lambda, double mu,        \
                                           double dx1, double dx2, double dx3, int sx1, int sx2, int sx3, R* sig) {  \
    return stressupdate(v1, v2, v3, lambda, mu, dx1, dx2, dx3, sx1, sx2, sx3, sig, (long)sx1 + 2);                   \
  }                                                                                                                  \
  ORC_EXPORT int orc_equivalentbodyforce_##SUFFIX(const R* sig, double dx1, double dx2, double dx3, int sx1,       \
                                                  int sx2, int sx3, R* c1, R* c2, R* c3, R* t1, R* t2, R* t3,       \
                                                  const R* mask) {                                                   \
    return equivalentbodyforce(sig, dx1, dx2, dx3, sx1, sx2, sx3, c1, c2, c3, t1, t2, t3, (long)sx1 + 2, mask);      \
  }                                                                                                                  \
  ORC_EXPORT int orc_viscouseigenstress_##SUFFIX(double mu, const orc_layer* st, const orc_tensor_layer* pre,      \
                                                 const R* sig, int sx1, int sx2, int sx3, R* moment,                \
                                                 double* maxwelltime, const R* dg0, R* gamma) {                      \
    return viscouseigenstress(mu, st, pre, sig, sx1, sx2, sx3, moment, maxwelltime, dg0, gamma, (long)sx1 + 2);      \
  }                                                                                                                  \
  ORC_EXPORT int orc_frictioneigenstress_##SUFFIX(double x, double y, double z, double L, double W, double strike, \
                                                  double dip, double rake, double beta, const R* sig, double mu,    \
                                                  const orc_layer* st, int sx1, int sx2, int sx3, double dx1,       \
                                                  double dx2, double dx3, R* moment, double* maxwelltime, R* vel) { \
    return frictioneigenstress(x, y, z, L, W, strike, dip, rake, beta, sig, mu, st, sx1, sx2, sx3, dx1, dx2, dx3,    \
                               moment, maxwelltime, vel, (long)sx1 + 2);                                             \
  }                                                                                                                  \
  ORC_EXPORT int orc_monitorfriction_##SUFFIX(double x, double y, double z, double L, double W, double strike,     \
                                              double dip, double rake, double beta, int sx1, int sx2, int sx3,      \
                                              int sx3_alloc, double dx1, double dx2, double dx3, const R* sig,      \
                                              const orc_layer* st, orc_patch* patch, int px2, int px3) {            \
    return monitorfriction(x, y, z, L, W, strike, dip, rake, beta, sx1, sx2, sx3, sx3_alloc, dx1, dx2, dx3, sig,     \
                           st, patch, px2, px3);                                                                     \
  }                                                                                                                  \
  ORC_EXPORT double orc_tensoramplitude_##SUFFIX(const R* t, long npts, double dx1, double dx2, double dx3) {      \
    return tensoramplitude(t, npts, dx1, dx2, dx3);                                                                  \
  }                                                                                                                  \
  ORC_EXPORT void orc_momentdensityshear_##SUFFIX(double mu, double slip, double x, double y, double z, double L,  \
                                                  double W, double strike, double dip, double rake, double beta,    \
                                                  int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,    \
                                                  R* sig) {                                                          \
    momentdensityshear(mu, slip, x, y, z, L, W, strike, dip, rake, beta, sx1, sx2, sx3, dx1, dx2, dx3, sig);         \
  }                                                                                                                  \
  ORC_EXPORT void orc_source_##SUFFIX(double mu, double s, double x, double y, double z, double L, double W,       \
                                      double strike, double dip, double rake, double beta, int sx1, int sx2,        \
                                      int sx3, double dx1, double dx2, double dx3, R* f1, R* f2, R* f3, R* t1,      \
                                      R* t2, R* t3) {                                                                \
    source(mu, s, x, y, z, L, W, strike, dip, rake, beta, sx1, sx2, sx3, dx1, dx2, dx3, f1, f2, f3, t1, t2, t3);     \
  }                                                                                                                  \
  ORC_EXPORT void orc_traction_##SUFFIX(const orc_source* l, int nl, int sx1, int sx2, double dx1, double dx2,      \
                                        double t, double Dt, R* t3, int rate) {                                     \
    traction(l, nl, sx1, sx2, dx1, dx2, t, Dt, t3, rate != 0);                                                       \
  }                                                                                                                  \
  ORC_EXPORT void orc_dislocations_##SUFFIX(const orc_source* s, int ns, const orc_source* ts, int nt,              \
                                            const orc_source* mg, int nm, const orc_source* l, int nl,              \
                                            double lambda, double mu, double beta, double nyquist, double t,        \
                                            int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, R* v1,   \
                                            R* v2, R* v3, R* t1, R* t2, R* t3, R* tau, const orc_weak* cb, int nc,  \
                                            const orc_tetra* th, int nth, const double* vtx, int nv) {              \
    dislocations(s, ns, ts, nt, mg, nm, l, nl, lambda, mu, beta, nyquist, t, sx1, sx2, sx3, dx1, dx2, dx3, v1, v2,  \
                 v3, t1, t2, t3, tau, cb, nc, th, nth, vtx, nv);                                                     \
  }                                                                                                                  \
  ORC_EXPORT void orc_tensilesource_##SUFFIX(double lambda, double mu, double s, double x, double y, double z,       \
                                             double L, double W, double strike, double dip, double beta, int sx1,  \
                                             int sx2, int sx3, double dx1, double dx2, double dx3, R* f1, R* f2,   \
                                             R* f3) {                                                                \
    tensilesource(lambda, mu, s, x, y, z, L, W, strike, dip, beta, sx1, sx2, sx3, dx1, dx2, dx3, f1, f2, f3);      \
  }                                                                                                                  \
  ORC_EXPORT void orc_mogisource_##SUFFIX(double lambda, double mu, double o, double x, double y, double z, int sx1, \
                                          int sx2, int sx3, double dx1, double dx2, double dx3, R* f1, R* f2,       \
                                          R* f3) {                                                                   \
    mogisource(lambda, mu, o, x, y, z, sx1, sx2, sx3, dx1, dx2, dx3, f1, f2, f3);                                  \
  }                                                                                                                  \
  ORC_EXPORT void orc_momentdensitytensile_##SUFFIX(double lambda, double mu, double s, double x, double y,         \
                                                    double z, double L, double W, double strike, double dip,        \
                                                    double beta, int sx1, int sx2, int sx3, double dx1, double dx2, \
                                                    double dx3, R* sig) {                                            \
    momentdensitytensile(lambda, mu, s, x, y, z, L, W, strike, dip, beta, sx1, sx2, sx3, dx1, dx2, dx3, sig);      \
  }                                                                                                                  \
  ORC_EXPORT void orc_momentdensitymogi_##SUFFIX(double lambda, double mu, double o, double x, double y, double z,  \
                                                 int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,    \
                                                 R* sig) {                                                           \
    momentdensitymogi(lambda, mu, o, x, y, z, sx1, sx2, sx3, dx1, dx2, dx3, sig);                                  \
  }                                                                                                                  \
  ORC_EXPORT int orc_exportpoints_##SUFFIX(const R* u1, const R* u2, const R* u3, const R* sig, int sx1, int sx2,  \
                                           int sx3, double dx1, double dx2, double dx3, const double* pts, int n,   \
                                           double rot, double* out) {                                               \
    return exportpoints(u1, u2, u3, sig, sx1, sx2, sx3, dx1, dx2, dx3, pts, n, rot, out);                           \
  }                                                                                                                  \
  ORC_EXPORT int orc_rk2_run_##SUFFIX(const orc_model* m, R* u1, R* u2, R* u3, R* tau, R* sig, R* gamma,           \
                                      R* epsilonik, double* t, int* oi, int* it, int* e, int nsteps,                \
                                      const double* dt_replay, double* info) {                                       \
    return rk2_run(m, u1, u2, u3, tau, sig, gamma, epsilonik, t, oi, it, e, nsteps, dt_replay, info);                \
  }                                                                                                                  \
  ORC_EXPORT int orc_transienteigenstress_##SUFFIX(double mu, const orc_layer* st, const R* sig, const R* eik,     \
                                                   int sx1, int sx2, int sx3, R* moment, R* eikdot,                 \
                                                   double* maxwelltime, const R* dg0) {                              \
    return transienteigenstress(mu, st, sig, eik, sx1, sx2, sx3, moment, eikdot, maxwelltime, dg0);                  \
  }                                                                                                                  \
  ORC_EXPORT void orc_builddgammadot0_##SUFFIX(int sx1, int sx2, int sx3, double dx1, double dx2, double dx3,      \
                                               double beta, int nz, const orc_weak* zones, R* dg) {                 \
    builddgammadot0(sx1, sx2, sx3, dx1, dx2, dx3, beta, nz, zones, dg);                                              \
  }

ORC_BOTH(f, float)
ORC_BOTH(d, double)

// Okada (1992): DC3D's displacement (single-precision arguments, as dc3d.f)
// and the green_space.f90 wrapper at one point
ORC_EXPORT int orc_dc3d(float alpha, float x, float y, float z, float depth, float dip, float al1, float al2,
                        float aw1, float aw2, float d1, float d2, float d3, float* u) {
  return okada::dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, d1, d2, d3, u, u + 1, u + 2);
}
ORC_EXPORT int orc_dc3d_core(double alpha, double x, double y, double z, double depth, double dip, double al1,
                             double al2, double aw1, double aw2, const double* U, double* u) {
  return okada::dc3d_core(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, U, u);
}
ORC_EXPORT void orc_okada_point(double lambda, double mu, double slip, double opening, double x1, double x2,
                                double x3, double width, double length, double strike, double dip, double rake,
                                double xrec, double yrec, double zrec, double* disp) {
  okada::point(lambda, mu, slip, opening, x1, x2, x3, width, length, strike, dip, rake, xrec, yrec, zrec, disp);
}
// dislocations_disp at a list of grid points only (0-based i1, i2, i3
// triples of the half grid): out (n x 3) receives the displacement each
// point would gain, accumulated in single precision as the grid's REAL*4
// field does, from zero.
ORC_EXPORT void orc_dislocations_disp_points(const orc_source* sc, int ns, const orc_source* tsc, int nt,
                                             double lambda, double mu, double nyquist, int sx1, int sx2, double dx1,
                                             double dx2, double dx3, const int* idx, long n, float* out) {
  const double dmin = nyquist * std::min(dx1, std::min(dx2, dx3));
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < n; ++p) {
    const double xr = shifted1(idx[3 * p], sx1, dx1), yr = shifted1(idx[3 * p + 1], sx2, dx2);
    const double zr = (double)idx[3 * p + 2] * dx3;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    double d[3];
    for (int k = 0; k < ns; ++k)
      if (dmin >= sc[k].length || dmin >= sc[k].width) {
        okada::point(lambda, mu, sc[k].slip, 0.0, sc[k].x, sc[k].y, sc[k].z, sc[k].width, sc[k].length,
                     sc[k].strike, sc[k].dip, sc[k].rake, xr, yr, zr, d);
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + (float)d[c];
      }
    for (int k = 0; k < nt; ++k)
      if (dmin >= tsc[k].length || dmin >= tsc[k].width) {
        okada::point(lambda, mu, 0.0, tsc[k].opening, tsc[k].x, tsc[k].y, tsc[k].z, tsc[k].width, tsc[k].length,
                     tsc[k].strike, tsc[k].dip, 0.0, xr, yr, zr, d);
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + (float)d[c];
      }
    for (int c = 0; c < 3; ++c) out[3 * p + c] = acc[c];
  }
}
#define ORC_OKADA(SUFFIX, R)                                                                                      \
  ORC_EXPORT void orc_grnfct_okada_##SUFFIX(double lambda, double mu, double slip, double opening, double x,      \
                                            double y, double z, double width, double length, double strike,       \
                                            double dip, double rake, int sx1, int sx2, int sx3, double dx1,       \
                                            double dx2, double dx3, R* u1, R* u2, R* u3) {                        \
    grnfct_okada(lambda, mu, slip, opening, x, y, z, width, length, strike, dip, rake, sx1, sx2, sx3, dx1, dx2,  \
                 dx3, u1, u2, u3);                                                                               \
  }                                                                                                              \
  ORC_EXPORT void orc_dislocations_disp_##SUFFIX(const orc_source* sc, int ns, const orc_source* tsc, int nt,    \
                                                 double lambda, double mu, double nyquist, int sx1, int sx2,     \
                                                 int h3, double dx1, double dx2, double dx3, R* u1, R* u2,       \
                                                 R* u3) {                                                        \
    dislocations_disp(sc, ns, tsc, nt, lambda, mu, nyquist, sx1, sx2, h3, dx1, dx2, dx3, u1, u2, u3);            \
  }
ORC_OKADA(f, float)
ORC_OKADA(d, double)

#define ORC_EIGEN(SUFFIX, R)                                                                                      \
  ORC_EXPORT void orc_eigenstrain_##SUFFIX(const orc_weak* cb, int nc, const orc_tetra* th, int nth,             \
                                           const double* vtx, int nv, double lambda, double mu, double beta,     \
                                           int sx1, int sx2, int sx3, double dx1, double dx2, double dx3, R* v1, \
                                           R* v2, R* v3, R* tau) {                                              \
    eigenstrain(cb, nc, th, nth, vtx, nv, lambda, mu, beta, sx1, sx2, sx3, dx1, dx2, dx3, v1, v2, v3, tau);     \
  }                                                                                                              \
  ORC_EXPORT void orc_cuboidsource_##SUFFIX(double lambda, double mu, const float* e, double x, double y,        \
                                            double z, double L, double W, double T, double strike, double dip,  \
                                            double beta, int sx1, int sx2, int sx3, double dx1, double dx2,     \
                                            double dx3, R* f1, R* f2, R* f3) {                                  \
    cuboidsource(lambda, mu, e, x, y, z, L, W, T, strike, dip, beta, sx1, sx2, sx3, dx1, dx2, dx3, f1, f2, f3);  \
  }                                                                                                              \
  ORC_EXPORT void orc_momentdensitycuboid_##SUFFIX(double mu, double lambda, const float* e, double x, double y, \
                                                   double z, double L, double W, double T, double strike,       \
                                                   double dip, double beta, int sx1, int sx2, int sx3,          \
                                                   double dx1, double dx2, double dx3, R* sig) {                \
    momentdensitycuboid(mu, lambda, e, x, y, z, L, W, T, strike, dip, beta, sx1, sx2, sx3, dx1, dx2, dx3, sig);  \
  }                                                                                                              \
  ORC_EXPORT void orc_tetrahedronsource_##SUFFIX(double lambda, double mu, const float* e, const double* A,      \
                                                 const double* B, const double* C, const double* D, int sx1,    \
                                                 int sx2, int sx3, double dx1, double dx2, double dx3, R* f1,   \
                                                 R* f2, R* f3) {                                                \
    tetrahedronsource(lambda, mu, e, A, B, C, D, sx1, sx2, sx3, dx1, dx2, dx3, f1, f2, f3);                      \
  }                                                                                                              \
  ORC_EXPORT void orc_momentdensitytetrahedron_##SUFFIX(double mu, double lambda, const float* e,                \
                                                        const double* A, const double* B, const double* C,      \
                                                        const double* D, int sx1, int sx2, int sx3, double dx1, \
                                                        double dx2, double dx3, R* sig) {                       \
    momentdensitytetrahedron(mu, lambda, e, A, B, C, D, sx1, sx2, sx3, dx1, dx2, dx3, sig);                      \
  }
ORC_EIGEN(f, float)
ORC_EIGEN(d, double)

ORC_EXPORT void orc_frictionadd(orc_patch* patch, long npatch, double dt) { frictionadd(patch, npatch, dt); }
ORC_EXPORT void orc_set_mutation(int k) { g_mut = k; }
ORC_EXPORT int orc_sizeof_patch(void) { return (int)sizeof(orc_patch); }
ORC_EXPORT int orc_sizeof_model(void) { return (int)sizeof(orc_model); }
ORC_EXPORT int orc_sizeof_weak(void) { return (int)sizeof(orc_weak); }
ORC_EXPORT int orc_sizeof_tetra(void) { return (int)sizeof(orc_tetra); }
ORC_EXPORT int orc_sizeof_source(void) { return (int)sizeof(orc_source); }
ORC_EXPORT int orc_shiftedindex(double x1, double x2, double x3, int sx1, int sx2, int sx3, double dx1, double dx2,
                                double dx3, int* idx) {
  return shiftedindex(x1, x2, x3, sx1, sx2, sx3, dx1, dx2, dx3, idx, idx + 1, idx + 2);
}
