// okada.hip -- Okada (1992) displacement of the shear patches and tensile
// cracks below the Nyquist size at the first event (relax.f90:357-360,
// dislocations_disp :1300-1349), on the device.
//
// grnfct_okada (okada/green_space.f90:119-146) evaluates, for every point
// of the half grid, Okada's closed-form displacement of a rectangular
// dislocation in a half-space (okada/dc3d.f, DC3D) and adds it to u.  Here
// one thread owns one grid point and walks the list of small sources in
// order, adding each one's displacement rounded to FP32 as the reference's
// REAL() does.  The closed form (Y. Okada, BSSA 82:1018-1040, Table 6) is
// evaluated in FP64 from single-precision arguments, as DC3D's REAL*4
// interface has them, with DC3D's treatment of singular points:
//  - |xi|, |eta|, |q| below 1e-6 count as 0;
//  - a point on a fault edge gives no displacement for that source;
//  - on the negative extension of an edge, ln(R + xi) becomes -ln(R - xi)
//    with X11 = X32 = 0 (likewise in eta);
//  - atan(xi eta / (q R)) is 0 for q = 0; |cos(dip)| < 1e-6 counts as 0.
#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.h"
#include "model.h"

namespace {

// One source in device form: the per-source constants of green_space.f90's
// okada() (its frame change and DC3D's single-precision arguments), formed
// on the host.
struct OkadaSource {
  double x1, x2;          // upper corner at the start of the strike (north, east)
  double csst, ssst;      // cos, sin of the strike
  double sd, cd;          // DC3D's dip constants (DCCON0)
  float depth, al2, aw1;  // DEPTH, AL2 = length, AW1 = -width (AL1 = AW2 = 0)
  float d1, d2, d3;       // DISL1..3: slip cos(rake), slip sin(rake), opening
};

constexpr double OK_EPS = 1.0e-6;
constexpr double OK_PI2 = 6.283185307179586;

// Displacement of one corner: the A part (infinite medium) and, for the
// image, the B and C parts, each summed over the three dislocation kinds and
// divided by 2 pi.
__device__ __forceinline__ void ok_corner(double xi, double et, double q, double z, double sd, double cd, double alpha,
                                          double U1, double U2, double U3, bool kxi, bool ket, bool image,
                                          double fA[3], double fB[3], double fC[3]) {
#pragma clang fp contract(off)
  const double R2 = xi * xi + et * et + q * q;
  const double R = sqrt(R2);
  const double yt = et * cd + q * sd, dt = et * sd - q * cd;
  const double th = (q == 0.0) ? 0.0 : atan(xi * et / (q * R));
  double lx, X11, X32, le, Y11, Y32;
  if (kxi) {
    lx = -log(R - xi); X11 = 0.0; X32 = 0.0;
  } else {
    const double rx = R + xi;
    lx = log(rx); X11 = 1.0 / (R * rx); X32 = (2.0 * R + xi) / (R * R * R * rx * rx);
  }
  if (ket) {
    le = -log(R - et); Y11 = 0.0; Y32 = 0.0;
  } else {
    const double re = R + et;
    le = log(re); Y11 = 1.0 / (R * re); Y32 = (2.0 * R + et) / (R * R * R * re * re);
  }
  const double a = alpha, b = 1.0 - alpha, q2 = q * q;
  // strike-slip, dip-slip and tensile rows of Table 6, A part
  {
    const double s1 = th / 2 + a / 2 * xi * q * Y11, s2 = a / 2 * q / R, s3 = b / 2 * le - a / 2 * q2 * Y11;
    const double d1 = a / 2 * q / R, d2 = th / 2 + a / 2 * et * q * X11, d3 = b / 2 * lx - a / 2 * q2 * X11;
    const double t1 = -b / 2 * le - a / 2 * q2 * Y11, t2 = -b / 2 * lx - a / 2 * q2 * X11;
    const double t3 = th / 2 - a / 2 * q * (et * X11 + xi * Y11);
    fA[0] = (U1 * s1 + U2 * d1 + U3 * t1) / OK_PI2;
    fA[1] = (U1 * s2 + U2 * d2 + U3 * t2) / OK_PI2;
    fA[2] = (U1 * s3 + U2 * d3 + U3 * t3) / OK_PI2;
  }
  if (!image) return;
  const double Rd = R + dt;
  double I3, I4;
  if (cd != 0.0) {
    if (xi == 0.0) {
      I4 = 0.0;
    } else {
      const double X = sqrt(xi * xi + q2);
      I4 = (xi / Rd * sd * cd + 2.0 * atan((et * (X + q * cd) + X * (R + X) * sd) / (xi * (R + X) * cd))) / (cd * cd);
    }
    I3 = (yt * cd / Rd - le + sd * log(Rd)) / (cd * cd);
  } else {
    I3 = (et / Rd + yt * q / (Rd * Rd) - le) / 2.0;
    I4 = xi * yt / (Rd * Rd) / 2.0;
  }
  const double I1 = -xi / Rd * cd - I4 * sd;
  const double I2 = log(Rd) + I3 * sd;
  const double k = b / a;
  {
    const double s1 = -xi * q * Y11 - th - k * I1 * sd, s2 = -q / R + k * yt / Rd * sd, s3 = q2 * Y11 - k * I2 * sd;
    const double d1 = -q / R + k * I3 * sd * cd, d2 = -et * q * X11 - th - k * xi / Rd * sd * cd;
    const double d3 = q2 * X11 + k * I4 * sd * cd;
    const double t1 = q2 * Y11 - k * I3 * sd * sd, t2 = q2 * X11 + k * xi / Rd * sd * sd;
    const double t3 = q * (et * X11 + xi * Y11) - th - k * I4 * sd * sd;
    fB[0] = (U1 * s1 + U2 * d1 + U3 * t1) / OK_PI2;
    fB[1] = (U1 * s2 + U2 * d2 + U3 * t2) / OK_PI2;
    fB[2] = (U1 * s3 + U2 * d3 + U3 * t3) / OK_PI2;
  }
  {
    const double R3 = R * R * R;
    const double ct = dt + z, h = q * cd - z;
    const double Z32 = sd / R3 - h * Y32;
    const double s1 = b * xi * Y11 * cd - a * xi * q * Z32;
    const double s2 = b * (cd / R + 2.0 * q * Y11 * sd) - a * ct * q / R3;
    const double s3 = b * q * Y11 * cd - a * (ct * et / R3 - z * Y11 + xi * xi * Z32);
    const double d1 = b * cd / R - q * Y11 * sd - a * ct * q / R3;
    const double d2 = b * yt * X11 - a * ct * et * q * X32;
    const double d3 = -dt * X11 - xi * Y11 * sd - a * ct * (X11 - q2 * X32);
    const double t1 = -b * (sd / R + q * Y11 * cd) - a * (z * Y11 - q2 * Z32);
    const double t2 = b * 2.0 * xi * Y11 * sd + dt * X11 - a * ct * (X11 - q2 * X32);
    const double t3 = b * (yt * X11 + xi * Y11 * cd) + a * q * (ct * et * X32 + xi * Z32);
    fC[0] = (U1 * s1 + U2 * d1 + U3 * t1) / OK_PI2;
    fC[1] = (U1 * s2 + U2 * d2 + U3 * t2) / OK_PI2;
    fC[2] = (U1 * s3 + U2 * d3 + U3 * t3) / OK_PI2;
  }
}

// DC3D's displacement at (X, Y, Z) (its single-precision arguments, widened);
// false at a singular point (on a fault edge).
__device__ bool ok_dc3d(const OkadaSource& S, double alpha, double x, double y, double z, double u[3]) {
#pragma clang fp contract(off)
  const double sd = S.sd, cd = S.cd, c = S.depth;
  const double U1 = S.d1, U2 = S.d2, U3 = S.d3;
  double xi[2] = {x - 0.0, x - (double)S.al2};
  for (int j = 0; j < 2; ++j)
    if (fabs(xi[j]) < OK_EPS) xi[j] = 0.0;
  u[0] = u[1] = u[2] = 0.0;
  for (int section = 0; section < 2; ++section) {
    const double d = section == 0 ? c + z : c - z;
    const double p = y * cd + d * sd;
    double q = y * sd - d * cd;
    double et[2] = {p - (double)S.aw1, p - 0.0};
    if (fabs(q) < OK_EPS) q = 0.0;
    for (int k = 0; k < 2; ++k)
      if (fabs(et[k]) < OK_EPS) et[k] = 0.0;
    if (q == 0.0 && ((xi[0] * xi[1] <= 0.0 && et[0] * et[1] == 0.0) || (et[0] * et[1] <= 0.0 && xi[0] * xi[1] == 0.0)))
      return false;
    const double r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + q * q);
    const double r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + q * q);
    const double r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + q * q);
    const bool kxi[2] = {xi[0] < 0.0 && r21 + xi[1] < OK_EPS, xi[0] < 0.0 && r22 + xi[1] < OK_EPS};
    const bool ket[2] = {et[0] < 0.0 && r12 + et[1] < OK_EPS, et[0] < 0.0 && r22 + et[1] < OK_EPS};
    for (int kk = 0; kk < 2; ++kk)
      for (int j = 0; j < 2; ++j) {
        double fA[3], fB[3] = {0.0, 0.0, 0.0}, fC[3] = {0.0, 0.0, 0.0};
        ok_corner(xi[j], et[kk], q, z, sd, cd, alpha, U1, U2, U3, kxi[kk], ket[j], section == 1, fA, fB, fC);
        double du0, du1, du2;
        if (section == 0) {
          // the real source: -A (dc3d.f's real-source sum)
          du0 = -fA[0];
          du1 = -fA[1] * cd + fA[2] * sd;
          du2 = -fA[1] * sd - fA[2] * cd;
        } else {
          // the image: A + B + z C, the vertical part of C reversed
          du0 = fA[0] + fB[0] + z * fC[0];
          du1 = (fA[1] + fB[1] + z * fC[1]) * cd - (fA[2] + fB[2] + z * fC[2]) * sd;
          du2 = (fA[1] + fB[1] - z * fC[1]) * sd + (fA[2] + fB[2] - z * fC[2]) * cd;
        }
        const double sgn = (j + kk == 1) ? -1.0 : 1.0;  // Chinnery's sum
        u[0] += sgn * du0;
        u[1] += sgn * du1;
        u[2] += sgn * du2;
      }
  }
  return true;
}

// One thread per point of the h3 top planes of u (rows of ld floats, fields
// fs apart): the sources' displacements added in their order.
__global__ __launch_bounds__(256) void k_okada(const OkadaSource* __restrict__ src, int ns, float alpha, int n1,
                                               int n2, int h3, double dx1, double dx2, double dx3, float* u, long ld,
                                               long fs) {
#pragma clang fp contract(off)
  const long N = (long)n1 * n2 * h3;
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= N) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  // shiftedcoordinates (elastic3d.f90:4027-4048) and x3 = (i3-1) dx3
  const double xr = (double)(i1 < n1 / 2 ? i1 : i1 - n1) * dx1;
  const double yr = (double)(i2 < n2 / 2 ? i2 : i2 - n2) * dx2;
  const double zr = (double)i3 * dx3;
  const long of = i1 + ld * (i2 + (long)n2 * i3);
  float u1 = u[of], u2 = u[fs + of], u3 = u[2 * fs + of];
  const double Z = (double)(float)(-zr);
  for (int k = 0; k < ns; ++k) {
    const OkadaSource S = src[k];
    const double X = (double)(float)((xr - S.x1) * S.csst + (yr - S.x2) * S.ssst);
    const double Y = (double)(float)((xr - S.x1) * S.ssst - (yr - S.x2) * S.csst);
    double d[3];
    if (!ok_dc3d(S, (double)alpha, X, Y, Z, d)) continue;  // singular: no displacement
    const double UX = (double)(float)d[0], UY = (double)(float)d[1], UZ = (double)(float)d[2];
    u1 = u1 + (float)(UX * S.csst + UY * S.ssst);
    u2 = u2 + (float)(UX * S.ssst - UY * S.csst);
    u3 = u3 + (float)(-UZ);
  }
  u[of] = u1;
  u[fs + of] = u2;
  u[2 * fs + of] = u3;
}

}  // namespace

// The device form of a source given in the deck's convention (upper corner
// at the start of the strike, north-east-down, degrees), as green_space.f90's
// okada() derives DC3D's arguments from it.
static OkadaSource okada_source(double slip, double opening, const relax_source_t& s) {
  const double degtorad = (double)1.745329252e-2f;  // a single-precision literal (green_space.f90:47)
  OkadaSource S;
  S.x1 = s.x;
  S.x2 = s.y;
  const double st = s.strike * degtorad, ra = s.rake * degtorad;
  S.csst = std::cos(st);
  S.ssst = std::sin(st);
  // DCCON0: the dip in degrees as DC3D's single-precision argument
  const double dip = (double)(float)s.dip, p18 = 6.283185307179586 / 360.0;
  S.sd = std::sin(dip * p18);
  S.cd = std::cos(dip * p18);
  if (std::fabs(S.cd) < 1.0e-6) {
    S.cd = 0.0;
    S.sd = S.sd > 0.0 ? 1.0 : (S.sd < 0.0 ? -1.0 : S.sd);
  }
  S.depth = (float)s.z;
  S.al2 = (float)s.length;
  S.aw1 = -(float)s.width;
  S.d1 = (float)(slip * std::cos(ra));
  S.d2 = (float)(slip * std::sin(ra));
  S.d3 = (float)opening;
  return S;
}

// Launch over a padded or packed field set; the sources are staged in src
// (device memory of ns OkadaSource entries) from the host list hs.
static int launch_okada(const std::vector<OkadaSource>& hs, void* src, double lambda, double mu, int n1, int n2,
                        int h3, double dx1, double dx2, double dx3, float* u, long ld, long fs, hipStream_t st) {
  if (hs.empty()) return RELAX_OK;
  HIP_TRY(hipMemcpyAsync(src, hs.data(), sizeof(OkadaSource) * hs.size(), hipMemcpyHostToDevice, st));
  const float alpha = (float)((lambda + mu) / (lambda + 2.0 * mu));
  const long N = (long)n1 * n2 * h3;
  k_okada<<<(unsigned)((N + 255) / 256), 256, 0, st>>>((const OkadaSource*)src, (int)hs.size(), alpha, n1, n2, h3,
                                                      dx1, dx2, dx3, u, ld, fs);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// dislocations_disp (relax.f90:1300-1349) on the model: the Okada
// displacement of the shear patches and tensile cracks at or below the
// nyquist size added to u, then sig = -tau + C:eps(u).
int relax_model_okada(relax_model_t* m, const relax_source_t* sc, int ns, const relax_source_t* tsc, int nt,
                      double nyquist) {
  if (!m || m->batch || ns < 0 || nt < 0 || (ns > 0 && !sc) || (nt > 0 && !tsc)) return RELAX_ERR_ARG;
  const double dmin = nyquist * std::min(m->desc.dx1, std::min(m->desc.dx2, m->desc.dx3));
  std::vector<OkadaSource> hs;
  for (int k = 0; k < ns; ++k)
    if (dmin >= sc[k].length || dmin >= sc[k].width) hs.push_back(okada_source(sc[k].slip, 0.0, sc[k]));
  for (int k = 0; k < nt; ++k)
    if (dmin >= tsc[k].length || dmin >= tsc[k].width) {
      relax_source_t t = tsc[k];
      t.rake = 0.0;
      hs.push_back(okada_source(0.0, tsc[k].opening, t));
    }
  if (hs.empty()) return RELAX_OK;
  void* src = nullptr;
  HIP_TRY(rx_malloc(&src, sizeof(OkadaSource) * hs.size()));
  const Grid& g = m->g;
  int rc = launch_okada(hs, src, m->desc.lambda, m->desc.mu, g.n1, g.n2, g.h3, m->desc.dx1, m->desc.dx2,
                        m->desc.dx3, m->u, g.n1, g.H, m->stream);
  if (!rc) rc = model_stress_set(m);
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
  (void)rx_free(src);
  return rc;
}

// grnfct_okada, okada/green_space.f90:119-146 (host arrays, added to)
int relax_grnfct_okada(double lambda, double mu, double slip, double opening, double x, double y, double z,
                       double width, double length, double strike, double dip, double rake, int sx1, int sx2, int sx3,
                       double dx1, double dx2, double dx3, float* u1, float* u2, float* u3, int align_data) {
  if (!u1 || !u2 || !u3 || sx1 < 1 || sx2 < 1 || sx3 < 1) return RELAX_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return RELAX_ERR_NODEVICE;
  const long ld = align_data ? sx1 + 2 : sx1;
  const long fs = ld * sx2 * sx3;
  const relax_source_t s{slip, x, y, z, width, length, strike, dip, rake, 0.0, 0.0, 0.0, opening};
  std::vector<OkadaSource> hs{okada_source(slip, opening, s)};
  float* d = nullptr;
  void* src = nullptr;
  HIP_TRY(rx_malloc(&d, sizeof(float) * 3 * fs));
  int rc = rx_malloc(&src, sizeof(OkadaSource)) == hipSuccess ? RELAX_OK : RELAX_ERR_NOMEM;
  float* h[3] = {u1, u2, u3};
  for (int k = 0; k < 3 && !rc; ++k)
    if (hipMemcpy(d + k * fs, h[k], sizeof(float) * fs, hipMemcpyHostToDevice) != hipSuccess) rc = RELAX_ERR_HIP;
  if (!rc) rc = launch_okada(hs, src, lambda, mu, sx1, sx2, sx3, dx1, dx2, dx3, d, ld, fs, 0);
  if (!rc && rx_device_sync() != hipSuccess) rc = RELAX_ERR_HIP;
  for (int k = 0; k < 3 && !rc; ++k)
    if (hipMemcpy(h[k], d + k * fs, sizeof(float) * fs, hipMemcpyDeviceToHost) != hipSuccess) rc = RELAX_ERR_HIP;
  (void)rx_free(d);
  if (src) (void)rx_free(src);
  return rc;
}
