// eigen.hip -- distributed eigenstrain sources of an event (--with-eigenstrain):
// cuboids (cuboidSource, elastic3d.f90:1680-1884; momentDensityCuboid
// :2927-3029) and tetrahedra (tetrahedronSource :1916-2039;
// momentDensityTetrahedron :2788-2873), called from dislocations
// (relax.f90:1151-1178, :1192-1223).
//
// The host stages one record per volume with everything that does not depend
// on the grid point (rotated centre, culling radius, face normals and
// centroids, tractions C:e.n); two kernels then run one thread per point of
// the top half: the body forces of all volumes into the Green solve's input,
// and their moment densities into tau.  Each volume adds in the reference's
// order with its rounding to single precision: two roundings per cuboid and
// component of the force, one per tetrahedron; the moment densities follow
// the TENSOR operators (.times., .plus., .sdyad., elastic3d.f90:186-255), which
// round after every operation.
#include <cmath>
#include <vector>

#include "kernels.h"
#include "model.h"

namespace {

struct CuboidDev {
  double x, y, z, L, W, T, beta, lim;
  double cs, ss, cd, sd, dip, xr, yr, zr;
  double le11, le22, le33, m12, m13, m23;  // Hooke's law terms of the body force
  float e[6];
  float SS[6], NS[6], SD[6], NN[6], ND[6], DD[6];  // sdyad of the strike (s), normal (n), dip (d) directions
};

struct TetraDev {
  double n[4][3], c[4][3], tr[4][3];  // outward unit normals, face centroids, tractions (C:e).n
  float e[6];
};

// the cull of both cuboid routines multiplies by SQRT(2.), a single-precision
// root (elastic3d.f90:1781, :2965)
const double SQRT2_SINGLE = (double)1.41421353816986083984375f;

void sdyad(const double a[3], const double b[3], float out[6]) {
  out[0] = (float)(a[0] * b[0]);
  out[1] = (float)((a[0] * b[1] + a[1] * b[0]) / 2.0);
  out[2] = (float)((a[0] * b[2] + a[2] * b[0]) / 2.0);
  out[3] = (float)(a[1] * b[1]);
  out[4] = (float)((a[1] * b[2] + a[2] * b[1]) / 2.0);
  out[5] = (float)(a[2] * b[2]);
}

CuboidDev make_cuboid(const relax_weak_t& c, double lambda, double mu, double beta) {
  // the driver passes (width, length) as the routines' (L, W)
  CuboidDev C{};
  C.x = c.x; C.y = c.y; C.z = c.z;
  C.L = c.width; C.W = c.length; C.T = c.thickness; C.beta = beta;
  C.cs = std::cos(c.strike); C.ss = std::sin(c.strike);
  C.cd = std::cos(c.dip); C.sd = std::sin(c.dip);
  C.dip = c.dip;
  const double Wp = C.W * (1.0 + 2.0 * beta) / 2.0, Lp = C.L * (1.0 + 2.0 * beta) / 2.0,
               Tp = C.T * (1.0 + 2.0 * beta) / 2.0;
  C.lim = SQRT2_SINGLE * std::fmax(Wp, std::fmax(Lp, Tp));
  const double x2r = C.cs * C.x - C.ss * C.y;
  C.xr = C.cd * x2r - C.sd * C.z;
  C.yr = C.ss * C.x + C.cs * C.y;
  C.zr = C.sd * x2r + C.cd * C.z;
  const relax_tensor_t& e = c.e;
  const float ef[6] = {e.s11, e.s12, e.s13, e.s22, e.s23, e.s33};
  for (int k = 0; k < 6; ++k) C.e[k] = ef[k];
  const double ekk = (double)e.s11 + (double)e.s22 + (double)e.s33;
  C.le11 = lambda * ekk + 2.0 * mu * e.s11;
  C.le22 = lambda * ekk + 2.0 * mu * e.s22;
  C.le33 = lambda * ekk + 2.0 * mu * e.s33;
  C.m12 = 2.0 * mu * e.s12; C.m13 = 2.0 * mu * e.s13; C.m23 = 2.0 * mu * e.s23;
  const double n[3] = {C.cd * C.cs, -C.cd * C.ss, -C.sd}, s[3] = {-C.ss, -C.cs, 0.0},
               d[3] = {C.cs * C.sd, -C.ss * C.sd, C.cd};
  sdyad(s, s, C.SS); sdyad(n, s, C.NS); sdyad(s, d, C.SD);
  sdyad(n, n, C.NN); sdyad(n, d, C.ND); sdyad(d, d, C.DD);
  return C;
}

// false for a tetrahedron without volume, which the reference skips
bool make_tetra(const relax_tetrahedron_t& t, const double* vtx, double lambda, double mu, TetraDev& T) {
  const double* P[4] = {vtx + 3L * (t.i1 - 1), vtx + 3L * (t.i2 - 1), vtx + 3L * (t.i3 - 1), vtx + 3L * (t.i4 - 1)};
  // face k is opposite vertex k: nA=(C-B)x(D-B), nB=(D-C)x(A-C),
  // nC=(A-D)x(B-D), nD=(B-A)x(C-A), centroids (B+C+D)/3 and so on
  double nrm[4];
  for (int k = 0; k < 4; ++k) {
    const double *p = P[(k + 1) % 4], *q = P[(k + 2) % 4], *r = P[(k + 3) % 4];
    const double u[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]}, v[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]};
    T.n[k][0] = u[1] * v[2] - u[2] * v[1];
    T.n[k][1] = u[2] * v[0] - u[0] * v[2];
    T.n[k][2] = u[0] * v[1] - u[1] * v[0];
    for (int j = 0; j < 3; ++j) T.c[k][j] = (p[j] + q[j] + r[j]) / 3.0;
    nrm[k] = std::sqrt(T.n[k][0] * T.n[k][0] + T.n[k][1] * T.n[k][1] + T.n[k][2] * T.n[k][2]);
  }
  for (int k = 0; k < 4; ++k)
    if (nrm[k] == 0.0) return false;
  for (int k = 0; k < 4; ++k) {
    for (int j = 0; j < 3; ++j) T.n[k][j] = T.n[k][j] / nrm[k];
    const double dot = T.n[k][0] * (P[k][0] - T.c[k][0]) + T.n[k][1] * (P[k][1] - T.c[k][1]) +
                       T.n[k][2] * (P[k][2] - T.c[k][2]);
    if (dot > 0)
      for (int j = 0; j < 3; ++j) T.n[k][j] = -T.n[k][j];
  }
  // m = C:e (isotropicstressstrain, elastic3d.f90:129-142), single precision
  const relax_tensor_t& e = t.e;
  const double epskk = (double)e.s11 + (double)e.s22 + (double)e.s33;
  const float tm = (float)(2.0 * mu);
  float m[6] = {tm * e.s11, tm * e.s12, tm * e.s13, tm * e.s22, tm * e.s23, tm * e.s33};
  m[0] = (float)(m[0] + lambda * epskk);
  m[3] = (float)(m[3] + lambda * epskk);
  m[5] = (float)(m[5] + lambda * epskk);
  for (int k = 0; k < 4; ++k) {
    const double* n = T.n[k];
    T.tr[k][0] = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
    T.tr[k][1] = m[1] * n[0] + m[3] * n[1] + m[4] * n[2];
    T.tr[k][2] = m[2] * n[0] + m[4] * n[1] + m[5] * n[2];
  }
  const float ef[6] = {e.s11, e.s12, e.s13, e.s22, e.s23, e.s33};
  for (int k = 0; k < 6; ++k) T.e[k] = ef[k];
  return true;
}

__device__ __forceinline__ double d_gaussi(double x, double sigma) {
  return (1.0 + erf(x / sqrt(2.0) / sigma)) / 2.0;
}

// signed distances of x to the four faces, (centroid - x).n
__device__ __forceinline__ void tetra_dist(const TetraDev& T, double x1, double x2, double x3, double dist[4]) {
  for (int k = 0; k < 4; ++k)
    dist[k] = (T.c[k][0] - x1) * T.n[k][0] + (T.c[k][1] - x2) * T.n[k][1] + (T.c[k][2] - x3) * T.n[k][2];
}

__global__ __launch_bounds__(256) void k_eigen_source(const CuboidDev* __restrict__ C, int nc,
                                                      const TetraDev* __restrict__ T, int nt, int n1, int n2, int h3,
                                                      double dx1, double dx2, double dx3, float* f, long ld,
                                                      long fs) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2 * h3) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = (double)i3 * dx3;
  const long of = i1 + ld * (i2 + (long)n2 * i3);
  float f1 = 0.f, f2 = 0.f, f3 = 0.f;
  bool any = false;
  for (int k = 0; k < nc; ++k) {
    const CuboidDev& c = C[k];
    if ((fabs(x1 - c.x) > c.lim) || (fabs(x2 - c.y) > c.lim) || (fabs(x3 - c.z) > c.lim)) continue;
    if (!any) { f1 = f[of]; f2 = f[fs + of]; f3 = f[2 * fs + of]; any = true; }
    const double cs = c.cs, ss = c.ss, cd = c.cd, sd = c.sd, b = c.beta;
    const double x2r = cs * x1 - ss * x2;
    const double x1s = cd * x2r - sd * x3, x1i = cd * x2r + sd * x3;
    const double x2s = ss * x1 + cs * x2, x3s = sd * x2r + cd * x3, x3i = -sd * x2r + cd * x3;
    double t1 = d_omega((x1s - c.xr) / c.T, b);
    const double t2 = d_omega((x2s - c.yr) / c.W, b);
    double t3 = d_omega((x3s - c.zr) / c.L, b);
    const double sourc = 1.0 / c.T * d_omegap((x1s - c.xr) / c.T, b) * t2 * t3;
    const double dblcp = 1.0 / c.W * t1 * d_omegap((x2s - c.yr) / c.W, b) * t3;
    const double dipcs = 1.0 / c.L * t1 * t2 * d_omegap((x3s - c.zr) / c.L, b);
    t1 = d_omega((x1i - c.xr) / c.T, b);
    t3 = d_omega((x3i + c.zr) / c.L, b);
    const double image = 1.0 / c.T * d_omegap((x1i - c.xr) / c.T, b) * t2 * t3;
    const double cplei = 1.0 / c.W * t1 * d_omegap((x2s - c.yr) / c.W, b) * t3;
    if (2.01 * RELAX_DEG2RAD > c.dip) {
      // subvertical: the method of images
      f1 = (float)((double)f1 + c.m12 * ss * (sourc + image) + c.m12 * cd * cs * (dblcp + cplei) -
                   c.le11 * ss * (dblcp + cplei) - c.le22 * cd * cs * (sourc + image));
      f2 = (float)((double)f2 + c.m12 * cs * (sourc + image) - c.m12 * cd * ss * (dblcp + cplei) -
                   c.le11 * cs * (dblcp + cplei) + c.le22 * cd * ss * (sourc + image));
      f3 = (float)((double)f3 - c.m12 * sd * (dblcp - cplei) + c.le22 * sd * (sourc - image));
    } else {
      f1 = (float)((double)f1 + c.m12 * ss * sourc + c.m12 * cd * cs * dblcp - c.le11 * ss * dblcp -
                   c.le22 * cd * cs * sourc);
      f2 = (float)((double)f2 + c.m12 * cs * sourc - c.m12 * cd * ss * dblcp - c.le11 * cs * dblcp +
                   c.le22 * cd * ss * sourc);
      f3 = (float)((double)f3 - c.m12 * sd * dblcp + c.le22 * sd * sourc);
    }
    // the vertical shear components, as written at elastic3d.f90:1851-1872
    f1 = (float)((double)f1 + c.m13 * cd * ss * dipcs + c.m13 * sd * cs * dblcp - c.m23 * cd * cs * dipcs -
                 c.m23 * sd * cs * sourc - c.le33 * sd * dipcs * cs);
    f2 = (float)((double)f2 + c.m13 * cd * cs * dipcs - c.m13 * sd * ss * dblcp + c.m23 * cd * ss * dipcs +
                 c.m23 * sd * ss * sourc + c.le33 * sd * dipcs * ss);
    f3 = (float)((double)f3 + c.m13 * cd * dblcp - c.m23 * cd * sourc - c.le33 * cd * dipcs);
  }
  for (int k = 0; k < nt; ++k) {
    const TetraDev& t = T[k];
    double dist[4];
    tetra_dist(t, x1, x2, x3, dist);
    if (fabs(dist[0]) > 7.0 * dx1 && fabs(dist[1]) > 7.0 * dx1 && fabs(dist[2]) > 7.0 * dx1 &&
        fabs(dist[3]) > 7.0 * dx1)
      continue;
    if (!any) { f1 = f[of]; f2 = f[fs + of]; f3 = f[2 * fs + of]; any = true; }
    double H[4], dl[4];
    for (int j = 0; j < 4; ++j) {
      H[j] = d_gaussi(dist[j], dx1);
      dl[j] = d_gauss(dist[j], dx1);
    }
    const double w0 = dl[0] * H[1] * H[2] * H[3], w1 = H[0] * dl[1] * H[2] * H[3], w2 = H[0] * H[1] * dl[2] * H[3],
                 w3 = H[0] * H[1] * H[2] * dl[3];
    f1 = (float)((double)f1 + t.tr[0][0] * w0 + t.tr[1][0] * w1 + t.tr[2][0] * w2 + t.tr[3][0] * w3);
    f2 = (float)((double)f2 + t.tr[0][1] * w0 + t.tr[1][1] * w1 + t.tr[2][1] * w2 + t.tr[3][1] * w3);
    f3 = (float)((double)f3 + t.tr[0][2] * w0 + t.tr[1][2] * w1 + t.tr[2][2] * w2 + t.tr[3][2] * w3);
  }
  if (any) { f[of] = f1; f[fs + of] = f2; f[2 * fs + of] = f3; }
}

// (q + ekk lambda I) + 2 mu E, rounded after each operation
__device__ __forceinline__ void add_hooke(float q[6], const float E[6], float lam, float rmu) {
  const float ekk = (float)((double)E[0] + (double)E[3] + (double)E[5]);
  const float le = ekk * lam;
  q[0] = q[0] + le; q[3] = q[3] + le; q[5] = q[5] + le;
  for (int c = 0; c < 6; ++c) q[c] = q[c] + rmu * E[c];
}

__global__ __launch_bounds__(256) void k_add_eigen(const CuboidDev* __restrict__ C, int nc,
                                                   const TetraDev* __restrict__ T, int nt, float lam, float rmu,
                                                   int n1, int n2, int h3, double dx1, double dx2, double dx3,
                                                   float* tau, long H) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2 * h3) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = (double)i3 * dx3;
  float q[6];
  bool any = false;
  for (int k = 0; k < nc; ++k) {
    const CuboidDev& c = C[k];
    if ((fabs(x1 - c.x) > c.lim) || (fabs(x2 - c.y) > c.lim) || (fabs(x3 - c.z) > c.lim)) continue;
    if (!any) {
      for (int j = 0; j < 6; ++j) q[j] = tau[j * H + o];
      any = true;
    }
    const double x2r = c.cs * x1 - c.ss * x2;
    const double x1s = c.cd * x2r - c.sd * x3, x2s = c.ss * x1 + c.cs * x2, x3s = c.sd * x2r + c.cd * x3;
    // the image aperture is formed but not added (elastic3d.f90:2993-2995)
    const float ap = (float)(d_omega((x1s - c.xr) / c.T, c.beta) * d_omega((x2s - c.yr) / c.W, c.beta) *
                             d_omega((x3s - c.zr) / c.L, c.beta));
    float ea[6], E[6];
    for (int j = 0; j < 6; ++j) ea[j] = ap * c.e[j];
    const float a12 = 2.f * ea[1], a13 = 2.f * ea[2], a23 = 2.f * ea[4];
    for (int j = 0; j < 6; ++j) {
      float v = ea[0] * c.SS[j];
      v = v + a12 * c.NS[j];
      v = v + a13 * c.SD[j];
      v = v + ea[3] * c.NN[j];
      v = v + a23 * c.ND[j];
      E[j] = v + ea[5] * c.DD[j];
    }
    add_hooke(q, E, lam, rmu);
  }
  for (int k = 0; k < nt; ++k) {
    const TetraDev& t = T[k];
    double dist[4];
    tetra_dist(t, x1, x2, x3, dist);
    // the reference skips the points within 7 dx1 of the inner side of
    // every face (elastic3d.f90:2838-2841), as written
    if (dist[0] < 7.0 * dx1 && dist[1] < 7.0 * dx1 && dist[2] < 7.0 * dx1 && dist[3] < 7.0 * dx1) continue;
    if (!any) {
      for (int j = 0; j < 6; ++j) q[j] = tau[j * H + o];
      any = true;
    }
    const float h = (float)(d_gaussi(dist[0], dx1) * d_gaussi(dist[1], dx1) * d_gaussi(dist[2], dx1) *
                            d_gaussi(dist[3], dx1));
    float E[6];
    for (int j = 0; j < 6; ++j) E[j] = h * t.e[j];
    add_hooke(q, E, lam, rmu);
  }
  if (any)
    for (int j = 0; j < 6; ++j) tau[j * H + o] = q[j];
}

}  // namespace

int eigen_stage(relax_model_t* m, const relax_event_t* ev) {
  m->cur_ncub = m->cur_ntet = 0;
  const int nc = ev->nc, nth = ev->nth;
  if (nc + nth == 0) return RELAX_OK;
  std::vector<CuboidDev> cub;
  std::vector<TetraDev> tet;
  const double lam = m->desc.lambda, mu = m->desc.mu, beta = m->desc.beta;
  for (int k = 0; k < nc; ++k) cub.push_back(make_cuboid(ev->c[k], lam, mu, beta));
  for (int k = 0; k < nth; ++k) {
    const relax_tetrahedron_t& t = ev->th[k];
    const int idx[4] = {t.i1, t.i2, t.i3, t.i4};
    for (int j = 0; j < 4; ++j)
      if (idx[j] < 1 || idx[j] > ev->nv) return RELAX_ERR_ARG;
    TetraDev T{};
    if (make_tetra(t, ev->v, lam, mu, T)) tet.push_back(T);
  }
  const size_t bytes = cub.size() * sizeof(CuboidDev) + tet.size() * sizeof(TetraDev);
  if (bytes > m->eig_cap) {
    if (m->eig_dev) (void)rx_free(m->eig_dev);
    m->eig_dev = nullptr;
    m->eig_cap = 0;
    HIP_TRY(rx_malloc(&m->eig_dev, bytes));
    m->eig_cap = bytes;
  }
  char* p = static_cast<char*>(m->eig_dev);
  if (!cub.empty())
    HIP_TRY(hipMemcpyAsync(p, cub.data(), cub.size() * sizeof(CuboidDev), hipMemcpyHostToDevice, m->stream));
  if (!tet.empty())
    HIP_TRY(hipMemcpyAsync(p + cub.size() * sizeof(CuboidDev), tet.data(), tet.size() * sizeof(TetraDev),
                           hipMemcpyHostToDevice, m->stream));
  // the host vectors go out of scope: wait for the copies
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->cur_ncub = (int)cub.size();
  m->cur_ntet = (int)tet.size();
  return RELAX_OK;
}

int eigen_body_force(relax_model_t* m, float* f, long ld, long fs) {
  if (m->cur_ncub + m->cur_ntet == 0) return RELAX_OK;
  const Grid& g = m->g;
  const CuboidDev* C = static_cast<const CuboidDev*>(m->eig_dev);
  const TetraDev* T = reinterpret_cast<const TetraDev*>(static_cast<const char*>(m->eig_dev) +
                                                         m->cur_ncub * sizeof(CuboidDev));
  const long N = (long)g.n1 * g.n2 * g.h3;
  k_eigen_source<<<(unsigned)((N + 255) / 256), 256, 0, m->stream>>>(
      C, m->cur_ncub, T, m->cur_ntet, g.n1, g.n2, g.h3, m->desc.dx1, m->desc.dx2, m->desc.dx3, f, ld, fs);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int eigen_moment(relax_model_t* m) {
  if (m->cur_ncub + m->cur_ntet == 0) return RELAX_OK;
  const Grid& g = m->g;
  const CuboidDev* C = static_cast<const CuboidDev*>(m->eig_dev);
  const TetraDev* T = reinterpret_cast<const TetraDev*>(static_cast<const char*>(m->eig_dev) +
                                                         m->cur_ncub * sizeof(CuboidDev));
  const long N = (long)g.n1 * g.n2 * g.h3;
  k_add_eigen<<<(unsigned)((N + 255) / 256), 256, 0, m->stream>>>(
      C, m->cur_ncub, T, m->cur_ntet, (float)m->desc.lambda, 2.0f * (float)m->desc.mu, g.n1, g.n2, g.h3,
      m->desc.dx1, m->desc.dx2, m->desc.dx3, m->tau, g.H);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
