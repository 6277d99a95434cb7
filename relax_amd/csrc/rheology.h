// rheology.h -- pointwise eigenstress rates, evaluated inside the stencil
// kernels so the moment-density field never round-trips through HBM.
//   viscous power law: viscouseigenstress, viscoelastic3d.f90:217-314
//   fault creep:       frictioneigenstress, friction3d.f90:82-241 (CPU path)
//   transient creep:   transienteigenstress, viscoelastic3d.f90:123-202
//   weak zones:        the dgammadot0 fields of builddgammadot0, :320-435
#pragma once
#include "common.h"

struct RheoParams {
  float mu;
  Mechanism mech[2];  // 0: linear (in%linearstruc), 1: power law (in%nonlinearstruc)
  // transient (Burgers) creep: 0: in%ltransientstruc, 1: in%nltransientstruc;
  // their cohesion table holds Gk
  Mechanism tr[2];
  Prestress pre;
  FrictionTables fr;
  int n1, n2;
  double dx1, dx2, dx3;
  // inter-seismic loading: the moment density of the loading segments,
  // six planes of H (SoA), or NULL.  The corrector adds it to the moment of
  // the rheology and, unscaled, to tau (relax.f90:795-799, :1199-1206).
  const float* inter;
};

// tensordeviatoric, elastic3d.f90:275-285
__device__ __forceinline__ T6 deviatoric(const T6& t) {
  const float diag = (t.s11 + t.s22 + t.s33) / 3.0f;
  return T6{t.s11 - diag, t.s12, t.s13, t.s22 - diag, t.s23, t.s33 - diag};
}
// tensornorm, elastic3d.f90:311-320: sqrt((sum of squares, off-diagonals twice)/2)
__device__ __forceinline__ float tnorm(const T6& t) {
  return sqrtf((t.s11 * t.s11 + 2.f * t.s12 * t.s12 + 2.f * t.s13 * t.s13 + t.s22 * t.s22 +
                2.f * t.s23 * t.s23 + t.s33 * t.s33) * 0.5f);
}
__device__ __forceinline__ double tnorm_d(const T6& t) {
  return sqrt(((double)t.s11 * t.s11 + 2.0 * t.s12 * t.s12 + 2.0 * t.s13 * t.s13 + (double)t.s22 * t.s22 +
               2.0 * t.s23 * t.s23 + (double)t.s33 * t.s33) * 0.5);
}

// gauss, omega, my_sinh: elastic3d.f90:525-531, :549-562, :89-97
__device__ __forceinline__ double d_gauss(double x, double sigma) {
  const double r = x / sigma;
  return exp(-0.5 * r * r) / sqrt(RELAX_PI2) / sigma;
}
__device__ __forceinline__ double d_omega(double x, double beta) {
  const double ax = fabs(x);
  if (ax <= (1.0 - 2.0 * beta) / (1.0 - beta) / 2.0) return 1.0;
  if (ax < 1.0 / (1 - beta) / 2.0) {
    const double c = cos(RELAX_PI * ((1.0 - beta) * ax - 0.5 + beta) / 2.0 / beta);
    return c * c;
  }
  return 0.0;
}
// the derivatives gaussp and omegap, elastic3d.f90:537-542, :570-581
__device__ __forceinline__ double d_gaussp(double x, double sigma) {
  const double r = x / sigma;
  return -x * exp(-0.5 * r * r) / sqrt(RELAX_PI2) / (sigma * sigma * sigma);
}
__device__ __forceinline__ double d_omegap(double x, double beta) {
  const double ax = fabs(x);
  if (ax > (1.0 - 2.0 * beta) / (1.0 - beta) / 2.0 && ax < 1.0 / (1 - beta) / 2.0)
    return -copysign(1.0, x) * RELAX_PI * (1.0 - beta) / 2.0 / beta *
           sin(RELAX_PI * ((1.0 - beta) * ax - 0.5 + beta) / beta);
  return 0.0;
}
__device__ __forceinline__ double d_my_sinh(double x) {
  if (fabs(x) > 11.0) return copysign(sinh(11.0), x);
  return sinh(x);
}
// shiftedcoordinates, elastic3d.f90:4027-4048 (0-based index)
__device__ __forceinline__ double d_shifted(int i, int n, double dx) {
  return (double)(i < n / 2 ? i : i - n) * dx;
}

// x^e for the stress exponents of the power law: exact forms for e = n-1 in
// {0, 1, 2, 2.5} (n = 1, 2, 3, 3.5), otherwise exp2(e log2 x) on the
// hardware transcendentals (relative error ~ |e log2 x| 2^-23).
__device__ __forceinline__ float pow_exponent(float x, float e) {
  if (e == 0.f) return 1.f;
  if (e == 1.f) return x;
  if (e == 2.f) return x * x;
  if (e == 2.5f) return x * x * sqrtf(x);
  if (e == 3.f) return x * x * x;
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}

// Viscous contribution of one mechanism; returns 1/gammadot (or +inf when
// the point is skipped) and adds gammadot*tauc/mu to *grate.  g0 is the
// fluidity gammadot0 of the point and gdp its prestress rate
// gammadot0 (taup/mu)^(n-1); without weak zones both depend on depth only
// and come from tables.
__device__ __forceinline__ float viscous_point(const Mechanism& M, const Prestress& P, float mu, int i3,
                                               float g0, float gdp, const T6& sd, T6& m, float* grate) {
  if (1.0e-20 > (double)g0) return __builtin_inff();
  const float* sp = P.sp + 6 * i3;
  const T6 ss{sd.s11 + sp[0], sd.s12 + sp[1], sd.s13 + sp[2], sd.s22 + sp[3], sd.s23 + sp[4], sd.s33 + sp[5]};
  const float tau = tnorm(ss);
  const float tauc = fmaxf(0.f, tau - M.cohesion[i3]);
  if ((double)tauc <= 1.0e-20) return __builtin_inff();
  const float ex = M.power[i3] - 1.f;
  const float gd = g0 * pow_exponent(tauc / mu, ex);
  const float a = 2.f * gd, b = 2.f * gdp;
  m.s11 += a * sd.s11 - b * sp[0];
  m.s12 += a * sd.s12 - b * sp[1];
  m.s13 += a * sd.s13 - b * sp[2];
  m.s22 += a * sd.s22 - b * sp[3];
  m.s23 += a * sd.s23 - b * sp[4];
  m.s33 += a * sd.s33 - b * sp[5];
  if (grate) *grate = gd * tauc / mu;
  return 1.f / gd;
}

// Fault-creep contribution of one plane at grid point (i1,i2,i3).
// Returns the candidate relaxation time (NaN-free: a NaN candidate becomes
// +inf, as MIN keeps the running value) and sets *status on overflow.
__device__ __forceinline__ float friction_point(const RheoParams& R, const PlaneDev& P, int i1, int i2, int i3,
                                                const T6& s, T6& m, float* vel, int* status) {
  const double dx1 = R.dx1;
  const double x3 = (double)i3 * R.dx3;
  if ((fabs(x3 - P.z) > P.Lp) && (fabs(x3 + P.z) > P.Lp)) return __builtin_inff();
  const double x1 = d_shifted(i1, R.n1, R.dx1), x2 = d_shifted(i2, R.n2, R.dx2);
  const double wl = fmax(P.Wp, P.Lp);
  if ((fabs(x1 - P.x) > wl) || (fabs(x2 - P.y) > wl)) return __builtin_inff();
  const double x2r = P.cstrike * x1 - P.sstrike * x2;
  const double x1s = P.cdip * x2r - P.sdip * x3;
  const double x1i = P.cdip * x2r + P.sdip * x3;
  if ((fabs(x1s - P.xr) > 7.01 * dx1) && (fabs(x1i - P.xr) > 7.01 * dx1)) return __builtin_inff();
  const double x2s = P.sstrike * x1 + P.cstrike * x2;
  const double x3s = P.sdip * x2r + P.cdip * x3;
  const double x3i = -P.sdip * x2r + P.cdip * x3;
  const double temp2 = d_omega((x2s - P.yr) / P.W, P.beta);
  const double sourc = d_gauss(x1s - P.xr, dx1) * temp2 * d_omega((x3s - P.zr) / P.L, P.beta);
  const double image = d_gauss(x1i - P.xr, dx1) * temp2 * d_omega((x3i + P.zr) / P.L, P.beta);
  const double impulse = sourc + image;
  const double* n = P.n;
  const double t0 = s.s11 * n[0] + s.s12 * n[1] + s.s13 * n[2];
  const double t1 = s.s12 * n[0] + s.s22 * n[1] + s.s23 * n[2];
  const double t2 = s.s13 * n[0] + s.s23 * n[1] + s.s33 * n[2];
  const double taun = t0 * n[0] + t1 * n[1] + t2 * n[2];
  double ts0 = t0 - taun * n[0], ts1 = t1 - taun * n[1], ts2 = t2 - taun * n[2];
  const double taus = sqrt(ts0 * ts0 + ts1 * ts1 + ts2 * ts2);
  const double fric = R.fr.friction[i3], coh = R.fr.cohesion[i3];
  const double tau = fmax(0.0, taus + fric * taun - coh);
  if ((ts0 * P.r[0] + ts1 * P.r[1] + ts2 * P.r[2]) < 0.0 && fabs(P.rake) < RELAX_PI2 * 1.5)
    return __builtin_inff();
  const double tauc = R.fr.tauc[i3];
  if ((tau / tauc) > 20) { *status = RELAX_ERR_FRICTION; return __builtin_inff(); }
  ts0 /= taus; ts1 /= taus; ts2 /= taus;
  const double gd = R.fr.v0[i3] * 2.0 * d_my_sinh(tau / tauc);
  float cand = (float)(tau / (double)R.mu / gd * P.minlw);
  if (cand != cand) cand = __builtin_inff();
  if (vel) *vel = (float)((double)*vel + gd * impulse * sqrt(RELAX_PI2) * dx1);
  const double a = 2.0 * (double)R.mu * impulse * gd;
  const double b0 = a * n[0], b1 = a * n[1], b2 = a * n[2];
  m.s11 += (float)(ts0 * b0);
  m.s12 += (float)((ts0 * b1 + ts1 * b0) / 2.0);
  m.s13 += (float)((ts0 * b2 + ts2 * b0) / 2.0);
  m.s22 += (float)(ts1 * b1);
  m.s23 += (float)((ts1 * b2 + ts2 * b1) / 2.0);
  m.s33 += (float)(ts2 * b2);
  return cand;
}

// Moment-density rate of all mechanisms at one point (relax.f90:568-612).
// tmin[3] (optional) receives the per-mechanism minima of the relaxation
// time; grate (optional) the viscous strain rate summed over mechanisms.
template <bool TM, bool FRIC = true>
__device__ __forceinline__ T6 rheology(const RheoParams& R, const T6& s, int i1, int i2, int i3, float* tmin,
                                       float* grate_lin, float* grate_nl, int* status) {
  T6 m{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (R.mech[0].active || R.mech[1].active) {
    const T6 sd = deviatoric(s);
    if (R.mech[0].active) {
      const Mechanism& M = R.mech[0];
      const float t = viscous_point(M, R.pre, R.mu, i3, M.gammadot0[i3], M.gdp[i3], sd, m, grate_lin);
      if (TM) tmin[0] = fminf(tmin[0], t);
    }
    if (R.mech[1].active) {
      const Mechanism& M = R.mech[1];
      const float t = viscous_point(M, R.pre, R.mu, i3, M.gammadot0[i3], M.gdp[i3], sd, m, grate_nl);
      if (TM) tmin[1] = fminf(tmin[1], t);
    }
  }
  if (FRIC && R.fr.active) {
    for (int k = 0; k < R.fr.np; ++k) {
      const float t = friction_point(R, R.fr.planes[k], i1, i2, i3, s, m, nullptr, status);
      if (TM) tmin[2] = fminf(tmin[2], t);
    }
  }
  return m;
}

// Transient contribution of one mechanism at a point of fluidity g0
// (transienteigenstress, viscoelastic3d.f90:170-196): Q = s' - 2 Gk E,
// gammadot = g0 (|Q|/mu)^(n-1), Edot = gammadot/mu Q, moment += 2 gammadot Q.
// Returns 1/gammadot; a point with g0 < 1e-20 gets Edot = 0 and +inf.
__device__ __forceinline__ float transient_point(const Mechanism& M, float mu, int i3, float g0, const T6& sd,
                                                 const T6& e, T6& m, T6& ed) {
  if (1.0e-20 > (double)g0) {
    ed = T6{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    return __builtin_inff();
  }
  const float k2 = 2.f * M.cohesion[i3];  // REAL(2 Gk)
  const T6 q{sd.s11 - k2 * e.s11, sd.s12 - k2 * e.s12, sd.s13 - k2 * e.s13,
             sd.s22 - k2 * e.s22, sd.s23 - k2 * e.s23, sd.s33 - k2 * e.s33};
  const float nq = tnorm(q);
  const float gd = g0 * pow_exponent(nq / mu, M.power[i3] - 1.f);
  const float a = gd / mu, b = 2.f * gd;
  ed = T6{a * q.s11, a * q.s12, a * q.s13, a * q.s22, a * q.s23, a * q.s33};
  m.s11 += b * q.s11; m.s12 += b * q.s12; m.s13 += b * q.s13;
  m.s22 += b * q.s22; m.s23 += b * q.s23; m.s33 += b * q.s33;
  return 1.f / gd;
}

// Point fluidity of a mechanism: the depth table plus the weak-zone field.
__device__ __forceinline__ float point_fluidity(const Mechanism& M, int i3, long o) {
  return M.dg ? M.gammadot0[i3] + M.dg[o] : M.gammadot0[i3];
}

// Moment-density rate of all mechanisms at point o = (i1,i2,i3), weak zones
// and transient creep included (relax.f90:568-650 predictor, :704-793
// corrector).  e is the transient strain the first transient mechanism
// reads: epsilonik in the predictor, the mid-step strain in the corrector.
// *ed receives the final epsilonikdot.  CORR reproduces the corrector's
// calls, which pass epsilonikdot as both the strain and the rate: the
// second mechanism reads the first one's rate.  In the predictor both read
// epsilonik and the second one's rate replaces the first one's.
template <bool TM, bool CORR>
__device__ __forceinline__ T6 rheology_ext(const RheoParams& R, const T6& s, int i1, int i2, int i3, long o,
                                           const T6& e, T6* ed, float* tmin, float* grate_lin, float* grate_nl,
                                           int* status) {
  T6 m{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const bool trans = R.tr[0].active || R.tr[1].active;
  if (R.mech[0].active || R.mech[1].active || trans) {
    const T6 sd = deviatoric(s);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const Mechanism& M = R.mech[k];
      if (!M.active) continue;
      const float g0 = point_fluidity(M, i3, o);
      const float gdp = M.dg ? g0 * M.rp[i3] : M.gdp[i3];
      const float t = viscous_point(M, R.pre, R.mu, i3, g0, gdp, sd, m, k ? grate_nl : grate_lin);
      if (TM) tmin[k] = fminf(tmin[k], t);
    }
    if (R.fr.active) {
      for (int k = 0; k < R.fr.np; ++k) {
        const float t = friction_point(R, R.fr.planes[k], i1, i2, i3, s, m, nullptr, status);
        if (TM) tmin[2] = fminf(tmin[2], t);
      }
    }
    if (trans) {
      T6 ein = e, r{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const Mechanism& M = R.tr[k];
        if (!M.active) continue;
        const float t = transient_point(M, R.mu, i3, point_fluidity(M, i3, o), sd, ein, m, r);
        if (TM) tmin[3 + k] = fminf(tmin[3 + k], t);
        if (CORR) ein = r;
      }
      *ed = r;
    }
  } else if (R.fr.active) {
    for (int k = 0; k < R.fr.np; ++k) {
      const float t = friction_point(R, R.fr.planes[k], i1, i2, i3, s, m, nullptr, status);
      if (TM) tmin[2] = fminf(tmin[2], t);
    }
  }
  return m;
}
