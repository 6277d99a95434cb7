// model.hip -- device-resident model and the RK2 predictor/corrector step
// (relax.f90:555-890, integrationstep :1420-1452, isoutput :1364-1376).
//
// The whole state (u, tau, sigma, gamma, fault patches, t, Dt) stays in HBM.
// One step is a fixed sequence of launches on the model's stream with no host
// synchronisation: the relaxation time is reduced on the device and the time
// step is computed by a one-thread kernel, so every later kernel reads Dt from
// device memory.
//
//   predictor  k_bodyforce(sig)+tm -> k_finalize_dt -> Green(W) -> v = u + Dt/2 G
//              -> k_stress: sig = -tau - Dt/2 m(sig) + C:eps(v)
//   corrector  k_bodyforce(sig) -> monitorfriction -> Green(W) -> u += Dt G
//              -> k_stress: tau += Dt m(sig), gamma += Dt gdot, sig = -tau + C:eps(u)
//              -> frictionadd -> k_end_step (t, oi, power, C:E^i)
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "kernels.h"
#include "model.h"
#include "stencil.h"

static const double STEP_MAX = 1e7;  // relax.f90:202

// ------------------------------------------------------------- kernels ----
__global__ void k_reset_scalars(StepScalars* sc, double t, int oi, int it) {
  sc->t = t; sc->Dt = 0; sc->tm = 0;
  sc->maxwell[0] = sc->maxwell[1] = sc->maxwell[2] = STEP_MAX;
  sc->power = sc->cei = 0;
  sc->oi = oi; sc->it = it; sc->status = 0;
  for (int k = 0; k < 3; ++k) sc->tmin_bits[k] = 0x7f800000u;
}

struct TimeLine {
  double interval, odt, tscale, next_event, etime;
  int skip, has_next;
  int mech[3];
};

__global__ void k_finalize_dt(StepScalars* sc, TimeLine tl, int use_override, double dt_override) {
  double maxwell[3];
  for (int k = 0; k < 3; ++k) {
    const float v = fminf(__uint_as_float(sc->tmin_bits[k]), (float)STEP_MAX);
    maxwell[k] = fmin((double)v, STEP_MAX);
    if (!tl.mech[k]) maxwell[k] = STEP_MAX;
  }
  double tm = 1.0 / ((double)tl.mech[0] / maxwell[0] + (double)tl.mech[1] / maxwell[1] +
                     (double)tl.mech[2] / maxwell[2]);
  tm = fmin(tm, STEP_MAX);
  const double t = sc->t;
  double Dt = tm / 10.0;
  if (tl.skip == 0) {
    if ((t + Dt) >= ((double)sc->oi * tl.odt) - Dt * 0.04) Dt = (double)sc->oi * tl.odt - t;
  } else {
    Dt = Dt * tl.tscale;
    if (tl.has_next && (t + Dt - tl.next_event) >= 0.0) Dt = tl.next_event - t;
  }
  if (use_override) Dt = dt_override;
  sc->Dt = Dt;
  sc->tm = tm;
  for (int k = 0; k < 3; ++k) sc->maxwell[k] = maxwell[k];
  sc->it = sc->it + 1;
}

// First level of the amplitude reduction: workgroup g sums the stress
// kernel's partials [g*RED_CHUNK, (g+1)*RED_CHUNK) in a fixed order.
static const int RED_CHUNK = 2048;
__global__ __launch_bounds__(256) void k_reduce_partials(const double2* __restrict__ partials, long nb,
                                                         double2* __restrict__ out) {
  __shared__ double2 red[4];
  const int tid = threadIdx.x;
  const long b0 = (long)blockIdx.x * RED_CHUNK;
  double2 a = make_double2(0.0, 0.0);
  for (int j = 0; j < RED_CHUNK / 256; ++j) {
    const long i = b0 + tid + 256L * j;
    if (i < nb) { a.x += partials[i].x; a.y += partials[i].y; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    a.x += __shfl_xor(a.x, off);
    a.y += __shfl_xor(a.y, off);
  }
  if ((tid & 63) == 0) red[tid >> 6] = a;
  __syncthreads();
  if (tid == 0) {
    double2 r = red[0];
    for (int w = 1; w < 4; ++w) { r.x += red[w].x; r.y += red[w].y; }
    out[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void k_end_step(StepScalars* sc, const double2* partials, long nb, double dv,
                                                  TimeLine tl) {
  __shared__ double2 red[256];
  const int tid = threadIdx.x;
  double2 a = make_double2(0.0, 0.0);
  for (long i = tid; i < nb; i += 256) { a.x += partials[i].x; a.y += partials[i].y; }
  red[tid] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) { red[tid].x += red[tid + s].x; red[tid].y += red[tid + s].y; }
    __syncthreads();
  }
  if (tid == 0) {
    sc->power = red[0].x * dv;
    sc->cei = red[0].y * dv;
    const double t = sc->t + sc->Dt;
    sc->t = t;
    const int it = sc->it, oi = sc->oi;
    const bool out = ((0 == tl.skip) && (fabs(t - oi * tl.odt) < 1e-6 * tl.odt)) ||
                     ((0 < tl.skip) && ((it - 1) % tl.skip == 0)) || (fabs(t - tl.etime) < 1e-6);
    if (out) sc->oi = oi + 1;
    for (int k = 0; k < 3; ++k) sc->tmin_bits[k] = 0x7f800000u;
  }
}

// fix, elastic3d.f90:70-83
__device__ __forceinline__ int d_fix(double number) {
  const double f = floor(number), c = ceil(number);
  return (number - f > 0.5) ? (int)c : (int)f;
}

// monitorfriction, friction3d.f90:263-416: one thread per patch
__global__ void k_monitorfriction(PlaneDev P, double rake, int px2, int px3, Grid g, double dx1, double dx2,
                                  double dx3, int sx3_range, const float* __restrict__ sig,
                                  const double* __restrict__ v0, const double* __restrict__ tauc_t,
                                  const double* __restrict__ fric, const double* __restrict__ coh,
                                  relax_slippatch_t* patch, StepScalars* sc) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= px2 * px3) return;
  const int j2 = idx % px2 + 1, j3 = idx / px2 + 1;
  relax_slippatch_t& Q = patch[idx];
  const double cs = P.cstrike, ss = P.sstrike, cd = P.cdip, sd = P.sdip;
  const double sv[3] = {ss, cs, 0.0};
  const double dv[3] = {-cs * sd, +ss * sd, -cd};
  // ref2local
  const double x2 = cs * P.x - ss * P.y;
  const double xr = cd * x2 - sd * P.z;
  double yr = ss * P.x + cs * P.y;
  double zr = sd * x2 + cd * P.z;
  yr = (double)((float)yr + (float)(((double)j2 - (double)px2 / 2.0 - 1.0) * dx2));
  zr = (double)((float)zr + (float)(((double)j3 - (double)px3 / 2.0 - 1.0) * dx3));
  // local2ref
  const double x2p = cd * xr + sd * zr;
  const double X1 = cs * x2p + ss * yr, X2 = -ss * x2p + cs * yr, X3 = -sd * xr + cd * zr;
  Q.x1 = X1; Q.x2 = X2; Q.x3 = X3; Q.lx = yr; Q.lz = zr;
  const int n1 = g.n1, n2 = g.n2;
  if ((X1 > (double)(n1 / 2 - 1) * dx1) || (X1 < -(double)(n1 / 2) * dx1) || (X2 > (double)(n2 / 2 - 1) * dx2) ||
      (X2 < -(double)(n2 / 2) * dx2) || (X3 > (double)(sx3_range - 1) * dx3) || (X3 < 0.0))
    return;
  const int i1 = (n1 + d_fix(X1 / dx1)) % n1, i2 = (n2 + d_fix(X2 / dx2)) % n2, i3 = d_fix(X3 / dx3);
  if (i3 >= g.h3) { atomicMax(&sc->status, RELAX_ERR_RANGE); return; }
  const long o = i1 + (long)n1 * (i2 + (long)n2 * i3);
  const T6 s = load_t6(sig, g.H, o);
  Q.sig.s11 = s.s11; Q.sig.s12 = s.s12; Q.sig.s13 = s.s13; Q.sig.s22 = s.s22; Q.sig.s23 = s.s23; Q.sig.s33 = s.s33;
  const double* n = P.n;
  const double t0 = s.s11 * n[0] + s.s12 * n[1] + s.s13 * n[2];
  const double t1 = s.s12 * n[0] + s.s22 * n[1] + s.s23 * n[2];
  const double t2 = s.s13 * n[0] + s.s23 * n[1] + s.s33 * n[2];
  const double taun = t0 * n[0] + t1 * n[1] + t2 * n[2];
  double ts0 = t0 - taun * n[0], ts1 = t1 - taun * n[1], ts2 = t2 - taun * n[2];
  const double taus = sqrt(ts0 * ts0 + ts1 * ts1 + ts2 * ts2);
  const double tau = fmax(0.0, taus + fric[i3] * taun - coh[i3]);
  if ((ts0 * P.r[0] + ts1 * P.r[1] + ts2 * P.r[2]) < 0.0 && fabs(rake) < RELAX_PI2 * 1.5) return;
  Q.taus = taus;
  const double v = v0[i3] * 2.0 * d_my_sinh(tau / tauc_t[i3]);
  Q.v = v;
  ts0 /= taus; ts1 /= taus; ts2 /= taus;
  Q.vss = v * (ts0 * sv[0] + ts1 * sv[1] + ts2 * sv[2]);
  Q.vds = v * (ts0 * dv[0] + ts1 * dv[1] + ts2 * dv[2]);
}

// frictionadd, friction3d.f90:424-453
__global__ void k_frictionadd(relax_slippatch_t* patch, long n, const StepScalars* sc, int use_dt, double dt) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double d = use_dt ? dt : sc->Dt;
  relax_slippatch_t& P = patch[k];
  P.ss = P.ss + d * P.vss;
  P.ds = P.ds + d * P.vds;
  P.slip = sqrt(P.ds * P.ds + P.ss * P.ss);
}
int launch_frictionadd(relax_slippatch_t* patch, long n, double dt, hipStream_t st) {
  if (n <= 0) return RELAX_OK;
  k_frictionadd<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(patch, n, nullptr, 1, dt);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
int launch_monitorfriction(const PlaneDev& P, double rake, int px2, int px3, const Grid& g, double dx1, double dx2,
                           double dx3, int sx3_range, const float* sig, const double* v0, const double* tauc,
                           const double* fric, const double* coh, relax_slippatch_t* patch, StepScalars* sc,
                           hipStream_t st) {
  const int n = px2 * px3;
  if (n <= 0) return RELAX_OK;
  k_monitorfriction<<<(n + 127) / 128, 128, 0, st>>>(P, rake, px2, px3, g, dx1, dx2, dx3, sx3_range, sig, v0, tauc,
                                                     fric, coh, patch, sc);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// AoS TENSOR <-> SoA planes
__global__ void k_aos_to_soa(const float* __restrict__ aos, float* __restrict__ soa, long H) {
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < H; o += (long)gridDim.x * blockDim.x)
    for (int c = 0; c < 6; ++c) soa[c * H + o] = aos[6 * o + c];
}
__global__ void k_soa_to_aos(const float* __restrict__ soa, float* __restrict__ aos, long H) {
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < H; o += (long)gridDim.x * blockDim.x)
    for (int c = 0; c < 6; ++c) aos[6 * o + c] = soa[c * H + o];
}
static dim3 grid_1d(long n) {
  long b = (n + 255) / 256;
  return dim3((unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b)));
}
int launch_aos_soa(const float* aos, float* soa, long H, bool to_soa, hipStream_t st) {
  if (to_soa) k_aos_to_soa<<<grid_1d(H), 256, 0, st>>>(aos, soa, H);
  else k_soa_to_aos<<<grid_1d(H), 256, 0, st>>>(soa, const_cast<float*>(aos), H);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// ------------------------------------------------------------ host side ----
static bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

size_t relax_model_footprint(const relax_model_desc_t* d) {
  if (!d) return 0;
  const size_t n1 = d->sx1, n2 = d->sx2, h3 = d->sx3 / 2, H = n1 * n2 * h3;
  size_t b = 0;
  b += 19 * H * 4;                  // u 3, sig 6, tau 6, gamma 1, work 3
  b += 3 * n2 * h3 * 8;             // Nyquist planes of the work spectrum
  b += 3 * n1 * n2 * 4 + 3 * n2 * 8;  // traction planes and their Nyquist bins
  b += ((H + 255) / 256 + (H + 255) / 256 / 2048 + 1) * 16;  // amplitude partials, two levels
  b += 19 * h3 * 4 + 4 * h3 * 8 + 4096;  // rheology and fault-creep tables, scalars
  return b;
}

static void free_model(relax_model_t* m) {
  if (!m) return;
  void* ptrs[] = {m->u, m->sig, m->tau, m->gam, m->W, m->NQ, m->T, m->TN, m->sc, m->partials, m->tables, m->ftab,
                  m->planes_dev, m->patches_dev};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (m->host_sc) (void)hipHostFree(m->host_sc);
  for (hipEvent_t e : m->prof_ev) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int relax_model_create(const relax_model_desc_t* d, relax_model_t** out) {
  if (!d || !out) return RELAX_ERR_ARG;
  *out = nullptr;
  if (!pow2_in(d->sx1, 32, 1024) || !pow2_in(d->sx2, 32, 1024) || !pow2_in(d->sx3, 32, 512)) return RELAX_ERR_SIZE;
  if (!(d->dx1 > 0 && d->dx2 > 0 && d->dx3 > 0 && d->mu > 0)) return RELAX_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RELAX_ERR_NODEVICE;
  size_t freeb = 0, totalb = 0;
  HIP_TRY(hipMemGetInfo(&freeb, &totalb));
  if (relax_model_footprint(d) > freeb) return RELAX_ERR_ADMISSION;

  relax_model_t* m = new (std::nothrow) relax_model_t();
  if (!m) return RELAX_ERR_NOMEM;
  HIP_TRY(hipGetDevice(&m->device));
  m->desc = *d;
  m->events.assign(d->event_times, d->event_times + (d->event_times ? d->ne : 0));
  m->desc.event_times = nullptr;
  Grid& g = m->g;
  g.n1 = d->sx1; g.n2 = d->sx2; g.n3 = d->sx3; g.h3 = d->sx3 / 2; g.m1 = d->sx1 / 2;
  g.H = (long)g.n1 * g.n2 * g.h3;
  g.dx1 = (float)d->dx1; g.dx2 = (float)d->dx2; g.dx3 = (float)d->dx3;
  const long H = g.H;
  bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
  auto alloc = [&](void** p, size_t bytes) { if (ok && hipMalloc(p, bytes) != hipSuccess) ok = false; };
  alloc((void**)&m->u, 3 * H * 4);
  alloc((void**)&m->sig, 6 * H * 4);
  alloc((void**)&m->tau, 6 * H * 4);
  alloc((void**)&m->gam, H * 4);
  alloc((void**)&m->W, 3 * H * 4);
  alloc((void**)&m->NQ, 3L * g.n2 * g.h3 * 8);
  alloc((void**)&m->T, 3L * g.n1 * g.n2 * 4);
  alloc((void**)&m->TN, 3L * g.n2 * 8);
  alloc((void**)&m->sc, sizeof(StepScalars));
  m->nblk = stress_blocks(g);
  m->nred = (m->nblk + RED_CHUNK - 1) / RED_CHUNK;
  alloc((void**)&m->partials, (m->nblk + m->nred) * sizeof(double2));
  alloc((void**)&m->tables, 19L * g.h3 * 4);
  alloc((void**)&m->ftab, 4L * g.h3 * 8);
  if (ok && hipHostMalloc((void**)&m->host_sc, sizeof(StepScalars) * 64) != hipSuccess) ok = false;
  m->host_cap = 64;
  if (!ok) { free_model(m); return RELAX_ERR_NOMEM; }
  // zero state; no mechanisms until set_structure
  if (hipMemsetAsync(m->u, 0, 3 * H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->sig, 0, 6 * H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->tau, 0, 6 * H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->gam, 0, H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->tables, 0, 19L * g.h3 * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->ftab, 0, 4L * g.h3 * 8, m->stream) != hipSuccess) {
    free_model(m);
    return RELAX_ERR_HIP;
  }
  k_reset_scalars<<<1, 1, 0, m->stream>>>(m->sc, 0.0, 1, 0);
  // rheology parameter block
  RheoParams& R = m->R;
  std::memset(&R, 0, sizeof(R));
  R.mu = (float)d->mu;
  const long h3 = g.h3;
  for (int k = 0; k < 2; ++k) {
    R.mech[k].gammadot0 = m->tables + (3 * k + 0) * h3;
    R.mech[k].power = m->tables + (3 * k + 1) * h3;
    R.mech[k].cohesion = m->tables + (3 * k + 2) * h3;
    R.mech[k].gdp = m->tables + (17 + k) * h3;
    R.mech[k].active = 0;
  }
  R.fr.v0 = m->ftab;
  R.fr.tauc = m->ftab + h3;
  R.fr.friction = m->ftab + 2 * h3;
  R.fr.cohesion = m->ftab + 3 * h3;
  R.pre.sp = m->tables + 10 * h3;  // 6*h3
  R.pre.taup = m->tables + 16 * h3;
  R.n1 = g.n1; R.n2 = g.n2;
  R.dx1 = d->dx1; R.dx2 = d->dx2; R.dx3 = d->dx3;
  m->S.k1h = (float)(0.5 / d->dx1);
  m->S.k2h = (float)(0.5 / d->dx2);
  m->S.k3h = (float)(0.5 / d->dx3);
  m->S.idx3 = (float)(1.0 / d->dx3);
  m->S.lam = (float)d->lambda;
  m->S.mu2 = (float)(2.0 * d->mu);
  m->gc.lambda = d->lambda; m->gc.mu = d->mu; m->gc.gamma = d->gamma;
  m->gc.dx1 = d->dx1; m->gc.dx2 = d->dx2; m->gc.dx3 = d->dx3;
  m->gc.n1 = g.n1; m->gc.n2 = g.n2; m->gc.n3 = g.n3;
  // twiddle tables are created outside any capture
  if (!twiddle_table(g.n1) || !twiddle_table(g.n2) || !twiddle_table(g.n3)) { free_model(m); return RELAX_ERR_NOMEM; }
  if (hipStreamSynchronize(m->stream) != hipSuccess) { free_model(m); return RELAX_ERR_HIP; }
  *out = m;
  return RELAX_OK;
}

int relax_model_destroy(relax_model_t* m) {
  if (!m) return RELAX_ERR_ARG;
  (void)hipStreamSynchronize(m->stream);
  free_model(m);
  return RELAX_OK;
}

void* relax_model_stream(relax_model_t* m) { return m ? (void*)m->stream : nullptr; }

int relax_model_synchronize(relax_model_t* m) {
  if (!m) return RELAX_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

// prestress rate of each depth, gammadot0 (taup/mu)^(n-1)
// (viscoelastic3d.f90:290), computed once in double
static int update_gdp(relax_model_t* m, int k) {
  const int h3 = m->g.h3;
  if (m->hg0[k].empty()) return RELAX_OK;
  std::vector<float> g(h3);
  for (int i = 0; i < h3; ++i) {
    const double taup = m->htaup.empty() ? 0.0 : m->htaup[i];
    g[i] = (float)(m->hg0[k][i] * std::pow(taup / m->desc.mu, m->hn[k][i] - 1.0));
  }
  HIP_TRY(hipMemcpyAsync(m->tables + (17L + k) * h3, g.data(), h3 * 4, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

int relax_model_set_structure(relax_model_t* m, int which, const relax_layer_t* layers, int n) {
  if (!m || which < 0 || which > 2) return RELAX_ERR_ARG;
  const int h3 = m->g.h3;
  if (!layers) {
    if (which < 2) m->R.mech[which].active = 0;
    else m->R.fr.active = 0;
    m->mech[which] = 0;
    return RELAX_OK;
  }
  // viscouseigenstress returns without effect when the structure does not
  // match the grid (viscoelastic3d.f90:240-241); here that is an error
  if (n != h3) return RELAX_ERR_ARG;
  std::vector<float> a(h3), b(h3), c(h3);
  std::vector<double> fd(4 * h3);
  for (int i = 0; i < h3; ++i) {
    if (which < 2 && layers[i].stressexponent < 0.999999) return RELAX_ERR_POWER;
    a[i] = (float)layers[i].gammadot0;
    b[i] = (float)layers[i].stressexponent;
    c[i] = (float)layers[i].cohesion;
    fd[i] = layers[i].gammadot0;
    fd[h3 + i] = layers[i].stressexponent;
    fd[2 * h3 + i] = layers[i].friction;
    fd[3 * h3 + i] = layers[i].cohesion;
  }
  if (which < 2) {
    float* base = m->tables + 3 * which * (long)h3;
    HIP_TRY(hipMemcpyAsync(base, a.data(), h3 * 4, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemcpyAsync(base + h3, b.data(), h3 * 4, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemcpyAsync(base + 2 * h3, c.data(), h3 * 4, hipMemcpyHostToDevice, m->stream));
    m->hg0[which].assign(h3, 0.0);
    m->hn[which].assign(h3, 1.0);
    for (int i = 0; i < h3; ++i) { m->hg0[which][i] = layers[i].gammadot0; m->hn[which][i] = layers[i].stressexponent; }
    m->R.mech[which].active = 1;
    int rc = update_gdp(m, which);
    if (rc) return rc;
  } else {
    HIP_TRY(hipMemcpyAsync(m->ftab, fd.data(), 4L * h3 * 8, hipMemcpyHostToDevice, m->stream));
    m->R.fr.active = m->R.fr.np > 0;
  }
  m->mech[which] = 1;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

int relax_model_set_prestress(relax_model_t* m, const relax_tensor_layer_t* p, int n) {
  if (!m || !p || n != m->g.h3) return RELAX_ERR_ARG;
  const int h3 = m->g.h3;
  std::vector<float> sp(6 * h3), taup(h3);
  for (int i = 0; i < h3; ++i) {
    const relax_tensor_t& t = p[i].t;
    const float diag = (float)((double)(t.s11 + t.s22 + t.s33) / 3.0);
    const float d[6] = {t.s11 - diag, t.s12, t.s13, t.s22 - diag, t.s23, t.s33 - diag};
    for (int k = 0; k < 6; ++k) sp[6 * i + k] = d[k];
    taup[i] = (float)std::sqrt(((double)(d[0] * d[0]) + 2.0 * (double)(d[1] * d[1]) + 2.0 * (double)(d[2] * d[2]) +
                                (double)(d[3] * d[3]) + 2.0 * (double)(d[4] * d[4]) + (double)(d[5] * d[5])) / 2.0);
  }
  float* spd = m->tables + 10L * h3;
  HIP_TRY(hipMemcpyAsync(spd, sp.data(), 6L * h3 * 4, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->tables + 16L * h3, taup.data(), h3 * 4, hipMemcpyHostToDevice, m->stream));
  m->htaup.assign(taup.begin(), taup.end());
  for (int k = 0; k < 2; ++k) {
    int rc = update_gdp(m, k);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

PlaneDev make_plane(const relax_plane_t& p, double beta, double dx1, double dx3) {
  // the reference passes a plane's width as L (down dip) and its length as W
  // (along strike): relax.f90:602-612, :1199-1206
  PlaneDev P;
  P.x = p.x; P.y = p.y; P.z = p.z; P.L = p.width; P.W = p.length; P.beta = beta;
  P.cstrike = std::cos(p.strike); P.sstrike = std::sin(p.strike);
  P.cdip = std::cos(p.dip); P.sdip = std::sin(p.dip);
  P.rake = p.rake;
  const double cr = std::cos(p.rake), sr = std::sin(p.rake);
  P.Wp = P.W * (1.0 + 2.0 * beta) / 2.0;
  P.Lp = P.L * (1.0 + 2.0 * beta) / 2.0;
  const double x2r = P.cstrike * p.x - P.sstrike * p.y;
  P.xr = P.cdip * x2r - P.sdip * p.z;
  P.yr = P.sstrike * p.x + P.cstrike * p.y;
  P.zr = P.sdip * x2r + P.cdip * p.z;
  P.n[0] = +P.cdip * P.cstrike; P.n[1] = -P.cdip * P.sstrike; P.n[2] = -P.sdip;
  P.r[0] = P.sstrike * cr + P.cstrike * P.sdip * sr;
  P.r[1] = P.cstrike * cr - P.sstrike * P.sdip * sr;
  P.r[2] = P.cdip * sr;
  P.minlw = std::min(p.length, p.width) / std::sqrt(dx1 * dx3);
  return P;
}

int relax_model_set_planes(relax_model_t* m, const relax_plane_t* planes, int np, const relax_slippatch_t* patches) {
  if (!m || np < 0 || (np > 0 && !planes)) return RELAX_ERR_ARG;
  if (m->planes_dev) { (void)hipFree(m->planes_dev); m->planes_dev = nullptr; }
  if (m->patches_dev) { (void)hipFree(m->patches_dev); m->patches_dev = nullptr; }
  m->host_planes.assign(planes, planes + np);
  m->plane_dev_host.clear();
  m->patch_off.assign(np + 1, 0);
  for (int k = 0; k < np; ++k) {
    m->plane_dev_host.push_back(make_plane(planes[k], m->desc.beta, m->desc.dx1, m->desc.dx3));
    m->patch_off[k + 1] = m->patch_off[k] + planes[k].px2 * planes[k].px3;
  }
  if (np > 0) {
    HIP_TRY(hipMalloc(&m->planes_dev, sizeof(PlaneDev) * np));
    HIP_TRY(hipMemcpy(m->planes_dev, m->plane_dev_host.data(), sizeof(PlaneDev) * np, hipMemcpyHostToDevice));
    const long npatch = m->patch_off[np];
    if (npatch > 0) {
      HIP_TRY(hipMalloc(&m->patches_dev, sizeof(relax_slippatch_t) * npatch));
      if (patches) HIP_TRY(hipMemcpy(m->patches_dev, patches, sizeof(relax_slippatch_t) * npatch, hipMemcpyHostToDevice));
      else HIP_TRY(hipMemset(m->patches_dev, 0, sizeof(relax_slippatch_t) * npatch));
    }
  }
  m->R.fr.planes = m->planes_dev;
  m->R.fr.np = np;
  m->R.fr.active = (np > 0) && m->mech[2];
  return RELAX_OK;
}

int relax_model_get_patches(relax_model_t* m, relax_slippatch_t* patches) {
  if (!m || !patches) return RELAX_ERR_ARG;
  const long n = m->patch_off.empty() ? 0 : m->patch_off.back();
  if (n == 0) return RELAX_OK;
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(patches, m->patches_dev, sizeof(relax_slippatch_t) * n, hipMemcpyDeviceToHost));
  return RELAX_OK;
}

static int copy_field(float* dst, const float* src, const Grid& g, int align, bool to_dev, hipStream_t st) {
  const size_t hp = (size_t)(align ? g.n1 + 2 : g.n1) * 4, dp = (size_t)g.n1 * 4;
  const size_t rows = (size_t)g.n2 * g.h3;
  if (to_dev) HIP_TRY(hipMemcpy2DAsync(dst, dp, src, hp, dp, rows, hipMemcpyHostToDevice, st));
  else HIP_TRY(hipMemcpy2DAsync(dst, hp, src, dp, dp, rows, hipMemcpyDeviceToHost, st));
  return RELAX_OK;
}

int relax_model_upload(relax_model_t* m, const float* u1, const float* u2, const float* u3,
                       const relax_tensor_t* tau, const relax_tensor_t* sig, const float* gamma, double t, int oi,
                       int i, int align_data) {
  if (!m || !u1 || !u2 || !u3 || !tau) return RELAX_ERR_ARG;
  const Grid& g = m->g;
  const long H = g.H;
  int rc;
  if ((rc = copy_field(m->u, u1, g, align_data, true, m->stream))) return rc;
  if ((rc = copy_field(m->u + H, u2, g, align_data, true, m->stream))) return rc;
  if ((rc = copy_field(m->u + 2 * H, u3, g, align_data, true, m->stream))) return rc;
  if (gamma) { if ((rc = copy_field(m->gam, gamma, g, align_data, true, m->stream))) return rc; }
  else HIP_TRY(hipMemsetAsync(m->gam, 0, H * 4, m->stream));
  // tensors: stage the AoS copy in the sig buffer's space via a temporary
  float* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, 6 * H * 4));
  rc = RELAX_OK;
  if (hipMemcpyAsync(tmp, tau, 6 * H * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
  if (!rc) rc = launch_aos_soa(tmp, m->tau, H, true, m->stream);
  if (!rc && sig) {
    if (hipMemcpyAsync(tmp, sig, 6 * H * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
    if (!rc) rc = launch_aos_soa(tmp, m->sig, H, true, m->stream);
  }
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
  (void)hipFree(tmp);
  if (rc) return rc;
  k_reset_scalars<<<1, 1, 0, m->stream>>>(m->sc, t, oi, i);
  if (!sig) {  // sig = -tau + C:eps(u), relax.f90:369-371
    rc = launch_stress(STRESS_SET, g, m->R, m->S, m->u, m->sig, m->tau, m->gam, m->sc, m->partials, m->stream);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

int relax_model_download(relax_model_t* m, float* u1, float* u2, float* u3, relax_tensor_t* tau,
                         relax_tensor_t* sig, float* gamma, double* t, int align_data) {
  if (!m) return RELAX_ERR_ARG;
  const Grid& g = m->g;
  const long H = g.H;
  int rc;
  if (u1 && (rc = copy_field(u1, m->u, g, align_data, false, m->stream))) return rc;
  if (u2 && (rc = copy_field(u2, m->u + H, g, align_data, false, m->stream))) return rc;
  if (u3 && (rc = copy_field(u3, m->u + 2 * H, g, align_data, false, m->stream))) return rc;
  if (gamma && (rc = copy_field(gamma, m->gam, g, align_data, false, m->stream))) return rc;
  if (tau || sig) {
    float* tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, 6 * H * 4));
    rc = RELAX_OK;
    if (tau) {
      rc = launch_aos_soa(tmp, m->tau, H, false, m->stream);
      if (!rc && hipMemcpyAsync(tau, tmp, 6 * H * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
      if (!rc && hipStreamSynchronize(m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
    }
    if (!rc && sig) {
      rc = launch_aos_soa(tmp, m->sig, H, false, m->stream);
      if (!rc && hipMemcpyAsync(sig, tmp, 6 * H * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
    }
    if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
    (void)hipFree(tmp);
    if (rc) return rc;
  }
  if (t) {
    StepScalars s;
    HIP_TRY(hipMemcpyAsync(&s, m->sc, sizeof(s), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    *t = s.t;
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

// Phase timers: when a profile is being taken, an event is recorded after
// every launch and the time since the previous event is charged to the
// launch's kernel (the GPU runs the stream back to back, so the difference is
// the kernel's duration plus a negligible launch gap).
static void mark(relax_model_t* m, int kind) {
  if (!m->prof_on) return;
  if (m->prof_n >= (int)m->prof_ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    m->prof_ev.push_back(e);
  }
  (void)hipEventRecord(m->prof_ev[m->prof_n], m->stream);
  m->prof_kind.resize(m->prof_ev.size());
  m->prof_kind[m->prof_n] = kind;
  ++m->prof_n;
}

// Green's function of the work field W (body force, upper half) with the
// traction planes T; the last pass writes u + c*G into `out`.
static int model_green(relax_model_t* m, int epi) {
  const Grid& g = m->g;
  hipStream_t st = m->stream;
  const int n1 = g.n1, n2 = g.n2, h3 = g.h3, m1 = g.m1;
  float2* Wc = reinterpret_cast<float2*>(m->W);
  float2* Tc = reinterpret_cast<float2*>(m->T);
  const double dv = m->desc.dx1 * m->desc.dx2 * m->desc.dx3;
  int rc;
  if ((rc = launch_fft_x_r2c(n1, m->W, n1, Wc, m1, m->NQ, 1, 3L * n2 * h3, (float)dv, st))) return rc;
  mark(m, RELAX_PROF_FFT_X_FWD);
  if ((rc = launch_fft_x_r2c(n1, m->T, n1, Tc, m1, m->TN, 1, 3L * n2, (float)(m->desc.dx1 * m->desc.dx2), st))) return rc;
  mark(m, RELAX_PROF_SMALL);
  if ((rc = launch_fft_cols(n2, -1, Wc, m1, (long)n2 * m1, m1, 3 * h3, st))) return rc;
  mark(m, RELAX_PROF_FFT_Y);
  if ((rc = launch_fft_cols(n2, -1, Tc, m1, (long)n2 * m1, m1, 3, st))) return rc;
  mark(m, RELAX_PROF_SMALL);
  if ((rc = launch_fft_rows(n2, -1, m->NQ, n2, 3L * h3, st))) return rc;
  mark(m, RELAX_PROF_SMALL);
  if ((rc = launch_fft_rows(n2, -1, m->TN, n2, 3, st))) return rc;
  mark(m, RELAX_PROF_SMALL);
  if ((rc = launch_green_columns(m->gc, h3, Wc, m->NQ, Tc, m->TN, st))) return rc;
  mark(m, RELAX_PROF_GREEN_Z);
  if ((rc = launch_fft_cols(n2, +1, Wc, m1, (long)n2 * m1, m1, 3 * h3, st))) return rc;
  mark(m, RELAX_PROF_FFT_Y);
  if ((rc = launch_fft_rows(n2, +1, m->NQ, n2, 3L * h3, st))) return rc;
  mark(m, RELAX_PROF_SMALL);
  const double inv = 1.0 / (n1 * m->desc.dx1 * n2 * m->desc.dx2 * g.n3 * m->desc.dx3);
  float* out = epi == EPI_PRED ? m->W : m->u;  // EPI_PLAIN writes G into u (coseismic set-up)
  rc = launch_fft_x_c2r(n1, epi, Wc, m1, m->NQ, 1, out, n1, 3L * n2 * h3, (float)inv, m->u, n1, m->sc, st);
  mark(m, RELAX_PROF_FFT_X_INV);
  return rc;
}

static TimeLine timeline(const relax_model_t* m) {
  TimeLine tl;
  tl.interval = m->desc.interval; tl.odt = m->desc.odt; tl.tscale = m->desc.tscale; tl.skip = m->desc.skip;
  tl.has_next = m->events.size() > 1;
  tl.next_event = tl.has_next ? m->events[1] : 0.0;
  tl.etime = m->events.empty() ? 0.0 : m->events[0];
  for (int k = 0; k < 3; ++k) tl.mech[k] = m->mech[k];
  return tl;
}

static int enqueue_step(relax_model_t* m, const double* dt_override) {
  hipStream_t st = m->stream;
  const Grid& g = m->g;
  const TimeLine tl = timeline(m);
  int rc;
  mark(m, -1);
  // predictor
  if ((rc = launch_bodyforce(g, m->R, m->S, m->sig, m->W, m->T, m->sc, BODY_PRED, st))) return rc;
  mark(m, RELAX_PROF_BODYFORCE);
  k_finalize_dt<<<1, 1, 0, st>>>(m->sc, tl, dt_override ? 1 : 0, dt_override ? *dt_override : 0.0);
  mark(m, RELAX_PROF_SMALL);
  if ((rc = model_green(m, EPI_PRED))) return rc;
  if ((rc = launch_stress(STRESS_PRED, g, m->R, m->S, m->W, m->sig, m->tau, m->gam, m->sc, m->partials, st))) return rc;
  mark(m, RELAX_PROF_STRESS_PRED);
  // corrector
  if ((rc = launch_bodyforce(g, m->R, m->S, m->sig, m->W, m->T, m->sc, BODY_CORR, st))) return rc;
  mark(m, RELAX_PROF_BODYFORCE);
  if (m->R.fr.active && m->patches_dev) {
    for (int k = 0; k < m->R.fr.np; ++k) {
      const relax_plane_t& p = m->host_planes[k];
      if ((rc = launch_monitorfriction(m->plane_dev_host[k], p.rake, p.px2, p.px3, g, m->desc.dx1, m->desc.dx2,
                                       m->desc.dx3, g.n3, m->sig, m->R.fr.v0, m->R.fr.tauc, m->R.fr.friction,
                                       m->R.fr.cohesion, m->patches_dev + m->patch_off[k], m->sc, st)))
        return rc;
      mark(m, RELAX_PROF_SMALL);
    }
  }
  if ((rc = model_green(m, EPI_CORR))) return rc;
  if ((rc = launch_stress(STRESS_CORR, g, m->R, m->S, m->u, m->sig, m->tau, m->gam, m->sc, m->partials, st))) return rc;
  mark(m, RELAX_PROF_STRESS_CORR);
  if (m->patches_dev && m->patch_off.back() > 0) {
    const long n = m->patch_off.back();
    k_frictionadd<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(m->patches_dev, n, m->sc, 0, 0.0);
    mark(m, RELAX_PROF_SMALL);
  }
  k_reduce_partials<<<(unsigned)m->nred, 256, 0, st>>>(m->partials, m->nblk, m->partials + m->nblk);
  k_end_step<<<1, 256, 0, st>>>(m->sc, m->partials + m->nblk, m->nred, m->desc.dx1 * m->desc.dx2 * m->desc.dx3, tl);
  mark(m, RELAX_PROF_SMALL);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int relax_model_profile(relax_model_t* m, int nsteps, double* ms, int* launches) {
  if (!m || nsteps < 1 || !ms || !launches) return RELAX_ERR_ARG;
  for (int k = 0; k < RELAX_PROF_KINDS; ++k) { ms[k] = 0.0; launches[k] = 0; }
  // all steps are enqueued back to back with their markers; one
  // synchronisation at the end, so the timed steps run as they do unprofiled
  m->prof_on = true;
  m->prof_n = 0;
  int rc = RELAX_OK;
  for (int s = 0; s < nsteps && !rc; ++s) rc = enqueue_step(m, nullptr);
  m->prof_on = false;
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  for (int i = 1; i < m->prof_n; ++i) {
    const int k = m->prof_kind[i];
    if (k < 0 || k >= RELAX_PROF_KINDS) continue;  // step-start markers
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, m->prof_ev[i - 1], m->prof_ev[i]));
    ms[k] += t;
    launches[k] += 1;
  }
  return RELAX_OK;
}

static void fill_info(const StepScalars& s, relax_step_info_t* info) {
  info->i = s.it; info->Dt = s.Dt; info->tm = s.tm;
  for (int k = 0; k < 3; ++k) info->maxwell[k] = s.maxwell[k];
  info->power = s.power; info->cei = s.cei; info->t = s.t; info->oi = s.oi; info->status = s.status;
}

int relax_model_step(relax_model_t* m, const double* dt_override, relax_step_info_t* info) {
  if (!m) return RELAX_ERR_ARG;
  int rc = enqueue_step(m, dt_override);
  if (rc) return rc;
  if (info) {
    HIP_TRY(hipMemcpyAsync(m->host_sc, m->sc, sizeof(StepScalars), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    fill_info(*m->host_sc, info);
    if (m->host_sc->status) return m->host_sc->status;
  }
  return RELAX_OK;
}

int relax_model_run(relax_model_t* m, int nsteps, const double* dt_override, relax_step_info_t* info, int* done) {
  if (!m || nsteps < 0) return RELAX_ERR_ARG;
  int k = 0;
  for (; k < nsteps; ++k) {
    HIP_TRY(hipMemcpyAsync(m->host_sc, m->sc, sizeof(StepScalars), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (m->host_sc->status) { if (done) *done = k; return m->host_sc->status; }
    if (m->host_sc->t >= m->desc.interval) break;
    relax_step_info_t tmp;
    const int rc = relax_model_step(m, dt_override ? dt_override + k : nullptr, info ? info + k : &tmp);
    if (rc) { if (done) *done = k; return rc; }
  }
  if (done) *done = k;
  return RELAX_OK;
}

// --------------------------------------------------- coseismic set-up ----
// momentdensityshear (elastic3d.f90:2661-2750) of a list of patches added to
// tau, as dislocations does at each event (relax.f90:1199-1206)
__global__ void k_add_shear(const PlaneDev* __restrict__ P, const double* __restrict__ slip, int np, float rmu,
                            Grid g, double dx1, double dx2, double dx3, float* tau) {
  const long H = g.H;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < H; o += (long)gridDim.x * blockDim.x) {
    const int i1 = o % g.n1, i2 = (o / g.n1) % g.n2, i3 = o / ((long)g.n1 * g.n2);
    const double x3 = (double)i3 * dx3;
    const double x1 = d_shifted(i1, g.n1, dx1), x2 = d_shifted(i2, g.n2, dx2);
    float acc[6] = {0, 0, 0, 0, 0, 0};
    bool any = false;
    for (int k = 0; k < np; ++k) {
      const PlaneDev& p = P[k];
      if (fabs(x3 - p.z) > p.Lp) continue;
      if ((fabs(x1 - p.x) > p.Wp) || (fabs(x2 - p.y) > p.Wp)) continue;
      const double x2r = p.cstrike * x1 - p.sstrike * x2;
      const double x1s = p.cdip * x2r - p.sdip * x3, x1i = p.cdip * x2r + p.sdip * x3;
      if ((fabs(x1s - p.xr) > 7.01 * dx1) && (fabs(x1i - p.xr) > 7.01 * dx1)) continue;
      const double x2s = p.sstrike * x1 + p.cstrike * x2;
      const double x3s = p.sdip * x2r + p.cdip * x3, x3i = -p.sdip * x2r + p.cdip * x3;
      const double t2 = d_omega((x2s - p.yr) / p.W, p.beta);
      double ap = d_gauss(x1s - p.xr, dx1) * t2 * d_omega((x3s - p.zr) / p.L, p.beta);
      ap = ap + d_gauss(x1i - p.xr, dx1) * t2 * d_omega((x3i + p.zr) / p.L, p.beta);
      const double n[3] = {p.n[0] * ap, p.n[1] * ap, p.n[2] * ap};
      const double s[3] = {p.r[0] * slip[k], p.r[1] * slip[k], p.r[2] * slip[k]};
      const float e[6] = {(float)(n[0] * s[0]), (float)((n[0] * s[1] + n[1] * s[0]) / 2.0),
                          (float)((n[0] * s[2] + n[2] * s[0]) / 2.0), (float)(n[1] * s[1]),
                          (float)((n[1] * s[2] + n[2] * s[1]) / 2.0), (float)(n[2] * s[2])};
      // each patch adds in turn, as the reference's sequence of calls does
      for (int c = 0; c < 6; ++c) acc[c] = (any ? acc[c] : tau[c * H + o]) + rmu * e[c];
      any = true;
    }
    if (any)
      for (int c = 0; c < 6; ++c) tau[c * H + o] = acc[c];
  }
}

int relax_model_add_shear(relax_model_t* m, const relax_plane_t* patches, const double* slip, int n) {
  if (!m || n < 0 || (n > 0 && (!patches || !slip))) return RELAX_ERR_ARG;
  if (n == 0) return RELAX_OK;
  std::vector<PlaneDev> hp(n);
  for (int k = 0; k < n; ++k) hp[k] = make_plane(patches[k], m->desc.beta, m->desc.dx1, m->desc.dx3);
  PlaneDev* dp = nullptr;
  double* ds = nullptr;
  HIP_TRY(hipMalloc(&dp, sizeof(PlaneDev) * n));
  if (hipMalloc(&ds, sizeof(double) * n) != hipSuccess) { (void)hipFree(dp); return RELAX_ERR_NOMEM; }
  int rc = RELAX_OK;
  if (hipMemcpyAsync(dp, hp.data(), sizeof(PlaneDev) * n, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemcpyAsync(ds, slip, sizeof(double) * n, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    rc = RELAX_ERR_HIP;
  if (!rc) {
    k_add_shear<<<grid_1d(m->g.H), 256, 0, m->stream>>>(dp, ds, n, 2.0f * (float)m->desc.mu, m->g, m->desc.dx1,
                                                        m->desc.dx2, m->desc.dx3, m->tau);
    if (hipGetLastError() != hipSuccess) rc = RELAX_ERR_HIP;
  }
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
  (void)hipFree(dp);
  (void)hipFree(ds);
  return rc;
}

int relax_model_coseismic(relax_model_t* m) {
  if (!m) return RELAX_ERR_ARG;
  const Grid& g = m->g;
  int rc;
  // body force and surface traction of tau, then u = G[f, t] and
  // sig = -tau + C:eps(u) (relax.f90:318-371 with the FIR divergence)
  if ((rc = launch_bodyforce(g, m->R, m->S, m->tau, m->W, m->T, m->sc, BODY_RAW, m->stream))) return rc;
  if ((rc = model_green(m, EPI_PLAIN))) return rc;
  if ((rc = launch_stress(STRESS_SET, g, m->R, m->S, m->u, m->sig, m->tau, m->gam, m->sc, m->partials, m->stream)))
    return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}
