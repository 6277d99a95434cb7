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
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "kernels.h"
#include "model.h"
#include "stencil.h"

static const double STEP_MAX = 1e7;  // relax.f90:202

// A member of a batch: wait for the batch's queued steps (defined with the
// batches at the end of the file).
static int wait_batch(relax_model_t* m);

// ------------------------------------------------------------- kernels ----
__global__ void k_reset_scalars(StepScalars* sc, double t, int oi, int it) {
  sc->t = t; sc->Dt = 0; sc->tm = 0; sc->e = 0;
  for (int k = 0; k < NMECH; ++k) sc->maxwell[k] = STEP_MAX;
  sc->power = sc->cei = 0;
  sc->oi = oi; sc->it = it; sc->status = 0;
  for (int k = 0; k < NMECH; ++k) sc->tmin_bits[k] = 0x7f800000u;
}

// One thread per model: the relaxation-time minima of the predictor's
// rheology pass become tm and Dt (relax.f90:659-677, integrationstep
// :1420-1452).
__global__ void k_finalize_dt(const ModelDev* __restrict__ md, int nm, int use_override, double dt_override) {
  const int mi = blockIdx.x * blockDim.x + threadIdx.x;
  if (mi >= nm) return;
  StepScalars* sc = md[mi].sc;
  const TimeLine tl = md[mi].tl;
  double maxwell[NMECH], rate = 0.0;
  for (int k = 0; k < NMECH; ++k) {
    const float v = fminf(__uint_as_float(sc->tmin_bits[k]), (float)STEP_MAX);
    maxwell[k] = fmin((double)v, STEP_MAX);
    if (!tl.mech[k]) maxwell[k] = STEP_MAX;
  }
  for (int k = 0; k < NMECH; ++k) rate = rate + (double)tl.mech[k] / maxwell[k];
  double tm = 1.0 / rate;
  tm = fmin(tm, STEP_MAX);
  // inter-seismic loading with nothing relaxing: a small step
  // (relax.f90:669-674)
  if (tl.inter && tm == STEP_MAX) tm = tl.interval / 20.0;
  const double t = sc->t;
  double Dt = tm / 10.0;
  if (tl.skip == 0) {
    if ((t + Dt) >= ((double)sc->oi * tl.odt) - Dt * 0.04) Dt = (double)sc->oi * tl.odt - t;
  } else {
    Dt = Dt * tl.tscale;
    // reduce the step so that t + Dt is the next event's time
    const int e = sc->e;
    if (e + 1 < tl.ne && (t + Dt - tl.ev[e + 1]) >= 0.0) Dt = tl.ev[e + 1] - t;
  }
  if (use_override) Dt = dt_override;
  sc->Dt = Dt;
  sc->tm = tm;
  for (int k = 0; k < NMECH; ++k) sc->maxwell[k] = maxwell[k];
  sc->it = sc->it + 1;
}

// First level of the amplitude reduction: workgroup g sums the stress
// kernel's partials [g*RED_CHUNK, (g+1)*RED_CHUNK) in a fixed order.
// The second-level partials follow the first nb in the model's buffer;
// the model is blockIdx.y.
static const int RED_CHUNK = 2048;
__global__ __launch_bounds__(256) void k_reduce_partials(const ModelDev* __restrict__ md, long nb) {
  __shared__ double2 red[4];
  const double2* __restrict__ partials = md[blockIdx.y].partials;
  double2* __restrict__ out = md[blockIdx.y].partials + nb;
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

// One workgroup per model (blockIdx.x): power and C:E^i from the nb
// second-level partials, then t += Dt and the output counter (isoutput,
// relax.f90:1364-1376).
__global__ __launch_bounds__(256) void k_end_step(const ModelDev* __restrict__ md, long nfirst, long nb) {
  __shared__ double2 red[256];
  const ModelDev& D = md[blockIdx.x];
  StepScalars* sc = D.sc;
  const double2* partials = D.partials + nfirst;
  const double dv = D.dv;
  const TimeLine& tl = D.tl;
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
    // the event switch (relax.f90:858-863): the next event is reached; its
    // slip is applied by relax_model_event
    int e = sc->e;
    if (e + 1 < tl.ne && fabs(t - tl.ev[e + 1]) < 1e-6) sc->e = ++e;
    const double etime = tl.ne > 0 ? tl.ev[e] : 0.0;
    const bool out = ((0 == tl.skip) && (fabs(t - oi * tl.odt) < 1e-6 * tl.odt)) ||
                     ((0 < tl.skip) && ((it - 1) % tl.skip == 0)) || (fabs(t - etime) < 1e-6);
    if (out) sc->oi = oi + 1;
    for (int k = 0; k < NMECH; ++k) sc->tmin_bits[k] = 0x7f800000u;
  }
}

// fix, elastic3d.f90:70-83 (rounds half down)
__device__ __forceinline__ int d_fix(double number) {
  const double f = floor(number), c = ceil(number);
  return (number - f > 0.5) ? (int)c : (int)f;
}
static int host_fix(double number) {
  const double f = std::floor(number), c = std::ceil(number);
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
                  m->planes_dev, m->patches_dev, m->md_dev, m->shear_dev, m->shear_slip, m->loads_dev, m->E, m->Em,
                  m->dg[0], m->dg[1], m->dg[2], m->dg[3], m->zones_dev, m->pts_dev, m->ev_dev, m->mi,
                  m->eig_dev};
  for (void* p : ptrs)
    if (p) (void)rx_free(p);
  if (m->host_sc) (void)rx_host_free(m->host_sc);
  for (hipEvent_t e : m->prof_ev) (void)hipEventDestroy(e);
  if (m->stream) (void)rx_stream_destroy(m->stream);
  delete m;
}

// --------------------------------------------------------- tuning table ----
// Launch-shape choices per device and grid, found by relax_model_autotune
// and used by every later model of that grid in the process.
struct TuneKey {
  int dev, n1, n2, n3;
  bool operator<(const TuneKey& o) const {
    return dev != o.dev ? dev < o.dev : n1 != o.n1 ? n1 < o.n1 : n2 != o.n2 ? n2 < o.n2 : n3 < o.n3;
  }
};
struct TuneChoice {
  int cols, green, body;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tune;

static void tune_lookup(relax_model_t* m) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(TuneKey{m->device, m->g.n1, m->g.n2, m->g.n3});
  if (it == g_tune.end()) return;
  m->tune_cols = it->second.cols;
  m->tune_green = it->second.green;
  m->tune_body = it->second.body;
}

static TimeLine timeline(const relax_model_t* m) {
  TimeLine tl;
  tl.interval = m->desc.interval; tl.odt = m->desc.odt; tl.tscale = m->desc.tscale; tl.skip = m->desc.skip;
  tl.ev = m->ev_dev;
  tl.ne = (int)m->events.size();
  for (int k = 0; k < NMECH; ++k) tl.mech[k] = m->mech[k];
  tl.inter = m->inter;
  return tl;
}

// The kernels a model needs: the extended rheology with weak zones or
// transient creep, the fault-creep kernels with afterslip planes.
static int rheo_class(const relax_model_t* m) {
  const RheoParams& R = m->R;
  if (R.tr[0].active || R.tr[1].active || R.mech[0].dg || R.mech[1].dg || R.inter) return RHEO_EXT;
  return (R.fr.active && R.fr.np > 0) ? RHEO_CREEP : RHEO_VISCOUS;
}

// Refresh the model's device record after a change of its rheology, planes
// or time line.  The copy is ordered on the model's stream before any later
// launch; the host image stays in the model, so it outlives the copy.
static int sync_md(relax_model_t* m) {
  if (!m->md_dirty) return RELAX_OK;
  ModelDev& d = m->md_host;
  d.u = m->u; d.sig = m->sig; d.tau = m->tau; d.gam = m->gam; d.W = m->W; d.NQ = m->NQ; d.T = m->T; d.TN = m->TN;
  d.E = m->E; d.Em = m->Em;
  d.sc = m->sc; d.partials = m->partials;
  d.dv = m->desc.dx1 * m->desc.dx2 * m->desc.dx3;
  d.tl = timeline(m);
  d.R = m->R;
  HIP_TRY(hipMemcpyAsync(m->md_dev, &d, sizeof(ModelDev), hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->md_dirty = false;
  ++m->md_version;
  return RELAX_OK;
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
  if (!m->events.empty()) {
    if (rx_malloc(&m->ev_dev, m->events.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(m->ev_dev, m->events.data(), m->events.size() * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess) {
      free_model(m);
      return RELAX_ERR_NOMEM;
    }
  }
  m->desc.event_times = nullptr;
  Grid& g = m->g;
  g.n1 = d->sx1; g.n2 = d->sx2; g.n3 = d->sx3; g.h3 = d->sx3 / 2; g.m1 = d->sx1 / 2;
  g.H = (long)g.n1 * g.n2 * g.h3;
  g.dx1 = (float)d->dx1; g.dx2 = (float)d->dx2; g.dx3 = (float)d->dx3;
  const long H = g.H;
  bool ok = rx_stream_create(&m->stream) == hipSuccess;
  auto alloc = [&](void** p, size_t bytes) { if (ok && rx_malloc(p, bytes) != hipSuccess) ok = false; };
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
  alloc((void**)&m->tables, TABLE_ROWS * g.h3 * 4);
  alloc((void**)&m->ftab, 4L * g.h3 * 8);
  alloc((void**)&m->md_dev, sizeof(ModelDev));
  m->self = m;
  if (ok && rx_host_malloc(&m->host_sc, sizeof(StepScalars) * 64) != hipSuccess) ok = false;
  m->host_cap = 64;
  if (!ok) { free_model(m); return RELAX_ERR_NOMEM; }
  // zero state; no mechanisms until set_structure
  if (hipMemsetAsync(m->u, 0, 3 * H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->sig, 0, 6 * H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->tau, 0, 6 * H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->gam, 0, H * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->tables, 0, TABLE_ROWS * g.h3 * 4, m->stream) != hipSuccess ||
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
    R.mech[k].rp = m->tables + (19 + k) * h3;
    R.mech[k].dg = nullptr;
    R.mech[k].active = 0;
    // transient: gammadot0, power, Gk (in the cohesion slot)
    R.tr[k].gammadot0 = m->tables + (21 + 3 * k) * h3;
    R.tr[k].power = m->tables + (22 + 3 * k) * h3;
    R.tr[k].cohesion = m->tables + (23 + 3 * k) * h3;
    R.tr[k].gdp = nullptr;
    R.tr[k].rp = nullptr;
    R.tr[k].dg = nullptr;
    R.tr[k].active = 0;
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
  if (!twiddle_table(g.n1) || !twiddle_table(g.n2) || !twiddle_table(g.n3) ||
      (g.n2 >= 128 && !twiddle_table(g.n2 / 8))) {
    free_model(m);
    return RELAX_ERR_NOMEM;
  }
  if (hipStreamSynchronize(m->stream) != hipSuccess) { free_model(m); return RELAX_ERR_HIP; }
  m->md_dirty = true;
  if (sync_md(m) != RELAX_OK) { free_model(m); return RELAX_ERR_HIP; }
  tune_lookup(m);
  *out = m;
  return RELAX_OK;
}

int relax_model_reset(relax_model_t* m) {
  if (!m) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const long H = m->g.H, h3 = m->g.h3;
  HIP_TRY(hipMemsetAsync(m->u, 0, 3 * H * 4, m->stream));
  HIP_TRY(hipMemsetAsync(m->sig, 0, 6 * H * 4, m->stream));
  HIP_TRY(hipMemsetAsync(m->tau, 0, 6 * H * 4, m->stream));
  HIP_TRY(hipMemsetAsync(m->gam, 0, H * 4, m->stream));
  HIP_TRY(hipMemsetAsync(m->tables, 0, TABLE_ROWS * h3 * 4, m->stream));
  if (m->E) {
    HIP_TRY(hipMemsetAsync(m->E, 0, 6 * H * 4, m->stream));
    HIP_TRY(hipMemsetAsync(m->Em, 0, 6 * H * 4, m->stream));
  }
  HIP_TRY(hipMemsetAsync(m->ftab, 0, 4L * h3 * 8, m->stream));
  k_reset_scalars<<<1, 1, 0, m->stream>>>(m->sc, 0.0, 1, 0);
  HIP_TRY(hipStreamSynchronize(m->stream));
  for (int k = 0; k < 2; ++k) {
    m->R.mech[k].active = 0;
    m->R.mech[k].dg = nullptr;
    m->R.tr[k].active = 0;
    m->R.tr[k].dg = nullptr;
    m->hg0[k].clear();
    m->hn[k].clear();
  }
  m->htaup.clear();
  m->cur_nloads = 0;
  for (int k = 0; k < NMECH; ++k) m->mech[k] = 0;
  m->host_planes.clear();
  m->plane_dev_host.clear();
  m->patch_off.assign(1, 0);
  m->R.fr.np = 0;
  m->R.fr.active = 0;
  m->R.inter = nullptr;
  m->inter = 0;
  m->md_dirty = true;
  return sync_md(m);
}

int relax_model_destroy(relax_model_t* m) {
  if (!m) return RELAX_ERR_ARG;
  if (m->batch) (void)relax_batch_remove(m->batch, m);
  (void)hipStreamSynchronize(m->stream);
  free_model(m);
  return RELAX_OK;
}

void* relax_model_stream(relax_model_t* m) { return m ? (void*)m->stream : nullptr; }

int relax_model_synchronize(relax_model_t* m) {
  if (!m) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

// prestress rate of each depth, gammadot0 (taup/mu)^(n-1)
// (viscoelastic3d.f90:290), computed once in double
static int update_gdp(relax_model_t* m, int k) {
  const int h3 = m->g.h3;
  if (m->hg0[k].empty()) return RELAX_OK;
  std::vector<float> g(h3), r(h3);
  for (int i = 0; i < h3; ++i) {
    const double taup = m->htaup.empty() ? 0.0 : m->htaup[i];
    const double ratio = std::pow(taup / m->desc.mu, m->hn[k][i] - 1.0);
    g[i] = (float)(m->hg0[k][i] * ratio);
    r[i] = (float)ratio;
  }
  HIP_TRY(hipMemcpyAsync(m->tables + (17L + k) * h3, g.data(), h3 * 4, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->tables + (19L + k) * h3, r.data(), h3 * 4, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

// Transient structure k (0: linear, 1: power law): fluidity, exponent and Gk
// per depth; the first one allocates the transient strain and its mid-step
// copy, both zero (relax.f90:299-303 allocates epsilonik).
static int set_transient(relax_model_t* m, int k, const relax_layer_t* layers, int n) {
  const int h3 = m->g.h3;
  const long H = m->g.H;
  // transienteigenstress returns without effect when the structure does
  // not match the grid (viscoelastic3d.f90:144); here that is an error
  if (n != h3) return RELAX_ERR_ARG;
  std::vector<float> t(3 * h3);
  for (int i = 0; i < h3; ++i) {
    if (layers[i].stressexponent < 0.999999) return RELAX_ERR_POWER;
    t[i] = (float)layers[i].gammadot0;
    t[h3 + i] = (float)layers[i].stressexponent;
    t[2 * h3 + i] = (float)layers[i].Gk;
  }
  if (!m->E) {
    if (rx_malloc(&m->E, 6 * H * 4) != hipSuccess) return RELAX_ERR_NOMEM;
    if (rx_malloc(&m->Em, 6 * H * 4) != hipSuccess) return RELAX_ERR_NOMEM;
    HIP_TRY(hipMemsetAsync(m->E, 0, 6 * H * 4, m->stream));
    HIP_TRY(hipMemsetAsync(m->Em, 0, 6 * H * 4, m->stream));
  }
  HIP_TRY(hipMemcpyAsync(m->tables + (21L + 3 * k) * h3, t.data(), 3L * h3 * 4, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->R.tr[k].active = 1;
  m->mech[3 + k] = 1;
  m->md_dirty = true;
  return RELAX_OK;
}

int relax_model_set_weakzones(relax_model_t* m, int which, const relax_weak_t* zones, int nz) {
  if (!m || nz < 0 || (nz > 0 && !zones) || !(which == 0 || which == 1 || which == 3 || which == 4))
    return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const int slot = which < 2 ? which : which - 1;  // dg[0..3]: linear, power law, two transients
  Mechanism& M = which < 2 ? m->R.mech[which] : m->R.tr[which - 3];
  m->md_dirty = true;
  if (nz == 0) {
    M.dg = nullptr;
    return RELAX_OK;
  }
  const Grid& g = m->g;
  if (!m->dg[slot] && rx_malloc(&m->dg[slot], g.H * 4) != hipSuccess) return RELAX_ERR_NOMEM;
  if (nz > m->zones_cap) {
    if (m->zones_dev) HIP_TRY(rx_free(m->zones_dev));
    m->zones_dev = nullptr;
    m->zones_cap = 0;
    if (rx_malloc(&m->zones_dev, (size_t)nz * sizeof(WeakDev)) != hipSuccess) return RELAX_ERR_NOMEM;
    m->zones_cap = nz;
  }
  // the taper of the reference's calls is 0 (relax.f90:487, :502, :525, :540)
  std::vector<WeakDev> wz(nz);
  for (int i = 0; i < nz; ++i) wz[i] = make_weak(zones[i], 0.0);
  HIP_TRY(hipMemcpyAsync(m->zones_dev, wz.data(), (size_t)nz * sizeof(WeakDev), hipMemcpyHostToDevice, m->stream));
  int rc = launch_builddgammadot0(g.n1, g.n2, g.h3, m->desc.dx1, m->desc.dx2, m->desc.dx3, 0.0, nz, m->zones_dev,
                                  m->dg[slot], m->stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  M.dg = m->dg[slot];
  return RELAX_OK;
}

int relax_model_set_transient_strain(relax_model_t* m, const relax_tensor_t* epsilonik) {
  if (!m || !epsilonik || !m->E) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const long H = m->g.H;
  float* tmp = nullptr;
  if (rx_malloc(&tmp, 6 * H * 4) != hipSuccess) return RELAX_ERR_NOMEM;
  int rc = hipMemcpyAsync(tmp, epsilonik, 6 * H * 4, hipMemcpyHostToDevice, m->stream) == hipSuccess
               ? launch_aos_soa(tmp, m->E, H, true, m->stream)
               : RELAX_ERR_HIP;
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
  (void)rx_free(tmp);
  return rc;
}

int relax_model_get_transient_strain(relax_model_t* m, relax_tensor_t* epsilonik) {
  if (!m || !epsilonik || !m->E) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const long H = m->g.H;
  float* tmp = nullptr;
  if (rx_malloc(&tmp, 6 * H * 4) != hipSuccess) return RELAX_ERR_NOMEM;
  int rc = launch_aos_soa(tmp, m->E, H, false, m->stream);
  if (!rc && hipMemcpyAsync(epsilonik, tmp, 6 * H * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    rc = RELAX_ERR_HIP;
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
  (void)rx_free(tmp);
  return rc;
}

int relax_model_set_structure(relax_model_t* m, int which, const relax_layer_t* layers, int n) {
  if (!m || which < 0 || which >= NMECH) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const int h3 = m->g.h3;
  m->md_dirty = true;
  if (!layers) {
    if (which < 2) m->R.mech[which].active = 0;
    else if (which == 2) m->R.fr.active = 0;
    else m->R.tr[which - 3].active = 0;
    m->mech[which] = 0;
    return RELAX_OK;
  }
  if (which >= 3) return set_transient(m, which - 3, layers, n);
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

// tau(:,:,k) = -prestress(k) over the half grid (relax.f90:295-297); p6 holds
// the six components of each depth
__global__ void k_background_stress(float* __restrict__ tau, const float* __restrict__ p6, long plane, long H) {
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < H; o += (long)gridDim.x * blockDim.x) {
    const long k = o / plane;
    for (int c = 0; c < 6; ++c) tau[c * H + o] = -p6[6 * k + c];
  }
}

int relax_model_set_prestress(relax_model_t* m, const relax_tensor_layer_t* p, int n) {
  if (!m || !p || n != m->g.h3) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const int h3 = m->g.h3;
  {
    // the initial tau: staged through the work field W (3H floats >= 6 h3)
    std::vector<float> p6(6 * h3);
    for (int i = 0; i < h3; ++i) {
      const relax_tensor_t& t = p[i].t;
      const float v[6] = {t.s11, t.s12, t.s13, t.s22, t.s23, t.s33};
      for (int c = 0; c < 6; ++c) p6[6 * i + c] = v[c];
    }
    HIP_TRY(hipMemcpyAsync(m->W, p6.data(), 6L * h3 * 4, hipMemcpyHostToDevice, m->stream));
    k_background_stress<<<grid_1d(m->g.H), 256, 0, m->stream>>>(m->tau, m->W, (long)m->g.n1 * m->g.n2, m->g.H);
    HIP_TRY(hipGetLastError());
  }
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
  P.dip = p.dip;
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
  if (int rc = wait_batch(m)) return rc;
  m->host_planes.assign(planes, planes + np);
  m->plane_dev_host.clear();
  m->patch_off.assign(np + 1, 0);
  for (int k = 0; k < np; ++k) {
    m->plane_dev_host.push_back(make_plane(planes[k], m->desc.beta, m->desc.dx1, m->desc.dx3));
    m->patch_off[k + 1] = m->patch_off[k] + planes[k].px2 * planes[k].px3;
  }
  // buffers only grow, so a model reused for another request allocates
  // (and frees, which would wait for the whole device) nothing
  const long npatch = m->patch_off[np];
  if (np > m->planes_cap) {
    if (m->planes_dev) (void)rx_free(m->planes_dev);
    m->planes_dev = nullptr;
    m->planes_cap = 0;
    HIP_TRY(rx_malloc(&m->planes_dev, sizeof(PlaneDev) * np));
    m->planes_cap = np;
  }
  if (npatch > m->patches_cap) {
    if (m->patches_dev) (void)rx_free(m->patches_dev);
    m->patches_dev = nullptr;
    m->patches_cap = 0;
    HIP_TRY(rx_malloc(&m->patches_dev, sizeof(relax_slippatch_t) * npatch));
    m->patches_cap = npatch;
  }
  if (np > 0)
    HIP_TRY(hipMemcpyAsync(m->planes_dev, m->plane_dev_host.data(), sizeof(PlaneDev) * np, hipMemcpyHostToDevice,
                           m->stream));
  if (npatch > 0) {
    if (patches)
      HIP_TRY(hipMemcpyAsync(m->patches_dev, patches, sizeof(relax_slippatch_t) * npatch, hipMemcpyHostToDevice,
                             m->stream));
    else HIP_TRY(hipMemsetAsync(m->patches_dev, 0, sizeof(relax_slippatch_t) * npatch, m->stream));
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->R.fr.planes = m->planes_dev;
  m->R.fr.np = np;
  m->R.fr.active = (np > 0) && m->mech[2];
  m->md_dirty = true;
  return RELAX_OK;
}

int relax_model_get_patches(relax_model_t* m, relax_slippatch_t* patches) {
  if (!m || !patches) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
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
  if (int rc = wait_batch(m)) return rc;
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
  HIP_TRY(rx_malloc(&tmp, 6 * H * 4));
  rc = RELAX_OK;
  if (hipMemcpyAsync(tmp, tau, 6 * H * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
  if (!rc) rc = launch_aos_soa(tmp, m->tau, H, true, m->stream);
  if (!rc && sig) {
    if (hipMemcpyAsync(tmp, sig, 6 * H * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
    if (!rc) rc = launch_aos_soa(tmp, m->sig, H, true, m->stream);
  }
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
  (void)rx_free(tmp);
  if (rc) return rc;
  k_reset_scalars<<<1, 1, 0, m->stream>>>(m->sc, t, oi, i);
  if (!sig) {  // sig = -tau + C:eps(u), relax.f90:369-371
    if ((rc = sync_md(m))) return rc;
    rc = launch_stress(STRESS_SET, g, m->S, m->md_dev, &m->md_host, 1, RHEO_VISCOUS, m->stream);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

int relax_model_download(relax_model_t* m, float* u1, float* u2, float* u3, relax_tensor_t* tau,
                         relax_tensor_t* sig, float* gamma, double* t, int align_data) {
  if (!m) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const Grid& g = m->g;
  const long H = g.H;
  int rc;
  if (u1 && (rc = copy_field(u1, m->u, g, align_data, false, m->stream))) return rc;
  if (u2 && (rc = copy_field(u2, m->u + H, g, align_data, false, m->stream))) return rc;
  if (u3 && (rc = copy_field(u3, m->u + 2 * H, g, align_data, false, m->stream))) return rc;
  if (gamma && (rc = copy_field(gamma, m->gam, g, align_data, false, m->stream))) return rc;
  if (tau || sig) {
    float* tmp = nullptr;
    HIP_TRY(rx_malloc(&tmp, 6 * H * 4));
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
    (void)rx_free(tmp);
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

int relax_model_surface(relax_model_t* m, float* u1, float* u2, float* u3, int align_data) {
  if (!m) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  const Grid& g = m->g;
  const size_t hp = (size_t)(align_data ? g.n1 + 2 : g.n1) * 4, dp = (size_t)g.n1 * 4;
  float* dst[3] = {u1, u2, u3};
  for (int c = 0; c < 3; ++c) {
    if (!dst[c]) continue;
    HIP_TRY(hipMemcpy2DAsync(dst[c], hp, m->u + c * g.H, dp, dp, (size_t)g.n2, hipMemcpyDeviceToHost, m->stream));
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

// Gather of u (3 planes) and sig (6 planes) at nq point offsets: out holds
// 9 floats per point.
__global__ void k_gather_points(const float* __restrict__ u, const float* __restrict__ sig, long H,
                                const long* __restrict__ off, int nq, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  const long o = off[k];
  float* r = out + 9 * k;
  for (int c = 0; c < 3; ++c) r[c] = u[c * H + o];
  for (int c = 0; c < 6; ++c) r[3 + c] = sig[c * H + o];
}

int relax_model_points(relax_model_t* m, const double* x, int n, double rot, double* u, relax_tensor_t* sig) {
  if (!m || n < 0 || (n > 0 && (!x || !u))) return RELAX_ERR_ARG;
  if (n == 0) return RELAX_OK;
  if (int rc = wait_batch(m)) return rc;
  const Grid& g = m->g;
  const double dx1 = m->desc.dx1, dx2 = m->desc.dx2, dx3 = m->desc.dx3;
  // shiftedindex (elastic3d.f90:4063-4101) over the half grid, as
  // relax.f90:894 passes sx3/2; out of range is an error, not a STOP
  std::vector<long> off(n);
  for (int k = 0; k < n; ++k) {
    const double x1 = x[3 * k], x2 = x[3 * k + 1], x3 = x[3 * k + 2];
    if (x1 > (double)(g.n1 / 2 - 1) * dx1 || x1 < -(double)(g.n1 / 2) * dx1 ||
        x2 > (double)(g.n2 / 2 - 1) * dx2 || x2 < -(double)(g.n2 / 2) * dx2 || x3 > (double)(g.h3 - 1) * dx3 ||
        x3 < 0.0)
      return RELAX_ERR_RANGE;
    const int i1 = (g.n1 + host_fix(x1 / dx1)) % g.n1, i2 = (g.n2 + host_fix(x2 / dx2)) % g.n2;
    const int i3 = host_fix(x3 / dx3);
    off[k] = i1 + (long)g.n1 * (i2 + (long)g.n2 * i3);
  }
  const size_t need = (size_t)n * (sizeof(long) + 9 * sizeof(float));
  if (need > m->pts_cap) {
    if (m->pts_dev) HIP_TRY(rx_free(m->pts_dev));
    m->pts_dev = nullptr;
    m->pts_cap = 0;
    if (rx_malloc(&m->pts_dev, need) != hipSuccess) return RELAX_ERR_NOMEM;
    m->pts_cap = need;
  }
  long* doff = static_cast<long*>(m->pts_dev);
  float* dout = reinterpret_cast<float*>(doff + n);
  std::vector<float> h(9L * n);
  HIP_TRY(hipMemcpyAsync(doff, off.data(), n * sizeof(long), hipMemcpyHostToDevice, m->stream));
  k_gather_points<<<(n + 255) / 256, 256, 0, m->stream>>>(m->u, m->sig, g.H, doff, n, dout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(h.data(), dout, 9L * n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  // exportpoints writes u in the user frame: rotation(v1, v2, -rot)
  // (util.f90:113-116, input.f90:1809-1822); sig as it is
  const double a = -rot * RELAX_DEG2RAD, c = std::cos(a), sn = std::sin(a);
  for (int k = 0; k < n; ++k) {
    const double v1 = h[9 * k], v2 = h[9 * k + 1];
    u[3 * k] = +v1 * c + v2 * sn;
    u[3 * k + 1] = -v1 * sn + v2 * c;
    u[3 * k + 2] = h[9 * k + 2];
    if (sig) {
      const float* t = &h[9 * k + 3];
      sig[k] = relax_tensor_t{t[0], t[1], t[2], t[3], t[4], t[5]};
    }
  }
  return RELAX_OK;
}

// Phase timers: when a profile is being taken, an event is recorded after
// every launch and the time since the previous event is charged to the
// launch's kernel (the GPU runs the stream back to back, so the difference is
// the kernel's duration plus a negligible launch gap).  The events only time:
// they are created without the system-scope fence, whose cache write-back
// and invalidation would otherwise slow the kernels that follow.
static void mark(relax_model_t* m, int kind) {
  if (!m->prof_on) return;
  if (m->prof_n >= (int)m->prof_ev.size()) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
    m->prof_ev.push_back(e);
  }
  (void)hipEventRecord(m->prof_ev[m->prof_n], m->stream);
  m->prof_kind.resize(m->prof_ev.size());
  m->prof_kind[m->prof_n] = kind;
  ++m->prof_n;
}

// A launch context: nm models of one grid and one set of elastic constants
// (those of `lead`), their device records, and the stream.
struct Ctx {
  relax_model_t* lead;
  const ModelDev* md;
  const ModelDev* one;            // host copy of the record of a single model, else nullptr
  int nm;
  int rclass;                     // rheology class of the launches (RHEO_*)
  hipStream_t st;
  relax_model_t* const* models;   // host handles, for the per-plane kernels
  relax_model_t* prof;            // phase timers (single model), or nullptr
  int tune_cols, tune_green, tune_body;  // launch-shape variants
};

static Ctx single(relax_model_t* m) {
  Ctx c;
  c.lead = m; c.md = m->md_dev; c.one = &m->md_host; c.nm = 1; c.rclass = rheo_class(m);
  c.st = m->stream; c.models = &m->self; c.prof = m;
  c.tune_cols = m->tune_cols; c.tune_green = m->tune_green; c.tune_body = m->tune_body;
  return c;
}

static void cmark(const Ctx& c, int kind) {
  if (c.prof) mark(c.prof, kind);
}

// The small transforms of a Green solve's y pass: forward, the traction
// planes' columns and the Nyquist rows of W and T; inverse, those of W.
static YRide ypass_ride(const Ctx& c, int sign) {
  const Grid& g = c.lead->g;
  const ModelDev* md = c.md;
  YRide r;
  r.r0 = MFIELD(float2, md, NQ);
  r.rp0 = g.n2; r.rn0 = 3L * g.h3;
  if (sign < 0) {
    r.tc = MFIELD(float2, md, T);
    r.tjs = g.m1; r.tps = (long)g.n2 * g.m1; r.tnc = g.m1; r.tnp = 3;
    r.r1 = MFIELD(float2, md, TN);
    r.rp1 = g.n2; r.rn1 = 3;
  }
  return r;
}

// Green's function of the work field W (body force, upper half) with the
// traction planes T of every model; the last pass writes u + c*G into W
// (predictor) or u (corrector and EPI_PLAIN).
static int model_green(const Ctx& c, int epi) {
  const relax_model_t* m = c.lead;
  const Grid& g = m->g;
  hipStream_t st = c.st;
  const int n1 = g.n1, n2 = g.n2, h3 = g.h3, m1 = g.m1, nm = c.nm;
  const ModelDev* md = c.md;
  const double dv = m->desc.dx1 * m->desc.dx2 * m->desc.dx3;
  // the y transform of W is split (radix 8 in the x kernels, n2/8 points in
  // the column kernel) from n2 = 128 up; the traction planes keep one pass
  const int ys = n2 >= 128 ? n2 : 0;
  const int l2 = ys ? n2 / 8 : 0;
  const int ly = ys ? l2 : n2, nyp = ys ? 3 * h3 * 8 : 3 * h3;
  int rc;
  // the traction planes ride along with the work field's x transform
  XSide tside;
  tside.in = MFIELD(const float, md, T);
  tside.out = MFIELD(float2, md, T);
  tside.nyq = MFIELD(float2, md, TN);
  tside.in_pitch = n1; tside.out_pitch = m1; tside.nyq_stride = 1; tside.nrows = 3L * n2;
  tside.scale = (float)(m->desc.dx1 * m->desc.dx2);
  if ((rc = launch_fft_x_r2c(n1, MFIELD(const float, md, W), n1, MFIELD(float2, md, W), m1, MFIELD(float2, md, NQ), 1,
                             3L * n2 * h3, (float)dv, nm, st, ys, &tside)))
    return rc;
  cmark(c, RELAX_PROF_FFT_X_FWD);
  // the y pass of W, with the traction planes' y transform and those of the
  // Nyquist planes of W and T in the same launch
  if ((rc = launch_fft_ypass(ly, n2, -1, MFIELD(float2, md, W), m1, (long)ly * m1, m1, nyp, ypass_ride(c, -1), nm, st,
                             c.tune_cols)))
    return rc;
  cmark(c, RELAX_PROF_FFT_Y);
  if ((rc = launch_green_columns(m->gc, h3, MFIELD(float2, md, W), MFIELD(float2, md, NQ), MFIELD(const float2, md, T),
                                 MFIELD(const float2, md, TN), nm, st, l2, c.tune_green)))
    return rc;
  cmark(c, RELAX_PROF_GREEN_Z);
  if ((rc = launch_fft_ypass(ly, n2, +1, MFIELD(float2, md, W), m1, (long)ly * m1, m1, nyp, ypass_ride(c, +1), nm, st,
                             c.tune_cols)))
    return rc;
  cmark(c, RELAX_PROF_FFT_Y);
  const double inv = 1.0 / (n1 * m->desc.dx1 * n2 * m->desc.dx2 * g.n3 * m->desc.dx3);
  const MPtr<float> out = epi == EPI_PRED ? MFIELD(float, md, W) : MFIELD(float, md, u);
  rc = launch_fft_x_c2r(n1, epi, MFIELD(const float2, md, W), m1, MFIELD(const float2, md, NQ), 1, out, n1,
                        3L * n2 * h3, (float)inv, MFIELD(const float, md, u), n1, MFIELD(const StepScalars, md, sc), nm,
                        st, ys);
  cmark(c, RELAX_PROF_FFT_X_INV);
  return rc;
}

__global__ void k_traction_rate(const relax_source_t* __restrict__ l, int nl, int n1, int n2, double dx1, double dx2,
                                const StepScalars* __restrict__ sc, int pred, float* t3);

// The surface-load rates of the models whose current event has loads
static int traction_rates(const Ctx& c, int pred) {
  for (int i = 0; i < c.nm; ++i) {
    relax_model_t* m = c.models[i];
    if (m->cur_nloads == 0) continue;
    const long P = (long)m->g.n1 * m->g.n2;
    k_traction_rate<<<(unsigned)((P + 255) / 256), 256, 0, c.st>>>(m->loads_dev, m->cur_nloads, m->g.n1, m->g.n2,
                                                                   m->desc.dx1, m->desc.dx2, m->sc, pred, m->T + 2 * P);
    if (hipGetLastError() != hipSuccess) return RELAX_ERR_HIP;
    cmark(c, RELAX_PROF_SMALL);
  }
  return RELAX_OK;
}

// The patch velocities of the context's afterslip planes (monitorfriction),
// after the corrector's divergence pass.
static int monitor_friction(const Ctx& c) {
  int rc;
  for (int i = 0; i < c.nm; ++i) {
    relax_model_t* m = c.models[i];
    if (!(m->R.fr.active && m->patches_dev)) continue;
    for (int k = 0; k < m->R.fr.np; ++k) {
      const relax_plane_t& p = m->host_planes[k];
      if ((rc = launch_monitorfriction(m->plane_dev_host[k], p.rake, p.px2, p.px3, m->g, m->desc.dx1, m->desc.dx2,
                                       m->desc.dx3, m->g.n3, m->sig, m->R.fr.v0, m->R.fr.tauc, m->R.fr.friction,
                                       m->R.fr.cohesion, m->patches_dev + m->patch_off[k], m->sc, c.st)))
        return rc;
      cmark(c, RELAX_PROF_SMALL);
    }
  }
  return RELAX_OK;
}

// The end of the step of every model of the context (all of one grid): the
// patch slip (frictionadd), the amplitude reductions, t += Dt.
static int end_steps(const Ctx& c) {
  const int nm = c.nm;
  for (int i = 0; i < nm; ++i) {
    relax_model_t* m = c.models[i];
    if (!(m->patches_dev && m->patch_off.back() > 0)) continue;
    const long n = m->patch_off.back();
    k_frictionadd<<<(unsigned)((n + 255) / 256), 256, 0, c.st>>>(m->patches_dev, n, m->sc, 0, 0.0);
    cmark(c, RELAX_PROF_SMALL);
  }
  k_reduce_partials<<<dim3((unsigned)c.lead->nred, (unsigned)nm), 256, 0, c.st>>>(c.md, c.lead->nblk);
  k_end_step<<<nm, 256, 0, c.st>>>(c.md, c.lead->nblk, c.lead->nred);
  cmark(c, RELAX_PROF_SMALL);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// One RK2 step of every model of the context, enqueued without host
// synchronisation.
static int enqueue_steps(const Ctx& c, const double* dt_override) {
  hipStream_t st = c.st;
  const Grid& g = c.lead->g;
  const StencilConst& S = c.lead->S;
  const int nm = c.nm;
  int rc;
  cmark(c, -1);
  // predictor
  if ((rc = launch_bodyforce(g, S, c.md, c.one, nm, c.rclass, BODY_PRED, st, c.tune_body))) return rc;
  cmark(c, RELAX_PROF_BODYFORCE);
  k_finalize_dt<<<(nm + 63) / 64, 64, 0, st>>>(c.md, nm, dt_override ? 1 : 0, dt_override ? *dt_override : 0.0);
  cmark(c, RELAX_PROF_SMALL);
  if ((rc = traction_rates(c, 1))) return rc;
  if ((rc = model_green(c, EPI_PRED))) return rc;
  if ((rc = launch_stress(STRESS_PRED, g, S, c.md, c.one, nm, c.rclass, st))) return rc;
  cmark(c, RELAX_PROF_STRESS_PRED);
  // corrector
  if ((rc = launch_bodyforce(g, S, c.md, c.one, nm, c.rclass, BODY_CORR, st, c.tune_body))) return rc;
  cmark(c, RELAX_PROF_BODYFORCE);
  if ((rc = monitor_friction(c))) return rc;
  if ((rc = traction_rates(c, 0))) return rc;
  if ((rc = model_green(c, EPI_CORR))) return rc;
  if ((rc = launch_stress(STRESS_CORR, g, S, c.md, c.one, nm, c.rclass, st))) return rc;
  cmark(c, RELAX_PROF_STRESS_CORR);
  return end_steps(c);
}

static int enqueue_step(relax_model_t* m, const double* dt_override) {
  int rc = sync_md(m);
  if (rc) return rc;
  return enqueue_steps(single(m), dt_override);
}

int relax_model_profile(relax_model_t* m, int nsteps, double* ms, int* launches) {
  if (!m || nsteps < 1 || !ms || !launches || m->batch) return RELAX_ERR_ARG;
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
  for (int k = 0; k < NMECH; ++k) info->maxwell[k] = s.maxwell[k];
  info->power = s.power; info->cei = s.cei; info->t = s.t; info->oi = s.oi; info->status = s.status;
  info->e = s.e;
}

int relax_model_step(relax_model_t* m, const double* dt_override, relax_step_info_t* info) {
  if (!m || m->batch) return RELAX_ERR_ARG;
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
  if (!m || nsteps < 0 || m->batch) return RELAX_ERR_ARG;
  int k = 0;
  for (; k < nsteps; ++k) {
    HIP_TRY(hipMemcpyAsync(m->host_sc, m->sc, sizeof(StepScalars), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (m->host_sc->status) { if (done) *done = k; return m->host_sc->status; }
    if (m->host_sc->t >= m->desc.interval) break;
    const int e0 = m->host_sc->e;
    relax_step_info_t tmp;
    relax_step_info_t* cur = info ? info + k : &tmp;
    const int rc = relax_model_step(m, dt_override ? dt_override + k : nullptr, cur);
    if (rc) { if (done) *done = k; return rc; }
    // an event was reached: the caller applies its slip before going on
    if (cur->e != e0) { ++k; break; }
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

// The patch list m->shear_host and its slips in device form (grow-only)
static int stage_planes(relax_model_t* m, const double* slip) {
  const long n = (long)m->shear_host.size();
  if (n > m->shear_cap) {
    if (m->shear_dev) (void)rx_free(m->shear_dev);
    if (m->shear_slip) (void)rx_free(m->shear_slip);
    m->shear_dev = nullptr;
    m->shear_slip = nullptr;
    m->shear_cap = 0;
    HIP_TRY(rx_malloc(&m->shear_dev, sizeof(PlaneDev) * n));
    HIP_TRY(rx_malloc(&m->shear_slip, sizeof(double) * n));
    m->shear_cap = n;
  }
  HIP_TRY(hipMemcpyAsync(m->shear_dev, m->shear_host.data(), sizeof(PlaneDev) * n, hipMemcpyHostToDevice,
                         m->stream));
  HIP_TRY(hipMemcpyAsync(m->shear_slip, slip, sizeof(double) * n, hipMemcpyHostToDevice, m->stream));
  return RELAX_OK;
}

// momentdensityshear of the staged patches [k0, k0+n) added to tau
static int add_shear(relax_model_t* m, int k0, int n) {
  if (n == 0) return RELAX_OK;
  k_add_shear<<<grid_1d(m->g.H), 256, 0, m->stream>>>(m->shear_dev + k0, m->shear_slip + k0, n,
                                                      2.0f * (float)m->desc.mu, m->g, m->desc.dx1, m->desc.dx2,
                                                      m->desc.dx3, m->tau);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int relax_model_add_shear(relax_model_t* m, const relax_plane_t* patches, const double* slip, int n) {
  if (!m || n < 0 || (n > 0 && (!patches || !slip))) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  if (n == 0) return RELAX_OK;
  m->shear_host.resize(n);
  for (int k = 0; k < n; ++k) m->shear_host[k] = make_plane(patches[k], m->desc.beta, m->desc.dx1, m->desc.dx3);
  int rc = stage_planes(m, slip);
  if (!rc) rc = add_shear(m, 0, n);
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = RELAX_ERR_HIP;
  return rc;
}

// source, elastic3d.f90:2064-2252: one thread per point of the top half,
// the planes' contributions added in their order with the reference's
// rounding (two FP32 roundings per plane and component); threads of the top
// plane also add the surface traction (n sdyad mu s b) e3.
__global__ void k_source(const PlaneDev* __restrict__ P, const double* __restrict__ slip, int np, double mu,
                         int n1, int n2, int h3, double dx1, double dx2, double dx3, float* f, long ld, long fs,
                         float* t, long ts) {
  const long N = (long)n1 * n2 * h3;
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= N) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = (double)i3 * dx3;
  const long of = i1 + ld * (i2 + (long)n2 * i3), ot = i1 + ld * i2;
  float f1 = f[of], f2 = f[fs + of], f3 = f[2 * fs + of];
  float t1 = 0.f, t2 = 0.f, t3 = 0.f;
  if (i3 == 0) { t1 = t[ot]; t2 = t[ts + ot]; t3 = t[2 * ts + ot]; }
  for (int k = 0; k < np; ++k) {
    const PlaneDev& p = P[k];
    const double wl = fmax(p.Lp, p.Wp);
    if ((fabs(x1 - p.x) > wl) || (fabs(x2 - p.y) > wl)) continue;
    const double cs = p.cstrike, ss = p.sstrike, cd = p.cdip, sd = p.sdip;
    const double xx = cs * x1 - ss * x2;
    const double x1s = cd * xx - sd * x3, x1i = cd * xx + sd * x3;
    if ((fabs(x1s - p.xr) > 7.01 * dx1) && (fabs(x1i - p.xr) > 7.01 * dx1)) continue;
    const double x2s = ss * x1 + cs * x2, x3s = sd * xx + cd * x3, x3i = -sd * xx + cd * x3;
    const double cr = cos(p.rake), sr = sin(p.rake), s = slip[k];
    const double om2 = d_omega((x2s - p.yr) / p.W, p.beta);
    if (i3 == 0) {
      // the moment density at the surface (source and image apertures)
      const double ap = d_gauss(x1s - p.xr, dx1) * om2 * d_omega((x3s - p.zr) / p.L, p.beta) +
                        d_gauss(x1i - p.xr, dx1) * om2 * d_omega((x3i + p.zr) / p.L, p.beta);
      const double n[3] = {p.n[0] * ap, p.n[1] * ap, p.n[2] * ap};
      const double b[3] = {mu * s * p.r[0], mu * s * p.r[1], mu * s * p.r[2]};
      t1 = t1 + (float)((n[0] * b[2] + n[2] * b[0]) / 2.0);
      t2 = t2 + (float)((n[1] * b[2] + n[2] * b[1]) / 2.0);
      t3 = t3 + (float)(n[2] * b[2]);
    }
    if ((fabs(x3 - p.z) > p.Lp) && (fabs(x3 + p.z) > p.Lp)) continue;
    const double scale = -mu * s;
    const double om2p = d_omegap((x2s - p.yr) / p.W, p.beta);
    double g = d_gauss(x1s - p.xr, dx1);
    double om3 = d_omega((x3s - p.zr) / p.L, p.beta);
    const double sourc = scale * d_gaussp(x1s - p.xr, dx1) * om2 * om3;
    const double dblcp = scale / p.W * g * om2p * om3;
    const double dipcs = scale / p.L * g * om2 * d_omegap((x3s - p.zr) / p.L, p.beta);
    g = d_gauss(x1i - p.xr, dx1);
    om3 = d_omega((x3i + p.zr) / p.L, p.beta);
    const double image = scale * d_gaussp(x1i - p.xr, dx1) * om2 * om3;
    const double cplei = scale / p.W * g * om2p * om3;
    if (2.01 * RELAX_DEG2RAD > p.dip) {
      // near-vertical planes (dip measured from the vertical) use the image
      f1 = (float)((double)f1 + (cr * ss * (sourc + image) + cr * cd * cs * (dblcp + cplei)));
      f2 = (float)((double)f2 + (cr * cs * (sourc + image) - cr * cd * ss * (dblcp + cplei)));
      f3 = (float)((double)f3 + (-cr * sd * (dblcp - cplei)));
    } else {
      f1 = (float)((double)f1 + (cr * ss * sourc + cr * cd * cs * dblcp));
      f2 = (float)((double)f2 + (cr * cs * sourc - cr * cd * ss * dblcp));
      f3 = (float)((double)f3 + (-cr * sd * dblcp));
    }
    f1 = (float)((double)f1 + (cd * sr * cs * dipcs + sd * sr * cs * sourc));
    f2 = (float)((double)f2 + (-cd * sr * ss * dipcs - sd * sr * ss * sourc));
    f3 = (float)((double)f3 + (cd * sr * sourc - sd * sr * dipcs));
  }
  f[of] = f1; f[fs + of] = f2; f[2 * fs + of] = f3;
  if (i3 == 0) { t[ot] = t1; t[ts + ot] = t2; t[2 * ts + ot] = t3; }
}

int launch_source(const PlaneDev* P, const double* slip, int np, double mu, int n1, int n2, int h3, double dx1,
                  double dx2, double dx3, float* f, long ld, long fs, float* t, long ts, hipStream_t st) {
  const long N = (long)n1 * n2 * h3;
  k_source<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(P, slip, np, mu, n1, n2, h3, dx1, dx2, dx3, f, ld, fs, t,
                                                        ts);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// tensilesource, elastic3d.f90:2315-2411: one thread per point of the full
// depth (x3 from the shifted coordinates, so that the images above the
// surface land in the lower half of the grid), the cracks added in their
// order with one rounding each
__global__ void k_tensile_source(const PlaneDev* __restrict__ P, const double* __restrict__ opening, int np,
                                 double lambda, double mu, int n1, int n2, int n3, double dx1, double dx2, double dx3,
                                 float* f, long ld, long fs) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2 * n3) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = d_shifted(i3, n3, dx3);
  const long of = i1 + ld * (i2 + (long)n2 * i3);
  float f1 = f[of], f2 = f[fs + of], f3 = f[2 * fs + of];
  bool any = false;
  for (int k = 0; k < np; ++k) {
    const PlaneDev& p = P[k];
    if ((fabs(x3 - p.z) > p.Lp) && (fabs(x3 + p.z) > p.Lp)) continue;
    if ((fabs(x1 - p.x) > p.Wp) || (fabs(x2 - p.y) > p.Wp)) continue;
    const double cs = p.cstrike, ss = p.sstrike, cd = p.cdip, sd = p.sdip;
    const double xx = cs * x1 - ss * x2;
    const double x1s = cd * xx - sd * x3, x1i = cd * xx + sd * x3;
    if ((fabs(x1s - p.xr) > 7.01 * dx1) && (fabs(x1i - p.xr) > 7.01 * dx1)) continue;
    const double x2s = ss * x1 + cs * x2, x3s = sd * xx + cd * x3, x3i = -sd * xx + cd * x3;
    const double scale1 = -opening[k] * (lambda + 2.0 * mu), scale2 = -opening[k] * lambda;
    double t1 = d_gauss(x1s - p.xr, dx1);
    const double t2 = d_omega((x2s - p.yr) / p.W, p.beta);
    double t3 = d_omega((x3s - p.zr) / p.L, p.beta);
    const double sourc = scale1 * d_gaussp(x1s - p.xr, dx1) * t2 * t3;
    const double dblcp = scale2 / p.W * t1 * d_omegap((x2s - p.yr) / p.W, p.beta) * t3;
    const double dipcs = scale2 / p.L * t1 * t2 * d_omegap((x3s - p.zr) / p.L, p.beta);
    t1 = d_gauss(x1i - p.xr, dx1);
    t3 = d_omega((x3i + p.zr) / p.L, p.beta);
    const double image = scale1 * d_gaussp(x1i - p.xr, dx1) * t2 * t3;
    const double cplei = scale2 / p.W * t1 * d_omegap((x2s - p.yr) / p.W, p.beta) * t3;
    const double dipci = scale2 / p.L * t1 * t2 * d_omegap((x3i + p.zr) / p.L, p.beta);
    f1 = (float)((double)f1 + cs * cd * (sourc + image) + ss * (dblcp + cplei) + cs * sd * (dipcs + dipci));
    f2 = (float)((double)f2 - ss * cd * (sourc + image) + cs * (dblcp + cplei) - ss * sd * (dipcs + dipci));
    f3 = (float)((double)f3 - sd * (sourc - image) + cd * (dipcs - dipci));
    any = true;
  }
  if (any) { f[of] = f1; f[fs + of] = f2; f[2 * fs + of] = f3; }
}

// mogisource, elastic3d.f90:2458-2509: point dilatation sources (volume
// change in slip) and their images over the full depth
__global__ void k_mogi_source(const relax_source_t* __restrict__ M, int nm, double lambda, double mu, int n1, int n2,
                              int n3, double dx1, double dx2, double dx3, float* f, long ld, long fs) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2 * n3) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = d_shifted(i3, n3, dx3);
  const double Wp = 6.0 * fmax(dx1, fmax(dx2, dx3)), Lp = Wp;
  const long of = i1 + ld * (i2 + (long)n2 * i3);
  float f1 = f[of], f2 = f[fs + of], f3 = f[2 * fs + of];
  bool any = false;
  for (int k = 0; k < nm; ++k) {
    const relax_source_t& s = M[k];
    if ((fabs(x3 - s.z) > Lp) && (fabs(x3 + s.z) > Lp)) continue;
    if ((fabs(x1 - s.x) > Wp) || (fabs(x2 - s.y) > Wp)) continue;
    const double scale = -(lambda + 2.0 * mu / 3.0) * s.slip;
    const double t1 = d_gauss(x1 - s.x, dx1), t2 = d_gauss(x2 - s.y, dx2);
    double t3 = d_gauss(x3 - s.z, dx3);
    const double s1 = scale * d_gaussp(x1 - s.x, dx1) * t2 * t3, s2 = scale * t1 * d_gaussp(x2 - s.y, dx2) * t3;
    const double s3 = scale * t1 * t2 * d_gaussp(x3 - s.z, dx3);
    t3 = d_gauss(x3 + s.z, dx3);
    const double m1 = scale * d_gaussp(x1 - s.x, dx1) * t2 * t3, m2 = scale * t1 * d_gaussp(x2 - s.y, dx2) * t3;
    const double m3 = scale * t1 * t2 * d_gaussp(x3 + s.z, dx3);
    f1 = (float)((double)f1 + (s1 + m1));
    f2 = (float)((double)f2 + (s2 + m2));
    f3 = (float)((double)f3 + (s3 - m3));
    any = true;
  }
  if (any) { f[of] = f1; f[fs + of] = f2; f[2 * fs + of] = f3; }
}

// momentdensitytensile, elastic3d.f90:3137-3215: Hooke's law of
// n sdyad (opening n), n carrying the aperture; each crack adds in turn
__global__ void k_add_tensile(const PlaneDev* __restrict__ P, const double* __restrict__ opening, int np,
                              double lambda, double mu, int n1, int n2, int h3, double dx1, double dx2, double dx3,
                              float* sig, long H) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2 * h3) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = (double)i3 * dx3;
  const float m2 = (float)(2.0 * mu);
  float acc[6];
  bool any = false;
  for (int k = 0; k < np; ++k) {
    const PlaneDev& p = P[k];
    if (fabs(x3 - p.z) > p.Lp) continue;
    if ((fabs(x1 - p.x) > p.Wp) || (fabs(x2 - p.y) > p.Wp)) continue;
    const double xx = p.cstrike * x1 - p.sstrike * x2;
    const double x1s = p.cdip * xx - p.sdip * x3, x1i = p.cdip * xx + p.sdip * x3;
    if ((fabs(x1s - p.xr) > 7.01 * dx1) && (fabs(x1i - p.xr) > 7.01 * dx1)) continue;
    const double x2s = p.sstrike * x1 + p.cstrike * x2;
    const double x3s = p.sdip * xx + p.cdip * x3, x3i = -p.sdip * xx + p.cdip * x3;
    const double t2 = d_omega((x2s - p.yr) / p.W, p.beta);
    double ap = d_gauss(x1s - p.xr, dx1) * t2 * d_omega((x3s - p.zr) / p.L, p.beta);
    ap = ap + d_gauss(x1i - p.xr, dx1) * t2 * d_omega((x3i + p.zr) / p.L, p.beta);
    const double n[3] = {p.n[0] * ap, p.n[1] * ap, p.n[2] * ap};
    const double b[3] = {opening[k] * n[0], opening[k] * n[1], opening[k] * n[2]};
    float e[6] = {(float)(n[0] * b[0]), (float)((n[0] * b[1] + n[1] * b[0]) / 2.0),
                  (float)((n[0] * b[2] + n[2] * b[0]) / 2.0), (float)(n[1] * b[1]),
                  (float)((n[1] * b[2] + n[2] * b[1]) / 2.0), (float)(n[2] * b[2])};
    // isotropicstressstrain, elastic3d.f90:129-142
    const double tr = (double)(e[0] + e[3] + e[5]);
    for (float& v : e) v = m2 * v;
    e[0] = (float)((double)e[0] + lambda * tr);
    e[3] = (float)((double)e[3] + lambda * tr);
    e[5] = (float)((double)e[5] + lambda * tr);
    if (!any)
      for (int c = 0; c < 6; ++c) acc[c] = sig[c * H + o];
    for (int c = 0; c < 6; ++c) acc[c] = acc[c] + e[c];
    any = true;
  }
  if (any)
    for (int c = 0; c < 6; ++c) sig[c * H + o] = acc[c];
}

// momentdensitymogi, elastic3d.f90:3240-3286: kappa o G on the diagonal
__global__ void k_add_mogi(const relax_source_t* __restrict__ M, int nm, double lambda, double mu, int n1, int n2,
                           int h3, double dx1, double dx2, double dx3, float* sig, long H) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2 * h3) return;
  const int i1 = o % n1, i2 = (o / n1) % n2, i3 = o / ((long)n1 * n2);
  const double x1 = d_shifted(i1, n1, dx1), x2 = d_shifted(i2, n2, dx2), x3 = (double)i3 * dx3;
  const double kappa = lambda + 2.0 / 3.0 * mu;
  const double Wp = 6.0 * fmax(dx1, fmax(dx2, dx3)), Lp = Wp;
  float a11 = 0.f, a22 = 0.f, a33 = 0.f;
  bool any = false;
  for (int k = 0; k < nm; ++k) {
    const relax_source_t& s = M[k];
    if (fabs(x3 - s.z) > Lp) continue;
    if ((fabs(x1 - s.x) > Wp) || (fabs(x2 - s.y) > Wp)) continue;
    double g = s.slip * d_gauss(x1 - s.x, dx1) * d_gauss(x2 - s.y, dx2) * d_gauss(x3 - s.z, dx3);
    g = kappa * g;
    const float m = (float)g;
    if (!any) { a11 = sig[o]; a22 = sig[3 * H + o]; a33 = sig[5 * H + o]; }
    a11 = a11 + m; a22 = a22 + m; a33 = a33 + m;
    any = true;
  }
  if (any) { sig[o] = a11; sig[3 * H + o] = a22; sig[5 * H + o] = a33; }
}

static unsigned blocks_of(long n) { return (unsigned)((n + 255) / 256); }

int launch_tensile_source(const PlaneDev* P, const double* opening, int np, double lambda, double mu, int n1, int n2,
                          int n3, double dx1, double dx2, double dx3, float* f, long ld, long fs, hipStream_t st) {
  k_tensile_source<<<blocks_of((long)n1 * n2 * n3), 256, 0, st>>>(P, opening, np, lambda, mu, n1, n2, n3, dx1, dx2,
                                                                  dx3, f, ld, fs);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
int launch_mogi_source(const relax_source_t* M, int nm, double lambda, double mu, int n1, int n2, int n3,
                       double dx1, double dx2, double dx3, float* f, long ld, long fs, hipStream_t st) {
  k_mogi_source<<<blocks_of((long)n1 * n2 * n3), 256, 0, st>>>(M, nm, lambda, mu, n1, n2, n3, dx1, dx2, dx3, f, ld,
                                                               fs);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
int launch_add_tensile(const PlaneDev* P, const double* opening, int np, double lambda, double mu, int n1, int n2,
                       int h3, double dx1, double dx2, double dx3, float* sig, long H, hipStream_t st) {
  k_add_tensile<<<blocks_of((long)n1 * n2 * h3), 256, 0, st>>>(P, opening, np, lambda, mu, n1, n2, h3, dx1, dx2, dx3,
                                                               sig, H);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
int launch_add_mogi(const relax_source_t* M, int nm, double lambda, double mu, int n1, int n2, int h3, double dx1,
                    double dx2, double dx3, float* sig, long H, hipStream_t st) {
  k_add_mogi<<<blocks_of((long)n1 * n2 * h3), 256, 0, st>>>(M, nm, lambda, mu, n1, n2, h3, dx1, dx2, dx3, sig, H);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// traction, elastic3d.f90:2517-2590: one thread per surface point, the
// loads subtracted in their order, each rounded to FP32 as the reference's
// REAL() does
__device__ __forceinline__ float traction_point(const relax_source_t* __restrict__ l, int nl, double x1, double x2,
                                                double t, double Dt, int rate, float v) {
  for (int k = 0; k < nl; ++k) {
    const relax_source_t& L = l[k];
    const double Lp = L.length * (1.0 + 2.0 * L.beta) / 2.0, Wp = L.width * (1.0 + 2.0 * L.beta) / 2.0;
    if (Lp == 0.0 || Wp == 0.0) continue;
    if ((fabs(x1 - L.x) > fmax(Lp, Wp)) || (fabs(x2 - L.y) > fmax(Lp, Wp))) continue;
    const double amp = d_omega((x1 - L.x) / L.length, L.beta) * d_omega((x2 - L.y) / L.width, L.beta) * L.slip;
    if (rate) {
      if (L.period != 0.0)
        v = (float)((double)v - amp * (sin(2 * RELAX_PI * (t + Dt) / L.period + L.phase) -
                                       sin(2 * RELAX_PI * t / L.period + L.phase)) / Dt);
    } else if (L.period == 0.0) {
      v = (float)((double)v - amp);
    } else {
      v = (float)((double)v - amp * sin(2 * RELAX_PI * t / L.period + L.phase));
    }
  }
  return v;
}

__global__ void k_traction(const relax_source_t* __restrict__ l, int nl, int n1, int n2, double dx1, double dx2,
                           double t, double Dt, int rate, float* t3, long ld) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2) return;
  const int i1 = o % n1, i2 = o / n1;
  t3[i1 + ld * i2] = traction_point(l, nl, d_shifted(i1, n1, dx1), d_shifted(i2, n2, dx2), t, Dt, rate,
                                    t3[i1 + ld * i2]);
}

// The rate of a model's time-dependent surface loads in a stage of the RK2
// step (relax.f90:684-685, :805-806): over [t, t + Dt] in the corrector,
// over [t, t + Dt/2e8] in the predictor (the reference's step there), with
// t and Dt of the step from the device
__global__ void k_traction_rate(const relax_source_t* __restrict__ l, int nl, int n1, int n2, double dx1, double dx2,
                                const StepScalars* __restrict__ sc, int pred, float* t3) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= (long)n1 * n2) return;
  const int i1 = o % n1, i2 = o / n1;
  const double dt = pred ? sc->Dt / 2.e8 : sc->Dt;
  t3[o] = traction_point(l, nl, d_shifted(i1, n1, dx1), d_shifted(i2, n2, dx2), sc->t, dt, 1, t3[o]);
}

int launch_traction(const relax_source_t* l, int nl, int n1, int n2, double dx1, double dx2, double t, double Dt,
                    int rate, float* t3, long ld, hipStream_t st) {
  const long N = (long)n1 * n2;
  k_traction<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(l, nl, n1, n2, dx1, dx2, t, Dt, rate, t3, ld);
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// A SOURCE_STRUCT patch in device form (its own taper beta)
static PlaneDev source_plane(const relax_source_t& s, double dx1, double dx3) {
  const relax_plane_t p{s.x, s.y, s.z, s.width, s.length, s.strike, s.dip, s.rake, 0, 0};
  return make_plane(p, s.beta, dx1, dx3);
}

// u += the top half of a full-depth field in the padded layout
__global__ void k_add_padded_top(float* __restrict__ u, const float* __restrict__ v, int n1, int n2, int h3, long H,
                                 long vs) {
  const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (o >= H) return;
  const int i1 = o % n1;
  const long r = o / n1;  // row (i2, i3)
  const long q = i1 + (long)(n1 + 2) * r;
  for (int c = 0; c < 3; ++c) u[c * H + o] = u[c * H + o] + v[c * vs + q];
}

int relax_model_event(relax_model_t* m, const relax_event_t* ev, double nyquist) {
  if (!m || m->batch || !ev || ev->ns < 0 || ev->nt < 0 || ev->nm < 0 || ev->nl < 0 || (ev->ns > 0 && !ev->s) ||
      (ev->nt > 0 && !ev->ts) || (ev->nm > 0 && !ev->m) || (ev->nl > 0 && !ev->l) || ev->nc < 0 || ev->nth < 0 ||
      ev->nv < 0 || (ev->nc > 0 && !ev->c) || (ev->nth > 0 && (!ev->th || !ev->v)))
    return RELAX_ERR_ARG;
  int rc;
  if ((rc = sync_md(m))) return rc;
  // the eigenstrain volumes (relax.f90:1151-1178, :1192-1223)
  if ((rc = eigen_stage(m, ev))) return rc;
  const Grid& g = m->g;
  const long H = g.H;
  const int ns = ev->ns, nt = ev->nt, nm = ev->nm, nl = ev->nl;
  // dislocations (relax.f90:1116-1289): the body forces of the shear
  // patches (source(), :1140-1150) and tensile cracks (tensilesource(),
  // :1228-1239) longer or wider than nyquist*min(dx) and of every point
  // dilatation source (mogisource(), :1262-1268); every source's moment
  // density into tau.  The staged plane list holds the large shear patches,
  // all shear patches, the large cracks and all cracks, in that order.
  const double dmin = std::min(m->desc.dx1, std::min(m->desc.dx2, m->desc.dx3)) * nyquist;
  std::vector<PlaneDev>& hp = m->shear_host;
  std::vector<double> amp;
  hp.clear();
  auto push = [&](const relax_source_t& x, double beta, double a) {
    relax_source_t y = x;
    y.beta = beta;
    hp.push_back(source_plane(y, m->desc.dx1, m->desc.dx3));
    amp.push_back(a);
  };
  for (int k = 0; k < ns; ++k)
    if (dmin < ev->s[k].length || dmin < ev->s[k].width) push(ev->s[k], ev->s[k].beta, ev->s[k].slip);
  const int nbig = (int)hp.size();
  for (int k = 0; k < ns; ++k) push(ev->s[k], ev->s[k].beta, ev->s[k].slip);
  // tensile cracks take the model's taper (dislocations passes its beta)
  const int t0 = (int)hp.size();
  for (int k = 0; k < nt; ++k)
    if (dmin < ev->ts[k].length || dmin < ev->ts[k].width) push(ev->ts[k], m->desc.beta, ev->ts[k].opening);
  const int ntbig = (int)hp.size() - t0;
  for (int k = 0; k < nt; ++k) push(ev->ts[k], m->desc.beta, ev->ts[k].opening);
  if (!hp.empty() && (rc = stage_planes(m, amp.data()))) return rc;
  // the surface loads and the point sources, staged together
  m->cur_nloads = 0;
  if (nl + nm > 0) {
    if (nl + nm > m->loads_cap) {
      if (m->loads_dev) (void)rx_free(m->loads_dev);
      m->loads_dev = nullptr;
      m->loads_cap = 0;
      HIP_TRY(rx_malloc(&m->loads_dev, sizeof(relax_source_t) * (nl + nm)));
      m->loads_cap = nl + nm;
    }
    if (nl) HIP_TRY(hipMemcpyAsync(m->loads_dev, ev->l, sizeof(relax_source_t) * nl, hipMemcpyHostToDevice, m->stream));
    // from now on the steps add the rate of this event's loads
    m->cur_nloads = nl;
    if (nm)
      HIP_TRY(hipMemcpyAsync(m->loads_dev + nl, ev->m, sizeof(relax_source_t) * nm, hipMemcpyHostToDevice, m->stream));
  }
  double t = 0.0;  // the model's time, for the loads (relax.f90:319, :874)
  HIP_TRY(hipMemcpyAsync(&t, &m->sc->t, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  const double dx1 = m->desc.dx1, dx2 = m->desc.dx2, dx3 = m->desc.dx3, lam = m->desc.lambda, mu = m->desc.mu;
  if (nt == 0 && nm == 0) {
    // body forces in the top half only: the pruned Green solve of the step,
    // adding to u (relax.f90:362-371, :877-884)
    HIP_TRY(hipMemsetAsync(m->W, 0, 3 * H * 4, m->stream));
    HIP_TRY(hipMemsetAsync(m->T, 0, 3L * g.n1 * g.n2 * 4, m->stream));
    if (nbig && (rc = launch_source(m->shear_dev, m->shear_slip, nbig, mu, g.n1, g.n2, g.h3, dx1, dx2, dx3, m->W,
                                    g.n1, H, m->T, (long)g.n1 * g.n2, m->stream)))
      return rc;
    if ((rc = eigen_body_force(m, m->W, g.n1, H))) return rc;
    if (nl && (rc = launch_traction(m->loads_dev, nl, g.n1, g.n2, dx1, dx2, t, 0.0, 0, m->T + 2L * g.n1 * g.n2, g.n1,
                                    m->stream)))
      return rc;
    Ctx c = single(m);
    c.prof = nullptr;
    if ((rc = model_green(c, EPI_ADD))) return rc;
  } else {
    // tensile cracks and point sources put images above the surface, in the
    // lower half of the grid: the full-depth Green solve on scratch fields
    // in the padded layout
    const long ld = g.n1 + 2, vs = ld * g.n2 * g.n3, ts = ld * g.n2;
    float* scratch = nullptr;
    const size_t bytes = (size_t)(3 * vs + 6 * ts) * 4;
    HIP_TRY(rx_malloc(&scratch, bytes));
    float* v[3] = {scratch, scratch + vs, scratch + 2 * vs};
    float* tr[3] = {scratch + 3 * vs, scratch + 3 * vs + ts, scratch + 3 * vs + 2 * ts};
    float* pp[3] = {scratch + 3 * vs + 3 * ts, scratch + 3 * vs + 4 * ts, scratch + 3 * vs + 5 * ts};
    rc = hipMemsetAsync(scratch, 0, bytes, m->stream) == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
    if (!rc && nbig)
      rc = launch_source(m->shear_dev, m->shear_slip, nbig, mu, g.n1, g.n2, g.h3, dx1, dx2, dx3, v[0], ld, vs, tr[0],
                         ts, m->stream);
    if (!rc) rc = eigen_body_force(m, v[0], ld, vs);
    if (!rc && ntbig)
      rc = launch_tensile_source(m->shear_dev + t0, m->shear_slip + t0, ntbig, lam, mu, g.n1, g.n2, g.n3, dx1, dx2,
                                 dx3, v[0], ld, vs, m->stream);
    if (!rc && nm)
      rc = launch_mogi_source(m->loads_dev + nl, nm, lam, mu, g.n1, g.n2, g.n3, dx1, dx2, dx3, v[0], ld, vs,
                              m->stream);
    if (!rc && nl)
      rc = launch_traction(m->loads_dev, nl, g.n1, g.n2, dx1, dx2, t, 0.0, 0, tr[2], ld, m->stream);
    if (!rc && hipStreamSynchronize(m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
    if (!rc) rc = green_full_device(v, tr, pp, g.n1, g.n2, g.n3, dx1, dx2, dx3, lam, mu, m->desc.gamma, m->stream, true);
    if (!rc) {
      k_add_padded_top<<<blocks_of(H), 256, 0, m->stream>>>(m->u, scratch, g.n1, g.n2, g.h3, H, vs);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess) rc = RELAX_ERR_HIP;
    }
    (void)hipStreamSynchronize(m->stream);
    (void)rx_free(scratch);
    if (rc) return rc;
  }
  // the moment densities of every source into tau, then
  // sig = -tau + C:eps(u) (relax.f90:887-889)
  if ((rc = add_shear(m, nbig, ns))) return rc;
  if ((rc = eigen_moment(m))) return rc;
  if (nt && (rc = launch_add_tensile(m->shear_dev + t0 + ntbig, m->shear_slip + t0 + ntbig, nt, lam, mu, g.n1, g.n2,
                                     g.h3, dx1, dx2, dx3, m->tau, H, m->stream)))
    return rc;
  if (nm && (rc = launch_add_mogi(m->loads_dev + nl, nm, lam, mu, g.n1, g.n2, g.h3, dx1, dx2, dx3, m->tau, H,
                                  m->stream)))
    return rc;
  Ctx c = single(m);
  c.prof = nullptr;
  if ((rc = launch_stress(STRESS_SET, g, m->S, c.md, c.one, 1, RHEO_VISCOUS, m->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

// Inter-seismic loading (in%inter, input.f90:1114-1250; relax.f90:669-674,
// :795-799): the moment density of the loading segments -- shear segments
// with their own taper, tensile segments with the model's -- is formed once
// on the device; every corrector adds it to the moment of the rheology
// before the divergence, and to tau unscaled as well as times Dt.  The
// model then runs the extended stencil class.
int relax_model_set_interseismic(relax_model_t* m, const relax_event_t* ev) {
  if (!m || m->batch) return RELAX_ERR_ARG;
  if (ev && (ev->ns < 0 || ev->nt < 0 || ev->nm != 0 || ev->nl != 0 || ev->nc != 0 || ev->nth != 0 ||
             (ev->ns > 0 && !ev->s) || (ev->nt > 0 && !ev->ts)))
    return RELAX_ERR_ARG;
  const int ns = ev ? ev->ns : 0, nt = ev ? ev->nt : 0;
  int rc;
  if (ns + nt == 0) {
    m->R.inter = nullptr;
    m->inter = 0;
    m->md_dirty = true;
    return sync_md(m);
  }
  const Grid& g = m->g;
  const long H = g.H;
  if (!m->mi) HIP_TRY(rx_malloc(&m->mi, 6 * H * sizeof(float)));
  HIP_TRY(hipMemsetAsync(m->mi, 0, 6 * H * sizeof(float), m->stream));
  std::vector<PlaneDev>& hp = m->shear_host;
  std::vector<double> amp;
  hp.clear();
  for (int k = 0; k < ns; ++k) {
    hp.push_back(source_plane(ev->s[k], m->desc.dx1, m->desc.dx3));
    amp.push_back(ev->s[k].slip);
  }
  for (int k = 0; k < nt; ++k) {
    relax_source_t y = ev->ts[k];
    y.beta = m->desc.beta;
    hp.push_back(source_plane(y, m->desc.dx1, m->desc.dx3));
    amp.push_back(ev->ts[k].opening);
  }
  if ((rc = stage_planes(m, amp.data()))) return rc;
  const double dx1 = m->desc.dx1, dx2 = m->desc.dx2, dx3 = m->desc.dx3;
  if (ns) {
    k_add_shear<<<grid_1d(H), 256, 0, m->stream>>>(m->shear_dev, m->shear_slip, ns, 2.0f * (float)m->desc.mu, g,
                                                   dx1, dx2, dx3, m->mi);
    if (hipGetLastError() != hipSuccess) return RELAX_ERR_HIP;
  }
  if (nt && (rc = launch_add_tensile(m->shear_dev + ns, m->shear_slip + ns, nt, m->desc.lambda, m->desc.mu, g.n1,
                                     g.n2, g.h3, dx1, dx2, dx3, m->mi, H, m->stream)))
    return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->R.inter = m->mi;
  m->inter = 1;
  m->md_dirty = true;
  return sync_md(m);
}

// sig = -tau + C:eps(u) on the model's stream (relax.f90:369-371)
int model_stress_set(relax_model_t* m) {
  int rc;
  if ((rc = sync_md(m))) return rc;
  Ctx c = single(m);
  return launch_stress(STRESS_SET, m->g, m->S, c.md, c.one, 1, RHEO_VISCOUS, m->stream);
}

int relax_model_coseismic(relax_model_t* m) {
  if (!m) return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  int rc;
  if ((rc = sync_md(m))) return rc;
  Ctx c = single(m);
  c.prof = nullptr;
  // body force and surface traction of tau, then u = G[f, t] and
  // sig = -tau + C:eps(u) (relax.f90:318-371 with the FIR divergence)
  if ((rc = launch_bodyforce(m->g, m->S, c.md, c.one, 1, RHEO_VISCOUS, BODY_RAW, m->stream))) return rc;
  if ((rc = model_green(c, EPI_PLAIN))) return rc;
  if ((rc = launch_stress(STRESS_SET, m->g, m->S, c.md, c.one, 1, RHEO_VISCOUS, m->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RELAX_OK;
}

// ------------------------------------------------------------- batches ----
struct relax_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<relax_model_t*> models;
  std::vector<unsigned> versions;   // member record versions in md_dev
  std::vector<ModelDev> md_host;
  ModelDev* md_dev = nullptr;       // one record per member
  StepScalars* gather = nullptr;    // device copies of the members' scalars
  StepScalars* host_sc = nullptr;   // pinned
  int cap = 0;
  bool dirty = true;                // membership or a record changed
  int stable = 0;                   // steps since the last change
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool tuned = false;               // tune holds relax_batch_autotune's choice
  int tune[3] = {0, 0, 0};
  int tune_grid[3] = {0, 0, 0};     // the grid it was made for
  // Members of several grid classes (grid, spacing, elastic constants): the
  // records in md_dev are grouped by class, member i's at rec[i]; `ordered`
  // lists the members in record order.  One class: rec is the identity.
  struct ClassRange {
    int off, n;
  };
  std::vector<int> rec;
  std::vector<relax_model_t*> ordered;
  std::vector<ClassRange> classes;
  // mixed batches: the work lists of the strain pass (entries [0, nm]) and
  // of the divergence pass ([nm + 1, 2 nm + 1]) over all members
  WorkItem* wl = nullptr;
  int wl_cap = 0;
  long wl_stress = 0, wl_body = 0;
  int body_variant = -1;
};

// A membership that changes every step or two (a server admitting and
// retiring requests) runs its first steps eagerly; the step is captured as a
// graph once the membership has been stable for this many steps.
static const int GRAPH_AFTER = 2;
// Batches of this many grid points or more run eagerly: their launches are
// long enough that replaying a graph saves nothing, while a capture (after
// every membership change) costs the host a few milliseconds.
static const long GRAPH_MAX_POINTS = 1L << 25;

static std::recursive_mutex g_capture_mu;
void capture_lock() { g_capture_mu.lock(); }
void capture_unlock() { g_capture_mu.unlock(); }

__global__ void k_gather_scalars(const ModelDev* __restrict__ md, int nm, StepScalars* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nm) out[i] = *md[i].sc;
}

static void drop_graph(relax_batch_t* b) {
  CaptureGuard guard;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static bool batch_compatible(const relax_model_t* a, const relax_model_t* m) {
  const relax_model_desc_t &x = a->desc, &y = m->desc;
  return a->device == m->device && x.sx1 == y.sx1 && x.sx2 == y.sx2 && x.sx3 == y.sx3 && x.dx1 == y.dx1 &&
         x.dx2 == y.dx2 && x.dx3 == y.dx3 && x.lambda == y.lambda && x.mu == y.mu && x.gamma == y.gamma;
}

int relax_batch_create(relax_batch_t** out) {
  if (!out) return RELAX_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RELAX_ERR_NODEVICE;
  relax_batch_t* b = new (std::nothrow) relax_batch_t();
  if (!b) return RELAX_ERR_NOMEM;
  if (hipGetDevice(&b->device) != hipSuccess ||
      rx_stream_create(&b->stream) != hipSuccess) {
    delete b;
    return RELAX_ERR_HIP;
  }
  *out = b;
  return RELAX_OK;
}

int relax_batch_destroy(relax_batch_t* b) {
  if (!b) return RELAX_ERR_ARG;
  (void)hipStreamSynchronize(b->stream);
  for (relax_model_t* m : b->models) m->batch = nullptr;
  drop_graph(b);
  if (b->wl) (void)rx_free(b->wl);
  if (b->md_dev) (void)rx_free(b->md_dev);
  if (b->gather) (void)rx_free(b->gather);
  if (b->host_sc) (void)rx_host_free(b->host_sc);
  (void)rx_stream_destroy(b->stream);
  delete b;
  return RELAX_OK;
}

int relax_batch_add(relax_batch_t* b, relax_model_t* m) {
  if (!b || !m || m->batch) return RELAX_ERR_ARG;
  if (m->device != b->device) return RELAX_ERR_ARG;
  int rc;
  if ((rc = sync_md(m))) return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));  // the model's own work is done before the batch's
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->models.push_back(m);
  m->batch = b;
  b->dirty = true;
  return RELAX_OK;
}

int relax_batch_remove(relax_batch_t* b, relax_model_t* m) {
  if (!b || !m || m->batch != b) return RELAX_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->models.size(); ++i)
    if (b->models[i] == m) { b->models.erase(b->models.begin() + i); break; }
  m->batch = nullptr;
  b->dirty = true;
  return RELAX_OK;
}

int relax_batch_size(const relax_batch_t* b) { return b ? (int)b->models.size() : 0; }

int relax_batch_synchronize(relax_batch_t* b) {
  if (!b) return RELAX_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RELAX_OK;
}

// Group the members by grid class (first-seen order of the classes, the
// members' order within each): rec, ordered and classes.
static void batch_classes(relax_batch_t* b) {
  const int nm = (int)b->models.size();
  std::vector<int> cls(nm, -1);
  std::vector<int> leads;  // a member of each class
  for (int i = 0; i < nm; ++i) {
    for (size_t k = 0; k < leads.size(); ++k)
      if (batch_compatible(b->models[leads[k]], b->models[i])) { cls[i] = (int)k; break; }
    if (cls[i] < 0) { cls[i] = (int)leads.size(); leads.push_back(i); }
  }
  b->rec.assign(nm, 0);
  b->ordered.clear();
  b->classes.clear();
  for (int k = 0; k < (int)leads.size(); ++k) {
    relax_batch_t::ClassRange r{(int)b->ordered.size(), 0};
    for (int i = 0; i < nm; ++i)
      if (cls[i] == k) { b->rec[i] = (int)b->ordered.size(); b->ordered.push_back(b->models[i]); ++r.n; }
    b->classes.push_back(r);
  }
}

static int batch_rclass(const relax_batch_t* b) {
  int rc = RHEO_VISCOUS;
  for (relax_model_t* m : b->models) rc = std::max(rc, rheo_class(m));
  return rc;
}

static bool tune_entry(int dev, const Grid& g, int rclass, TuneChoice* out);

// The divergence variant of a mixed batch: the choice for its largest class
// (relax_batch_autotune's table, else that class's first member's), if it
// tiles every member's grid, else the 32x8 tiles (variant 1), which tile
// every grid a model can have.
static int mixed_body_variant(const relax_batch_t* b, int rclass) {
  long best = -1;
  const relax_model_t* lead = nullptr;
  for (const auto& k : b->classes) {
    const long pts = (long)k.n * b->ordered[k.off]->g.H;
    if (pts > best) { best = pts; lead = b->ordered[k.off]; }
  }
  TuneChoice t;
  int v = tune_entry(b->device, lead->g, rclass, &t) ? t.body : lead->tune_body;
  for (const relax_model_t* m : b->models)
    if (!body_variant_fits(m->g, v)) return 1;
  return v;
}

// Bring the device records up to date; a change drops the captured step.
static int batch_refresh(relax_batch_t* b) {
  const int nm = (int)b->models.size();
  bool changed = b->dirty || (int)b->versions.size() != nm;
  for (int i = 0; i < nm; ++i) {
    relax_model_t* m = b->models[i];
    if (m->md_dirty) {
      int rc = sync_md(m);
      if (rc) return rc;
    }
    if (!changed && b->versions[i] != m->md_version) changed = true;
  }
  // a mixed batch's divergence variant follows the tuning table
  if (!changed && b->classes.size() > 1 && mixed_body_variant(b, batch_rclass(b)) != b->body_variant)
    changed = true;
  if (!changed) return RELAX_OK;
  HIP_TRY(hipStreamSynchronize(b->stream));
  drop_graph(b);
  if (nm > b->cap) {
    if (b->md_dev) { (void)rx_free(b->md_dev); b->md_dev = nullptr; }
    if (b->gather) { (void)rx_free(b->gather); b->gather = nullptr; }
    if (b->host_sc) { (void)rx_host_free(b->host_sc); b->host_sc = nullptr; }
    b->cap = 0;
    const int cap = nm < 8 ? 8 : 2 * nm;  // room to grow without reallocating
    HIP_TRY(rx_malloc(&b->md_dev, sizeof(ModelDev) * cap));
    HIP_TRY(rx_malloc(&b->gather, sizeof(StepScalars) * cap));
    HIP_TRY(rx_host_malloc(&b->host_sc, sizeof(StepScalars) * cap));
    b->cap = cap;
  }
  batch_classes(b);
  b->md_host.resize(nm);
  b->versions.resize(nm);
  for (int i = 0; i < nm; ++i) {
    b->md_host[b->rec[i]] = b->models[i]->md_host;
    b->versions[i] = b->models[i]->md_version;
  }
  HIP_TRY(hipMemcpy(b->md_dev, b->md_host.data(), sizeof(ModelDev) * nm, hipMemcpyHostToDevice));
  b->body_variant = -1;
  if (b->classes.size() > 1) {
    // the work lists, in record order
    const int v = mixed_body_variant(b, batch_rclass(b));
    std::vector<WorkItem> wl(2 * (nm + 1));
    long ns = 0, nb = 0;
    for (int r = 0; r <= nm; ++r) {
      WorkItem& a = wl[r];
      WorkItem& c = wl[nm + 1 + r];
      a.first = ns;
      c.first = nb;
      if (r == nm) break;
      const relax_model_t* m = b->ordered[r];
      a.g = c.g = m->g;
      a.S = c.S = m->S;
      ns += stress_blocks(m->g);
      nb += body_items(m->g, v);
    }
    if ((int)wl.size() > b->wl_cap) {
      if (b->wl) { (void)rx_free(b->wl); b->wl = nullptr; }
      b->wl_cap = 0;
      HIP_TRY(rx_malloc(&b->wl, sizeof(WorkItem) * wl.size() * 2));
      b->wl_cap = (int)wl.size() * 2;
    }
    HIP_TRY(hipMemcpy(b->wl, wl.data(), sizeof(WorkItem) * wl.size(), hipMemcpyHostToDevice));
    b->wl_stress = ns;
    b->wl_body = nb;
    b->body_variant = v;
  }
  b->dirty = false;
  b->stable = 0;
  return RELAX_OK;
}

// Launch shapes of batches, per device, grid and rheology class, found by
// relax_batch_autotune and used by every later batch of that grid and class in
// the process.  A batch without an entry launches with its first member's
// shapes (the single-model choice).
struct BatchTuneKey {
  int dev, n1, n2, n3, rclass;
  bool operator<(const BatchTuneKey& o) const {
    return dev != o.dev ? dev < o.dev
           : n1 != o.n1 ? n1 < o.n1
           : n2 != o.n2 ? n2 < o.n2
           : n3 != o.n3 ? n3 < o.n3
                        : rclass < o.rclass;
  }
};
static std::map<BatchTuneKey, TuneChoice> g_btune;  // guarded by g_tune_mu

static bool tune_entry(int dev, const Grid& g, int rclass, TuneChoice* out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_btune.find(BatchTuneKey{dev, g.n1, g.n2, g.n3, rclass});
  if (it == g_btune.end()) return false;
  *out = it->second;
  return true;
}

// The launch context of grid class k of a batch (records off .. off+n-1):
// the batch's own choice when it was made for this grid, else the table's
// for the grid and the batch's rheology class, else the class's first
// member's shapes.
static Ctx class_ctx(relax_batch_t* b, const relax_batch_t::ClassRange& k, int rclass) {
  Ctx c;
  c.lead = b->ordered[k.off]; c.md = b->md_dev + k.off; c.one = nullptr; c.nm = k.n; c.st = b->stream;
  c.models = b->ordered.data() + k.off;
  c.prof = nullptr;
  c.rclass = rclass;
  const Grid& g = c.lead->g;
  c.tune_cols = c.lead->tune_cols; c.tune_green = c.lead->tune_green; c.tune_body = c.lead->tune_body;
  TuneChoice t;
  if (b->tuned && b->tune_grid[0] == g.n1 && b->tune_grid[1] == g.n2 && b->tune_grid[2] == g.n3) {
    c.tune_cols = b->tune[0]; c.tune_green = b->tune[1]; c.tune_body = b->tune[2];
  } else if (tune_entry(b->device, g, rclass, &t)) {
    c.tune_cols = t.cols; c.tune_green = t.green; c.tune_body = t.body;
  }
  return c;
}

// The launch context of a batch of one grid class.
static Ctx batch_ctx(relax_batch_t* b) { return class_ctx(b, b->classes[0], batch_rclass(b)); }

// Workgroups of a mixed batch's work-list launches: 0 gives every work item
// its own workgroup; a positive count makes them walk the list.
#ifndef RELAX_WL_GRID
#define RELAX_WL_GRID 0
#endif

// One RK2 step of every member of a mixed batch: the divergence and strain
// passes as one work-list launch each over all members, the Green passes
// (transforms of their grid's lengths) and the end of the step per class.
static int enqueue_steps_mixed(relax_batch_t* b, int rclass) {
  hipStream_t st = b->stream;
  const int nm = (int)b->models.size();
  const ModelDev* md = b->md_dev;
  const WorkItem* wls = b->wl;
  const WorkItem* wlb = b->wl + (nm + 1);
  std::vector<Ctx> cs;
  for (const auto& k : b->classes) cs.push_back(class_ctx(b, k, rclass));
  int rc;
  // predictor
  if ((rc = launch_bodyforce_list(wlb, md, nm, b->wl_body, RELAX_WL_GRID, rclass, BODY_PRED, st, b->body_variant))) return rc;
  k_finalize_dt<<<(nm + 63) / 64, 64, 0, st>>>(md, nm, 0, 0.0);
  for (const Ctx& c : cs)
    if ((rc = traction_rates(c, 1))) return rc;
  for (const Ctx& c : cs)
    if ((rc = model_green(c, EPI_PRED))) return rc;
  if ((rc = launch_stress_list(STRESS_PRED, wls, md, nm, b->wl_stress, RELAX_WL_GRID, rclass, st))) return rc;
  // corrector
  if ((rc = launch_bodyforce_list(wlb, md, nm, b->wl_body, RELAX_WL_GRID, rclass, BODY_CORR, st, b->body_variant))) return rc;
  for (const Ctx& c : cs)
    if ((rc = monitor_friction(c))) return rc;
  for (const Ctx& c : cs)
    if ((rc = traction_rates(c, 0))) return rc;
  for (const Ctx& c : cs)
    if ((rc = model_green(c, EPI_CORR))) return rc;
  if ((rc = launch_stress_list(STRESS_CORR, wls, md, nm, b->wl_stress, RELAX_WL_GRID, rclass, st))) return rc;
  for (const Ctx& c : cs)
    if ((rc = end_steps(c))) return rc;
  return RELAX_OK;
}

int relax_batch_step(relax_batch_t* b, relax_step_info_t* infos) {
  if (!b) return RELAX_ERR_ARG;
  const int nm = (int)b->models.size();
  if (nm == 0) return RELAX_OK;
  int rc;
  if ((rc = batch_refresh(b))) return rc;
  const bool mixed = b->classes.size() > 1;
  const int rclass = batch_rclass(b);
  const Ctx c = batch_ctx(b);
  auto enqueue = [&]() { return mixed ? enqueue_steps_mixed(b, rclass) : enqueue_steps(c, nullptr); };
  long points = 0;
  for (relax_model_t* m : b->models) points += m->g.H;
  if (!b->exec && (b->stable < GRAPH_AFTER || points >= GRAPH_MAX_POINTS)) {
    ++b->stable;
    if ((rc = enqueue())) return rc;
  } else if (!b->exec) {
    // capture one step of all members; replayed until the membership or a
    // member's record changes
    // (no other thread frees memory or synchronises the device meanwhile)
    CaptureGuard guard;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : relax_hip_error(e, "hipStreamEndCapture");
    }
    b->graph = g;
    const hipError_t ei = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      drop_graph(b);
      return relax_hip_error(ei, "hipGraphInstantiate");
    }
  }
  if (b->exec) HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  if (!infos) return RELAX_OK;
  k_gather_scalars<<<(nm + 63) / 64, 64, 0, b->stream>>>(b->md_dev, nm, b->gather);
  HIP_TRY(hipMemcpyAsync(b->host_sc, b->gather, sizeof(StepScalars) * nm, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  int first = RELAX_OK;
  for (int i = 0; i < nm; ++i) {
    fill_info(b->host_sc[b->rec[i]], infos + i);
    if (!first && b->host_sc[b->rec[i]].status) first = b->host_sc[b->rec[i]].status;
  }
  return first;
}

int relax_batch_busy(relax_batch_t* b) {
  if (!b) return 0;
  return hipStreamQuery(b->stream) == hipErrorNotReady ? 1 : 0;
}

int relax_batch_scalars(relax_batch_t* b, relax_step_info_t* infos) {
  if (!b || !infos) return RELAX_ERR_ARG;
  const int nm = (int)b->models.size();
  if (nm == 0) return RELAX_OK;
  int rc;
  if ((rc = batch_refresh(b))) return rc;
  k_gather_scalars<<<(nm + 63) / 64, 64, 0, b->stream>>>(b->md_dev, nm, b->gather);
  HIP_TRY(hipMemcpyAsync(b->host_sc, b->gather, sizeof(StepScalars) * nm, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (int i = 0; i < nm; ++i) fill_info(b->host_sc[b->rec[i]], infos + i);
  return RELAX_OK;
}

static int wait_batch(relax_model_t* m) {
  if (m->batch) HIP_TRY(hipStreamSynchronize(m->batch->stream));
  return RELAX_OK;
}

// ------------------------------------------------------------ autotuner ----
// Times every variant of the tunable kernels on the model's own work
// buffers (W, NQ, T, TN are scratch between steps; the divergence pass in its
// corrector form only reads the stress and writes W and T) and keeps the
// fastest per device and grid.
int relax_model_set_tuning(relax_model_t* m, const int* choice) {
  if (!m || !choice) return RELAX_ERR_ARG;
  if (choice[0] < 0 || choice[0] >= COLS_VARIANTS || choice[1] < 0 || choice[1] >= GREEN_VARIANTS ||
      choice[2] < 0 || choice[2] >= BODY_VARIANTS)
    return RELAX_ERR_ARG;
  if (int rc = wait_batch(m)) return rc;
  m->tune_cols = choice[0];
  m->tune_green = choice[1];
  m->tune_body = choice[2];
  return RELAX_OK;
}

// Times every variant of the three tunable launches of context c (one launch
// over all of its models each) and returns the fastest in best; times (3 rows
// of RELAX_TUNE_VARIANTS, ms per launch) holds -1 for a variant that does not
// fit the grid.
static int tune_time(const Ctx& c, int* best, float (*times)[RELAX_TUNE_VARIANTS]) {
  static_assert(COLS_VARIANTS <= RELAX_TUNE_VARIANTS && GREEN_VARIANTS <= RELAX_TUNE_VARIANTS &&
                    BODY_VARIANTS <= RELAX_TUNE_VARIANTS,
                "the timing table holds every variant");
  const relax_model_t* m = c.lead;
  const Grid& g = m->g;
  hipStream_t st = c.st;
  const int n2 = g.n2, h3 = g.h3, m1 = g.m1, nm = c.nm;
  const int ys = n2 >= 128 ? n2 : 0, l2 = ys ? n2 / 8 : 0;
  const int ly = ys ? l2 : n2, nyp = ys ? 3 * h3 * 8 : 3 * h3;
  const ModelDev* md = c.md;
  int rc;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return RELAX_ERR_HIP; }
  auto run = [&](int kind, int v) -> int {
    if (kind == 0)
      return launch_fft_ypass(ly, n2, -1, MFIELD(float2, md, W), m1, (long)ly * m1, m1, nyp, ypass_ride(c, -1), nm,
                              st, v);
    if (kind == 1)
      return launch_green_columns(m->gc, h3, MFIELD(float2, md, W), MFIELD(float2, md, NQ),
                                  MFIELD(const float2, md, T), MFIELD(const float2, md, TN), nm, st, l2, v);
    return launch_bodyforce(g, m->S, md, c.one, nm, c.rclass, BODY_CORR, st, v);
  };
  const int nvar[3] = {COLS_VARIANTS, GREEN_VARIANTS, BODY_VARIANTS};
  constexpr int NREP = 3;
  best[0] = best[1] = best[2] = 0;
  for (int k = 0; k < 3; ++k)
    for (int v = 0; v < RELAX_TUNE_VARIANTS; ++v) times[k][v] = -1.f;
  rc = RELAX_OK;
  for (int kind = 0; kind < 3 && !rc; ++kind) {
    float tbest = -1.f;
    for (int v = 0; v < nvar[kind] && !rc; ++v) {
      rc = run(kind, v);  // warm-up (and the shape check)
      if (rc == RELAX_ERR_SIZE) { rc = RELAX_OK; continue; }  // does not fit this grid
      if (rc) break;
      if (hipEventRecord(e0, st) != hipSuccess) { rc = RELAX_ERR_HIP; break; }
      for (int r = 0; r < NREP && !rc; ++r) rc = run(kind, v);
      if (rc) break;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = RELAX_ERR_HIP; break; }
      float t = 0.f;
      (void)hipEventElapsedTime(&t, e0, e1);
      times[kind][v] = t / NREP;
      if (tbest < 0.f || t < tbest) { tbest = t; best[kind] = v; }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  return RELAX_OK;
}

int relax_model_autotune(relax_model_t* m, int* choice, float* ms) {
  if (!m || m->batch) return RELAX_ERR_ARG;
  int rc;
  if ((rc = sync_md(m))) return rc;
  const Grid& g = m->g;
  int best[3];
  float times[3][RELAX_TUNE_VARIANTS];
  if ((rc = tune_time(single(m), best, times))) return rc;
  m->tune_cols = best[0];
  m->tune_green = best[1];
  m->tune_body = best[2];
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[TuneKey{m->device, g.n1, g.n2, g.n3}] = TuneChoice{best[0], best[1], best[2]};
  }
  if (ms)
    for (int k = 0; k < 3; ++k)
      for (int v = 0; v < RELAX_TUNE_VARIANTS; ++v) ms[k * RELAX_TUNE_VARIANTS + v] = times[k][v];
  if (choice) { choice[0] = best[0]; choice[1] = best[1]; choice[2] = best[2]; }
  return RELAX_OK;
}

int relax_batch_autotune(relax_batch_t* b, int* choice, float* ms) {
  if (!b || b->models.empty()) return RELAX_ERR_ARG;
  int rc;
  if ((rc = batch_refresh(b))) return rc;
  if (b->classes.size() > 1) return RELAX_ERR_ARG;  // tune each grid class as a batch of its own
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->tuned = false;
  const Ctx c = batch_ctx(b);
  int best[3];
  float times[3][RELAX_TUNE_VARIANTS];
  if ((rc = tune_time(c, best, times))) return rc;
  drop_graph(b);  // a captured step keeps the shapes it was captured with
  b->tuned = true;
  for (int k = 0; k < 3; ++k) b->tune[k] = best[k];
  b->tune_grid[0] = c.lead->g.n1; b->tune_grid[1] = c.lead->g.n2; b->tune_grid[2] = c.lead->g.n3;
  {
    const Grid& g = c.lead->g;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_btune[BatchTuneKey{b->device, g.n1, g.n2, g.n3, c.rclass}] = TuneChoice{best[0], best[1], best[2]};
  }
  if (ms)
    for (int k = 0; k < 3; ++k)
      for (int v = 0; v < RELAX_TUNE_VARIANTS; ++v) ms[k * RELAX_TUNE_VARIANTS + v] = times[k][v];
  if (choice) { choice[0] = best[0]; choice[1] = best[1]; choice[2] = best[2]; }
  return RELAX_OK;
}
