// model.h -- the device-resident model handle (internal).
#pragma once
#include <vector>

#include "kernels.h"
#include "rheology.h"
#include "stencil.h"

// Rows of the per-depth table block (each sx3/2 floats): viscous mechanism k
// at 3k (gammadot0, power, cohesion), deviatoric prestress at 10 (six rows),
// its norm at 16, prestress rates at 17+k, (taup/mu)^(n-1) at 19+k,
// transient mechanism k at 21+3k (gammadot0, power, Gk).
static const long TABLE_ROWS = 27;

struct relax_model {
  int device = 0;
  hipStream_t stream = nullptr;
  relax_model_desc_t desc{};
  std::vector<double> events;
  Grid g{};
  // state and work buffers (device)
  float* u = nullptr;      // 3 planes of H: u1, u2, u3 (packed rows of sx1)
  float* sig = nullptr;    // 6 planes of H (SoA tensor)
  float* tau = nullptr;    // 6 planes of H
  float* gam = nullptr;    // H
  float* W = nullptr;      // 3H: body force -> spectrum -> predictor displacement
  float2* NQ = nullptr;    // 3 * sx2 * sx3/2: Nyquist bins of the work spectrum
  float* T = nullptr;      // 3 * sx1 * sx2: surface traction, then its spectrum
  float2* TN = nullptr;    // 3 * sx2: Nyquist bins of the traction spectrum
  StepScalars* sc = nullptr;
  double2* partials = nullptr;
  long nblk = 0;  // stress-kernel workgroups (first-level amplitude partials)
  long nred = 0;  // second-level partials
  float* E = nullptr;       // 6 planes of H: transient strain (with a transient structure)
  float* Em = nullptr;      // 6 planes of H: its mid-step value
  float* dg[4] = {nullptr, nullptr, nullptr, nullptr};  // weak-zone fields (H each)
  float* mi = nullptr;      // 6 planes of H: moment density of the inter-seismic loading
  int inter = 0;            // inter-seismic loading set (RheoParams::inter = mi)
  WeakDev* zones_dev = nullptr;  // staging of set_weakzones' zones
  void* pts_dev = nullptr;        // staging of relax_model_points (grow-only)
  double* ev_dev = nullptr;       // event times on the device
  size_t pts_cap = 0;
  int zones_cap = 0;
  float* tables = nullptr;  // per-depth rheology tables (TABLE_ROWS * sx3/2 floats)
  double* ftab = nullptr;   // fault-creep tables: v0, (a-b)sigma, friction, cohesion

  PlaneDev* planes_dev = nullptr;
  relax_slippatch_t* patches_dev = nullptr;
  long planes_cap = 0, patches_cap = 0;  // grow-only: a reused model allocates nothing
  PlaneDev* shear_dev = nullptr;         // add_shear's patch list
  double* shear_slip = nullptr;
  long shear_cap = 0;
  std::vector<PlaneDev> shear_host;
  relax_source_t* loads_dev = nullptr;  // an event's surface loads, then its point sources
  long loads_cap = 0;
  int cur_nloads = 0;  // the current event's loads (loads_dev[0..)), whose rate each step adds
  void* eig_dev = nullptr;  // an event's eigenstrain volumes (cuboid then tetrahedron records, grow-only)
  size_t eig_cap = 0;
  int cur_ncub = 0, cur_ntet = 0;
  std::vector<relax_plane_t> host_planes;
  std::vector<PlaneDev> plane_dev_host;
  std::vector<long> patch_off;
  StepScalars* host_sc = nullptr;  // pinned
  int host_cap = 0;
  int mech[NMECH] = {0, 0, 0, 0, 0};
  std::vector<double> hg0[2], hn[2], htaup;  // host copies for the prestress-rate table
  RheoParams R{};
  StencilConst S{};
  GreenConst gc{};
  // the device record the step kernels read (a batch of one), its host
  // image, and a version that batches compare to refresh their copies
  ModelDev* md_dev = nullptr;
  ModelDev md_host{};
  bool md_dirty = true;
  unsigned md_version = 0;
  relax_model* self = nullptr;
  relax_batch* batch = nullptr;  // the batch the model is a member of
  // launch-shape variants of the tunable kernels (relax_model_autotune):
  // y transform columns, Green column pass, divergence tiles
  int tune_cols = 0, tune_green = 0, tune_body = 0;
  // phase timers (relax_model_profile)
  bool prof_on = false;
  int prof_n = 0;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_kind;
};

PlaneDev make_plane(const relax_plane_t& p, double beta, double dx1, double dx3);
int launch_frictionadd(relax_slippatch_t* patch, long n, double dt, hipStream_t st);
int launch_monitorfriction(const PlaneDev& P, double rake, int px2, int px3, const Grid& g, double dx1, double dx2,
                           double dx3, int sx3_range, const float* sig, const double* v0, const double* tauc,
                           const double* fric, const double* coh, relax_slippatch_t* patch, StepScalars* sc,
                           hipStream_t st);
int launch_aos_soa(const float* aos, float* soa, long H, bool to_soa, hipStream_t st);
// sig = -tau + C:eps(u) on the model's stream
int model_stress_set(relax_model_t* m);
// eigen.hip: the eigenstrain volumes of an event (cuboids, tetrahedra):
// staged on the device, their body forces added to the top half of f (row
// stride ld, component stride fs), their moment densities added to tau
int eigen_stage(relax_model_t* m, const relax_event_t* ev);
int eigen_body_force(relax_model_t* m, float* f, long ld, long fs);
int eigen_moment(relax_model_t* m);
