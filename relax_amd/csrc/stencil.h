// stencil.h -- constants and launchers of the fused stencil kernels.
#pragma once
#include "common.h"
#include "kernels.h"
#include "modeldev.h"
#include "rheology.h"

enum { STRESS_SET = 0, STRESS_PRED = 1, STRESS_CORR = 2 };
// k_bodyforce: predictor (rheology + relaxation time), corrector, or the
// divergence of a given moment density (no rheology)
enum { BODY_PRED = 0, BODY_CORR = 1, BODY_RAW = 2 };
// Rheology class of a launch: viscous mechanisms only; with fault creep;
// extended (also weak zones and transient creep).  A batch takes the
// largest class of its members.
enum { RHEO_VISCOUS = 0, RHEO_CREEP = 1, RHEO_EXT = 2 };

struct StencilConst {
  float k1h, k2h, k3h;  // fir1 taps divided by the spacing: 0.5/dx
  float idx3;           // 1/dx3 of the one-sided boundary differences
  float lam, mu2;       // lambda and REAL(2 mu) of Hooke's law
};

// Launches over the nm models of the device array md (all of grid g).
// rclass selects the kernels of a rheology class (RHEO_*).
// BODY_RAW takes the divergence of tau (coseismic set-up), the other modes
// of the rheology of sigma.  STRESS_PRED strains the work field W (v),
// STRESS_CORR and STRESS_SET the displacement u.
// `one` (host copy of the record, or NULL) lets a launch over a single model
// pass the record by value instead of through md.
int launch_bodyforce(const Grid& g, const StencilConst& S, const ModelDev* md, const ModelDev* one, int nm,
                     int rclass, int mode, hipStream_t st, int variant = 0);
long stress_blocks(const Grid& g);

// Launches over models of different grids (a mixed batch): wl is a device
// array of nm + 1 entries, entry m model m's grid, stencil constants and
// first work item, entry nm's first the total; `grid` workgroups walk the
// items (0, or the total: one item per workgroup).
struct WorkItem {
  long first;
  Grid g;
  StencilConst S;
};
// whether divergence variant `variant` tiles grid g, and its work items
// (tiles) on g; the strain pass has stress_blocks(g) items
int body_variant_fits(const Grid& g, int variant);
long body_items(const Grid& g, int variant);
int launch_bodyforce_list(const WorkItem* wl, const ModelDev* md, int nm, long total, long grid, int rclass,
                          int mode, hipStream_t st, int variant);
int launch_stress_list(int mode, const WorkItem* wl, const ModelDev* md, int nm, long total, long grid, int rclass,
                       hipStream_t st);
int launch_stress(int mode, const Grid& g, const StencilConst& S, const ModelDev* md, const ModelDev* one, int nm,
                  int rclass, hipStream_t st);
