// stencil.h -- constants and launchers of the fused stencil kernels.
#pragma once
#include "common.h"
#include "modeldev.h"
#include "rheology.h"

enum { STRESS_SET = 0, STRESS_PRED = 1, STRESS_CORR = 2 };
// k_bodyforce: predictor (rheology + relaxation time), corrector, or the
// divergence of a given moment density (no rheology)
enum { BODY_PRED = 0, BODY_CORR = 1, BODY_RAW = 2 };

struct StencilConst {
  float k1h, k2h, k3h;  // fir1 taps divided by the spacing: 0.5/dx
  float idx3;           // 1/dx3 of the one-sided boundary differences
  float lam, mu2;       // lambda and REAL(2 mu) of Hooke's law
};

// Launches over the nm models of the device array md (all of grid g).  fric
// selects the kernels with fault creep (some model has afterslip planes).
// BODY_RAW takes the divergence of tau (coseismic set-up), the other modes
// of the rheology of sigma.  STRESS_PRED strains the work field W (v),
// STRESS_CORR and STRESS_SET the displacement u.
// `one` (host copy of the record, or NULL) lets a launch over a single model
// pass the record by value instead of through md.
int launch_bodyforce(const Grid& g, const StencilConst& S, const ModelDev* md, const ModelDev* one, int nm,
                     bool fric, int mode, hipStream_t st);
long stress_blocks(const Grid& g);
int launch_stress(int mode, const Grid& g, const StencilConst& S, const ModelDev* md, const ModelDev* one, int nm,
                  bool fric, hipStream_t st);
