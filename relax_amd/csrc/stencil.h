// stencil.h -- constants and launchers of the fused stencil kernels.
#pragma once
#include "common.h"
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

int launch_bodyforce(const Grid& g, const RheoParams& R, const StencilConst& S, const float* sig, float* W,
                     float* T, StepScalars* sc, int mode, hipStream_t st);
long stress_blocks(const Grid& g);
int launch_stress(int mode, const Grid& g, const RheoParams& R, const StencilConst& S, const float* X, float* sig,
                  float* tau, float* gam, StepScalars* sc, double2* partials, hipStream_t st);
