// stencil.hip -- FIR strain and divergence stencils of the RK2 step, with the
// rheology and the stage updates fused in.
//
//   k_bodyforce: moment = rheology(sig) evaluated on an LDS tile with a
//     one-point halo, marched down x3 through a three-plane ring, then
//     f = -div(moment) (equivalentbodyforce, elastic3d.f90:1214-1506, with
//     the 1-tap centred differentiator fir1, kernel1.inc) and the surface
//     traction t = moment(:,:,1).e3.  The predictor pass also reduces the
//     relaxation time of every mechanism (viscoelastic3d.f90:303,
//     friction3d.f90:223).
//   k_stress: strain from the centred differences of u or v with one-sided
//     differences at the top and bottom planes (stressupdate,
//     elastic3d.f90:956-1180), Hooke's law, and the stage update of the RK2
//     loop (relax.f90:679-701, :833-890) in the same pass.
#include <atomic>

#include "kernels.h"
#include "rheology.h"
#include "stencil.h"

// The kernel bodies take the model's buffers as __restrict__ parameters:
// after inlining, the compiler keeps them as alias scopes, so the per-depth
// rheology tables (read-only) are known not to alias the stores and are
// read through the scalar cache.
//
// RC is the rheology class: RHEO_VISCOUS, RHEO_CREEP (with fault creep) or
// RHEO_EXT (also weak zones and transient creep, which reads the transient
// strain E: epsilonik in the predictor, the mid-step strain in the
// corrector).  TM marks the predictor pass.
template <int TXB, int TYB, bool TM, bool RAW, int RC>
__device__ __forceinline__ void bodyforce_tile(const Grid& g, const RheoParams& R, const StencilConst& S,
                                               const float* __restrict__ sig, const float* __restrict__ E,
                                               float* __restrict__ W, float* __restrict__ T, StepScalars* sc,
                                               int bx, int by, int bz, int zc, int gz) {
  constexpr int NT = TXB * TYB;
  constexpr int HX = TXB + 2, HY = TYB + 2, HP = HX * HY;
  constexpr int NTM = RC == RHEO_EXT ? NMECH : 3;  // mechanisms whose minima are reduced
  __shared__ float ring[3][6][HP];
  __shared__ float rmin[NTM][NT / 64];
  const int tid = threadIdx.x;
  const int tx = tid % TXB, ty = tid / TXB;
  const int n1 = g.n1, n2 = g.n2, h3 = g.h3;
  const long H = g.H;
  const int x0 = bx * TXB, y0 = by * TYB;
  const int z0 = (bz % gz) * zc, z1 = min(z0 + zc, h3);
  float tmin[NTM];
#pragma unroll
  for (int k = 0; k < NTM; ++k) tmin[k] = __builtin_inff();
  int status = 0;
  const bool trans = RC == RHEO_EXT && (R.tr[0].active || R.tr[1].active);
  const int e0 = (ty + 1) * HX + (tx + 1);
  const int i1 = x0 + tx, i2 = y0 + ty;

  auto output = [&](int zo) {
    const float* cur = &ring[zo % 3][0][0];
    const float* up = &ring[(zo + 2) % 3][0][0];
    const float* dn = &ring[(zo + 1) % 3][0][0];
    float v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int ci = (c == 0) ? 2 : (c == 1 ? 4 : 5);  // s13, s23, s33
      const float* a = cur + ci * HP;
      if (zo == 0) v[c] = (dn[ci * HP + e0] - a[e0]) * S.idx3;
      else if (zo == h3 - 1) v[c] = (a[e0] - up[ci * HP + e0]) * S.idx3;
      else v[c] = (dn[ci * HP + e0] - up[ci * HP + e0]) * S.k3h;
    }
    const float* s11 = cur, * s12 = cur + HP, * s13 = cur + 2 * HP, * s22 = cur + 3 * HP, * s23 = cur + 4 * HP;
    const float f1 = (s11[e0 + 1] - s11[e0 - 1]) * S.k1h + (s12[e0 + HX] - s12[e0 - HX]) * S.k2h + v[0];
    const float f2 = (s12[e0 + 1] - s12[e0 - 1]) * S.k1h + (s22[e0 + HX] - s22[e0 - HX]) * S.k2h + v[1];
    const float f3 = (s13[e0 + 1] - s13[e0 - 1]) * S.k1h + (s23[e0 + HX] - s23[e0 - HX]) * S.k2h + v[2];
    const long o = i1 + (long)n1 * (i2 + (long)n2 * zo);
    gst(W + o, -f1);
    gst(W + H + o, -f2);
    gst(W + 2 * H + o, -f3);
    if (zo == 0) {
      const long q = i1 + (long)n1 * i2, P = (long)n1 * n2;
      gst(T + q, s13[e0]);
      gst(T + P + q, s23[e0]);
      gst(T + 2 * P + q, cur[5 * HP + e0]);
    }
  };

  // The tile points (with halo) of one plane that this thread evaluates.
  // Their stresses for plane z+1 are loaded into registers while plane z is
  // evaluated and the ring is synchronised, so the load latency overlaps work.
  constexpr int KP = (HP + NT - 1) / NT;
  int pj1[KP], pj2[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int e = tid + k * NT;
    const int hx = e % HX, hy = e / HX;
    pj1[k] = (x0 + hx - 1 + n1) % n1;
    pj2[k] = (y0 + hy - 1 + n2) % n2;
  }
  T6 pre[KP], pre_e[RC == RHEO_EXT ? KP : 1];
  auto load_plane = [&](int z) {
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      if (tid + k * NT < HP) {
        const long o = pj1[k] + (long)n1 * (pj2[k] + (long)n2 * z);
        pre[k] = load_t6(sig, H, o);
        if (RC == RHEO_EXT) pre_e[k] = trans ? load_t6(E, H, o) : T6{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      }
    }
  };
  const int zs = max(z0 - 1, 0), ze = min(z1, h3 - 1);
  load_plane(zs);
  for (int z = zs; z <= ze; ++z) {
    float* slot = &ring[z % 3][0][0];
    T6 cur[KP], cur_e[RC == RHEO_EXT ? KP : 1];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      cur[k] = pre[k];
      if (RC == RHEO_EXT) cur_e[k] = pre_e[k];
    }
    if (z < ze) load_plane(z + 1);
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int e = tid + k * NT;
      if (e < HP) {
        T6 m;
        if (RAW) {
          m = cur[k];
        } else if (RC == RHEO_EXT) {
          T6 ed;
          const long o = pj1[k] + (long)n1 * (pj2[k] + (long)n2 * z);
          m = rheology_ext<TM, !TM>(R, cur[k], pj1[k], pj2[k], z, o, cur_e[RC == RHEO_EXT ? k : 0], &ed, tmin,
                                    nullptr, nullptr, &status);
          // the corrector's moment includes the inter-seismic loading
          if (!TM && R.inter) {
            const T6 a = load_t6(R.inter, H, o);
            m = T6{m.s11 + a.s11, m.s12 + a.s12, m.s13 + a.s13, m.s22 + a.s22, m.s23 + a.s23, m.s33 + a.s33};
          }
        } else {
          m = rheology<TM, RC == RHEO_CREEP>(R, cur[k], pj1[k], pj2[k], z, tmin, nullptr, nullptr, &status);
        }
        slot[e] = m.s11; slot[HP + e] = m.s12; slot[2 * HP + e] = m.s13;
        slot[3 * HP + e] = m.s22; slot[4 * HP + e] = m.s23; slot[5 * HP + e] = m.s33;
      }
    }
    __syncthreads();
    const int zo = z - 1;
    if (zo >= z0 && zo < z1) output(zo);
    __syncthreads();
  }
  if (z1 == h3 && h3 - 1 >= z0) output(h3 - 1);

  if (status) atomicMax(&sc->status, status);
  if (TM) {
    // wavefront minima, then one atomic per block (positive floats order as
    // unsigned integers, so the result does not depend on the order)
#pragma unroll
    for (int k = 0; k < NTM; ++k) {
      float v = tmin[k];
      for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
      if ((tid & 63) == 0) rmin[k][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < NTM) {
      float v = rmin[tid][0];
      for (int w = 1; w < NT / 64; ++w) v = fminf(v, rmin[tid][w]);
      if (v < __builtin_inff()) atomicMin(&sc->tmin_bits[tid], __float_as_uint(v));
    }
  }
}

// The same divergence with the x3 differences taken in registers: every
// thread owns one point of the tile and keeps the s13, s23, s33 moments of
// its column at the two planes above the current one.  Only the current
// plane's s11, s12, s13, s22, s23 go through LDS, with a one-point halo along
// x1 and x2 and no corners (the divergence uses none), double-buffered so
// that one barrier per plane suffices.  The first 2*(TXB+TYB) threads also
// evaluate one halo point each, on the tile's output planes only; the planes
// just above and below the depth chunk need only the tile's own points.
// Every output is the same expression as in bodyforce_tile.
template <int TXB, int TYB, bool TM, bool RAW, int RC>
__device__ __forceinline__ void bodyforce_tile_zr(const Grid& g, const RheoParams& R, const StencilConst& S,
                                                  const float* __restrict__ sig, const float* __restrict__ E,
                                                  float* __restrict__ W, float* __restrict__ T, StepScalars* sc,
                                                  int bx, int by, int bz, int zc, int gz) {
  constexpr int NT = TXB * TYB;
  constexpr int HX = TXB + 2, HY = TYB + 2, HP = HX * HY;
  constexpr int NH = 2 * (TXB + TYB);  // halo points of one plane
  static_assert(NH <= NT, "one halo point per thread at most");
  constexpr int NTM = RC == RHEO_EXT ? NMECH : 3;  // mechanisms whose minima are reduced
  __shared__ float pl[2][5][HP];
  __shared__ float rmin[NTM][NT / 64];
  const int tid = threadIdx.x;
  const int tx = tid % TXB, ty = tid / TXB;
  const int n1 = g.n1, n2 = g.n2, h3 = g.h3;
  const long H = g.H, P = (long)n1 * n2;
  const int x0 = bx * TXB, y0 = by * TYB;
  const int z0 = (bz % gz) * zc, z1 = min(z0 + zc, h3);
  float tmin[NTM];
#pragma unroll
  for (int k = 0; k < NTM; ++k) tmin[k] = __builtin_inff();
  int status = 0;
  const bool trans = RC == RHEO_EXT && (R.tr[0].active || R.tr[1].active);
  // own point
  const int c1 = x0 + tx, c2 = y0 + ty;
  const int ec = (ty + 1) * HX + (tx + 1);
  const long oc0 = c1 + (long)n1 * c2;
  // halo point: rows above and below the tile, then the columns left and right
  const bool hal = tid < NH;
  int hx = 0, hy = 0;
  if (tid < TXB) { hx = tid; hy = -1; }
  else if (tid < 2 * TXB) { hx = tid - TXB; hy = TYB; }
  else if (tid < 2 * TXB + TYB) { hx = -1; hy = tid - 2 * TXB; }
  else if (hal) { hx = TXB; hy = tid - 2 * TXB - TYB; }
  const int h1 = (x0 + hx + n1) % n1, h2 = (y0 + hy + n2) % n2;
  const int eh = (hy + 1) * HX + (hx + 1);
  const long oh0 = h1 + (long)n1 * h2;

  const T6 zero{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  T6 pc = zero, ph = zero, pce = zero, phe = zero;  // inputs of the next plane
  auto load_plane = [&](int z, bool halo) {
    pc = load_t6(sig, H, oc0 + P * z);
    if (RC == RHEO_EXT && trans) pce = load_t6(E, H, oc0 + P * z);
    if (halo && hal) {
      ph = load_t6(sig, H, oh0 + P * z);
      if (RC == RHEO_EXT && trans) phe = load_t6(E, H, oh0 + P * z);
    }
  };
  auto moment = [&](const T6& s, const T6& e, int p1, int p2, int z, long o) -> T6 {
    if (RAW) return s;
    if (RC == RHEO_EXT) {
      T6 ed;
      T6 m = rheology_ext<TM, !TM>(R, s, p1, p2, z, o, e, &ed, tmin, nullptr, nullptr, &status);
      // the corrector's moment includes the inter-seismic loading
      if (!TM && R.inter) {
        const T6 a = load_t6(R.inter, H, o);
        m = T6{m.s11 + a.s11, m.s12 + a.s12, m.s13 + a.s13, m.s22 + a.s22, m.s23 + a.s23, m.s33 + a.s33};
      }
      return m;
    }
    return rheology<TM, RC == RHEO_CREEP>(R, s, p1, p2, z, tmin, nullptr, nullptr, &status);
  };

  const int zs = max(z0 - 1, 0), ze = min(z1, h3 - 1);
  // moments 13, 23, 33 of the own column at planes z-2 (a) and z-1 (b), and
  // the x1/x2 part of the divergence at plane z-1
  float a13 = 0.f, a23 = 0.f, a33 = 0.f, b13 = 0.f, b23 = 0.f, b33 = 0.f;
  float f1 = 0.f, f2 = 0.f, f3 = 0.f;
  load_plane(zs, zs >= z0);
  for (int z = zs; z <= ze; ++z) {
    const bool outp = z >= z0 && z < z1;
    const T6 sc0 = pc, sce = pce, sh0 = ph, she = phe;
    if (z < ze) load_plane(z + 1, z + 1 >= z0 && z + 1 < z1);
    const T6 mc = moment(sc0, sce, c1, c2, z, oc0 + P * z);
    float* buf = &pl[z & 1][0][0];
    if (outp) {
      buf[ec] = mc.s11; buf[HP + ec] = mc.s12; buf[2 * HP + ec] = mc.s13;
      buf[3 * HP + ec] = mc.s22; buf[4 * HP + ec] = mc.s23;
      if (hal) {
        const T6 mh = moment(sh0, she, h1, h2, z, oh0 + P * z);
        buf[eh] = mh.s11; buf[HP + eh] = mh.s12; buf[2 * HP + eh] = mh.s13;
        buf[3 * HP + eh] = mh.s22; buf[4 * HP + eh] = mh.s23;
      }
    }
    // plane z-1 is complete: its x3 differences take plane z below it
    const int zo = z - 1;
    if (zo >= z0 && zo < z1) {
      float v1, v2, v3;
      if (zo == 0) {
        v1 = (mc.s13 - b13) * S.idx3; v2 = (mc.s23 - b23) * S.idx3; v3 = (mc.s33 - b33) * S.idx3;
      } else {
        v1 = (mc.s13 - a13) * S.k3h; v2 = (mc.s23 - a23) * S.k3h; v3 = (mc.s33 - a33) * S.k3h;
      }
      const long o = oc0 + P * zo;
      gst(W + o, -(f1 + v1));
      gst(W + H + o, -(f2 + v2));
      gst(W + 2 * H + o, -(f3 + v3));
      if (zo == 0) {
        gst(T + oc0, b13);
        gst(T + P + oc0, b23);
        gst(T + 2 * P + oc0, b33);
      }
    }
    __syncthreads();
    if (outp) {
      const float* s11 = buf, * s12 = buf + HP, * s13 = buf + 2 * HP, * s22 = buf + 3 * HP, * s23 = buf + 4 * HP;
      f1 = (s11[ec + 1] - s11[ec - 1]) * S.k1h + (s12[ec + HX] - s12[ec - HX]) * S.k2h;
      f2 = (s12[ec + 1] - s12[ec - 1]) * S.k1h + (s22[ec + HX] - s22[ec - HX]) * S.k2h;
      f3 = (s13[ec + 1] - s13[ec - 1]) * S.k1h + (s23[ec + HX] - s23[ec - HX]) * S.k2h;
    }
    a13 = b13; a23 = b23; a33 = b33;
    b13 = mc.s13; b23 = mc.s23; b33 = mc.s33;
  }
  // the deepest plane of the grid: one-sided difference with the plane above
  if (z1 == h3 && h3 - 1 >= z0) {
    const int zo = h3 - 1;
    const long o = oc0 + P * zo;
    gst(W + o, -(f1 + (b13 - a13) * S.idx3));
    gst(W + H + o, -(f2 + (b23 - a23) * S.idx3));
    gst(W + 2 * H + o, -(f3 + (b33 - a33) * S.idx3));
    if (zo == 0) {
      gst(T + oc0, b13);
      gst(T + P + oc0, b23);
      gst(T + 2 * P + oc0, b33);
    }
  }

  if (status) atomicMax(&sc->status, status);
  if (TM) {
#pragma unroll
    for (int k = 0; k < NTM; ++k) {
      float v = tmin[k];
      for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
      if ((tid & 63) == 0) rmin[k][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < NTM) {
      float v = rmin[tid][0];
      for (int w = 1; w < NT / 64; ++w) v = fminf(v, rmin[tid][w]);
      if (v < __builtin_inff()) atomicMin(&sc->tmin_bits[tid], __float_as_uint(v));
    }
  }
}

// BATCH: the models' records are in the device array md (one model per
// nx*ny*gz tiles); otherwise the one model's record is the argument Dv.
template <int TXB, int TYB, bool TM, bool RAW, int RC, bool BATCH>
__global__ __launch_bounds__(TXB* TYB) void k_bodyforce(Grid g, ModelDev Dv, const ModelDev* __restrict__ md,
                                                        StencilConst S, int zc, int gz) {
  // 1-D grid of tiles (x fastest, then y, depth chunk, model), taken in the
  // XCD-aware order so that neighbouring tiles, which share halo lines, meet
  // in one L2
  const int nx = g.n1 / TXB, ny = g.n2 / TYB;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = tile % nx, by = (tile / nx) % ny, bz = tile / (nx * ny);
  const ModelDev& D = BATCH ? md[bz / gz] : Dv;
  bodyforce_tile<TXB, TYB, TM, RAW, RC>(g, D.R, S, RAW ? D.tau : D.sig, TM ? D.E : D.Em, D.W, D.T, D.sc, bx, by,
                                         bz, zc, gz);
}

template <int TXB, int TYB, bool TM, bool RAW, int RC, bool BATCH>
__global__ __launch_bounds__(TXB* TYB) void k_bodyforce_zr(Grid g, ModelDev Dv, const ModelDev* __restrict__ md,
                                                           StencilConst S, int zc, int gz) {
  const int nx = g.n1 / TXB, ny = g.n2 / TYB;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = tile % nx, by = (tile / nx) % ny, bz = tile / (nx * ny);
  const ModelDev& D = BATCH ? md[bz / gz] : Dv;
  bodyforce_tile_zr<TXB, TYB, TM, RAW, RC>(g, D.R, S, RAW ? D.tau : D.sig, TM ? D.E : D.Em, D.W, D.T, D.sc, bx,
                                            by, bz, zc, gz);
}

// Hooke's law and the stage update of one point from its strain e (the
// differences before the off-diagonal halving).  With transient creep
// (RC == RHEO_EXT) the predictor also writes the mid-step strain
// Em = E + Dt/2 Edot (relax.f90:690-692) and the corrector advances
// E += Dt Edot (:792-793), Edot from the mid-step strain.  pw and ce collect
// |moment| and |tau| of the corrector.
// The pointwise streams (tau, sig, gamma, the transient strains) are read
// and written once per pass with non-temporal accesses, which leaves the
// caches to the stencil's neighbour reads of u or v (0.83 -> 0.75 ms for
// the corrector at 512x512x256).  The divergence pass keeps cached loads of
// the stress: its tiles re-read their halo rows.
template <int MODE, int RC>
__device__ __forceinline__ void stress_point(const RheoParams& R, const StencilConst& S, long H, long o, int i1,
                                             int i2, int i3, T6 e, float* __restrict__ sig, float* __restrict__ tau,
                                             float* __restrict__ gam, float* __restrict__ E,
                                             float* __restrict__ Em, const StepScalars* sc, double& pw, double& ce,
                                             int& status) {
  e.s12 *= 0.5f; e.s13 *= 0.5f; e.s23 *= 0.5f;
  // isotropicstressstrain, elastic3d.f90:129-142
  const float tr = e.s11 + e.s22 + e.s33;
  e.s11 = e.s11 * S.mu2 + S.lam * tr;
  e.s12 *= S.mu2; e.s13 *= S.mu2;
  e.s22 = e.s22 * S.mu2 + S.lam * tr;
  e.s23 *= S.mu2;
  e.s33 = e.s33 * S.mu2 + S.lam * tr;
  const T6 t = load_t6_nt(tau, H, o);
  T6 out;
  if (MODE == STRESS_SET) {
    out = T6{-t.s11 + e.s11, -t.s12 + e.s12, -t.s13 + e.s13, -t.s22 + e.s22, -t.s23 + e.s23, -t.s33 + e.s33};
  } else if (MODE == STRESS_PRED) {
    const T6 s = load_t6_nt(sig, H, o);
    T6 m;
    if (RC == RHEO_EXT) {
      const bool trans = R.tr[0].active || R.tr[1].active;
      const T6 ev = trans ? load_t6_nt(E, H, o) : T6{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      T6 ed;
      m = rheology_ext<false, false>(R, s, i1, i2, i3, o, ev, &ed, nullptr, nullptr, nullptr, &status);
      if (trans) {
        const float h = (float)(sc->Dt / 2);
        store_t6_nt(Em, H, o, T6{h * ed.s11 + ev.s11, h * ed.s12 + ev.s12, h * ed.s13 + ev.s13,
                              h * ed.s22 + ev.s22, h * ed.s23 + ev.s23, h * ed.s33 + ev.s33});
      }
    } else {
      m = rheology<false, RC == RHEO_CREEP>(R, s, i1, i2, i3, nullptr, nullptr, nullptr, &status);
    }
    const float c = -(float)(sc->Dt / 2);
    out = T6{(c * m.s11 - t.s11) + e.s11, (c * m.s12 - t.s12) + e.s12, (c * m.s13 - t.s13) + e.s13,
             (c * m.s22 - t.s22) + e.s22, (c * m.s23 - t.s23) + e.s23, (c * m.s33 - t.s33) + e.s33};
  } else {
    const T6 s = load_t6_nt(sig, H, o);
    float gl = 0.f, gn = 0.f;
    const float dt = (float)sc->Dt;
    T6 m;
    if (RC == RHEO_EXT) {
      const bool trans = R.tr[0].active || R.tr[1].active;
      const T6 em = trans ? load_t6_nt(Em, H, o) : T6{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      T6 ed;
      m = rheology_ext<false, true>(R, s, i1, i2, i3, o, em, &ed, nullptr, &gl, &gn, &status);
      if (trans) {
        const T6 ev = load_t6_nt(E, H, o);
        store_t6_nt(E, H, o, T6{ev.s11 + dt * ed.s11, ev.s12 + dt * ed.s12, ev.s13 + dt * ed.s13,
                             ev.s22 + dt * ed.s22, ev.s23 + dt * ed.s23, ev.s33 + dt * ed.s33});
      }
    } else {
      m = rheology<false, RC == RHEO_CREEP>(R, s, i1, i2, i3, nullptr, &gl, &gn, &status);
    }
    T6 t0 = t;
    if (RC == RHEO_EXT && R.inter) {
      // inter-seismic loading: its moment density joins the corrector's
      // moment and, unscaled, tau (relax.f90:795-799, :1199-1206), before
      // tau += Dt moment (:833)
      const T6 a = load_t6_nt(R.inter, H, o);
      t0 = T6{t.s11 + a.s11, t.s12 + a.s12, t.s13 + a.s13, t.s22 + a.s22, t.s23 + a.s23, t.s33 + a.s33};
      m = T6{m.s11 + a.s11, m.s12 + a.s12, m.s13 + a.s13, m.s22 + a.s22, m.s23 + a.s23, m.s33 + a.s33};
    }
    const T6 tn{t0.s11 + dt * m.s11, t0.s12 + dt * m.s12, t0.s13 + dt * m.s13,
                t0.s22 + dt * m.s22, t0.s23 + dt * m.s23, t0.s33 + dt * m.s33};
    store_t6_nt(tau, H, o, tn);
    if (R.mech[0].active || R.mech[1].active) {
      float gv = gld_nt(gam + o);
      if (R.mech[0].active) gv = gv + dt * gl;
      if (R.mech[1].active) gv = gv + dt * gn;
      gst_nt(gam + o, gv);
    }
    out = T6{-tn.s11 + e.s11, -tn.s12 + e.s12, -tn.s13 + e.s13, -tn.s22 + e.s22, -tn.s23 + e.s23,
             -tn.s33 + e.s33};
    pw += tnorm_d(m);
    ce += tnorm_d(tn);
  }
  store_t6_nt(sig, H, o, out);
}

// The corrector's |moment| and |tau| sums of a workgroup, in a fixed order.
template <int NT>
__device__ __forceinline__ void block_partials(double pw, double ce, double2* __restrict__ partials, int idx) {
  __shared__ double2 red[NT / 64];
  const int tid = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) {
    pw += __shfl_xor(pw, off);
    ce += __shfl_xor(ce, off);
  }
  if ((tid & 63) == 0) red[tid >> 6] = make_double2(pw, ce);
  __syncthreads();
  if (tid == 0) {
    double2 a = red[0];
    for (int w = 1; w < NT / 64; ++w) { a.x += red[w].x; a.y += red[w].y; }
    partials[idx] = a;
  }
}

// One point per thread.  (A z-marching tile form, reading X once per plane
// through LDS, fetched 1.6x less but ran 10 % slower at 512x512x256: its
// lower occupancy cost more than the re-reads, which the last-level cache
// serves.  64x4 tiles marched down 8 or 16 planes with the x3 neighbours in
// registers lost the same way: 104 VGPRs against 68 in the predictor form,
// 0.63 ms against 0.56 ms.)
template <int MODE, int RC>
__device__ __forceinline__ void stress_block(const Grid& g, const RheoParams& R, const StencilConst& S,
                                             const float* __restrict__ X, float* __restrict__ sig,
                                             float* __restrict__ tau, float* __restrict__ gam,
                                             float* __restrict__ E, float* __restrict__ Em, StepScalars* sc,
                                             double2* __restrict__ partials, int blk) {
  const int tid = threadIdx.x;
  // block blk: 256 points in address order (the corrector's partials go to
  // entry blk, so their sum keeps one order whatever the launch order of
  // the blocks; an order in which each XCD walks a contiguous eighth of the
  // grid measured 9-15 % slower at 512x512x256, see slab_tile)
  const long o = (long)blk * 256 + tid;
  const int n1 = g.n1, n2 = g.n2, h3 = g.h3;
  const long H = g.H;
  double pw = 0.0, ce = 0.0;
  int status = 0;
  if (o < H) {
    // n1 and n2 are powers of two (relax_model_create): shifts and masks
    const int l1 = __builtin_ctz((unsigned)n1), l2 = __builtin_ctz((unsigned)n2);
    const int i1 = (int)(o & (n1 - 1)), i2 = (int)((o >> l1) & (n2 - 1)), i3 = (int)(o >> (l1 + l2));
    const long row = o - i1;
    const long xp = row + (i1 + 1 == n1 ? 0 : i1 + 1), xm = row + (i1 == 0 ? n1 - 1 : i1 - 1);
    const long plane = (long)n1 * n2;
    const long pbase = o - (long)i2 * n1;
    const long yp = pbase + (long)(i2 + 1 == n2 ? 0 : i2 + 1) * n1, ym = pbase + (long)(i2 == 0 ? n2 - 1 : i2 - 1) * n1;
    long zp, zm;
    float wz;
    if (i3 == 0) { zp = o + plane; zm = o; wz = S.idx3; }
    else if (i3 == h3 - 1) { zp = o; zm = o - plane; wz = S.idx3; }
    else { zp = o + plane; zm = o - plane; wz = S.k3h; }
    const float* X0 = X;
    const float* X1 = X + H;
    const float* X2 = X + 2 * H;
    T6 e;
    e.s11 = (gld(X0 + xp) - gld(X0 + xm)) * S.k1h;
    e.s12 = (gld(X1 + xp) - gld(X1 + xm)) * S.k1h;
    e.s13 = (gld(X2 + xp) - gld(X2 + xm)) * S.k1h;
    e.s12 = e.s12 + (gld(X0 + yp) - gld(X0 + ym)) * S.k2h;
    e.s22 = (gld(X1 + yp) - gld(X1 + ym)) * S.k2h;
    e.s23 = (gld(X2 + yp) - gld(X2 + ym)) * S.k2h;
    e.s13 = e.s13 + (gld(X0 + zp) - gld(X0 + zm)) * wz;
    e.s23 = e.s23 + (gld(X1 + zp) - gld(X1 + zm)) * wz;
    e.s33 = (gld(X2 + zp) - gld(X2 + zm)) * wz;
    stress_point<MODE, RC>(R, S, H, o, i1, i2, i3, e, sig, tau, gam, E, Em, sc, pw, ce, status);
  }
  if (MODE == STRESS_CORR) block_partials<256>(pw, ce, partials, blk);
  if (status) atomicMax(&sc->status, status);
}

// Block order of the strain pass: workgroup b of nb, with P workgroups per
// plane (P % 8 == 0), takes block slab_tile(b): XCD b % 8 walks the eighth
// [x P/8, (x+1) P/8) of every plane (n2/8 whole rows), plane after plane.
// The x2 and x3 neighbours of a point are then read by workgroups of the
// same XCD a few hundred workgroups apart, and its L2 serves the re-reads:
// at 512x512x256 the pass fetches 1.01x the bytes of its inputs, against
// 1.45x (predictor) and 1.42x (corrector) in address order, where the
// neighbouring rows of a block belong to other XCDs.  The predictor runs
// 0.502 ms against 0.524, the corrector is unchanged (0.738 ms: its
// 45 % share of writes holds it at the HBM rate of its mix).
__device__ __forceinline__ int slab_tile(int b, int nb, int P) {
  if (P % 8 || nb % P) return b;
  const int x = b % 8, l = b / 8, q = P / 8;
  return (l / q) * P + x * q + l % q;
}

template <int MODE, int RC, bool BATCH>
__global__ __launch_bounds__(256) void k_stress(Grid g, ModelDev Dv, const ModelDev* __restrict__ md,
                                               StencilConst S) {
  // model blockIdx.y of the launch (BATCH) or the argument Dv
  const ModelDev& D = BATCH ? md[blockIdx.y] : Dv;
  const int blk = slab_tile(blockIdx.x, gridDim.x, (g.n1 * g.n2) / 256);
  stress_block<MODE, RC>(g, D.R, S, MODE == STRESS_PRED ? D.W : D.u, D.sig, D.tau, D.gam, D.E, D.Em, D.sc,
                         D.partials, blk);
}

// ---------------------------------------------- models of uneven grids ----
// A launch over the nm models of md whose grids differ walks a work list:
// entry m gives model m's grid, stencil constants and its first work item
// (a workgroup's share: 256 points of the strain pass, one tile of the
// divergence pass); entry nm's `first` is the total.  Workgroup b takes
// items b, b + gridDim.x, ... (with one item per workgroup the XCD-aware
// order of the divergence tiles is kept).  The item's model is found by
// binary search over the entries (uniform across the workgroup: scalar
// loads).
__device__ __forceinline__ int work_model(const WorkItem* __restrict__ wl, int nm, long item) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (wl[mid].first <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int MODE, int RC>
__global__ __launch_bounds__(256) void k_stress_list(const ModelDev* __restrict__ md,
                                                     const WorkItem* __restrict__ wl, int nm) {
  const long total = wl[nm].first;
  for (long item = blockIdx.x; item < total; item += gridDim.x) {
    const int m = work_model(wl, nm, item);
    const ModelDev& D = md[m];
    stress_block<MODE, RC>(wl[m].g, D.R, wl[m].S, MODE == STRESS_PRED ? D.W : D.u, D.sig, D.tau, D.gam, D.E, D.Em,
                           D.sc, D.partials, (int)(item - wl[m].first));
    if (MODE == STRESS_CORR) __syncthreads();  // the partials' LDS is reused by the next item
  }
}

template <int TXB, int TYB, bool TM, bool RAW, int RC, bool ZR>
__global__ __launch_bounds__(TXB* TYB) void k_bodyforce_list(const ModelDev* __restrict__ md,
                                                             const WorkItem* __restrict__ wl, int nm, int zc) {
  const long total = wl[nm].first;
  for (long b = blockIdx.x; b < total; b += gridDim.x) {
    const long item = gridDim.x == total ? (long)xcd_tile((int)b, (int)total) : b;
    const int m = work_model(wl, nm, item);
    const ModelDev& D = md[m];
    const Grid& g = wl[m].g;
    const int nx = g.n1 / TXB, ny = g.n2 / TYB, gz = (g.h3 + zc - 1) / zc;
    const int lt = (int)(item - wl[m].first);
    const int bx = lt % nx, by = (lt / nx) % ny, bz = lt / (nx * ny);
    if constexpr (ZR)
      bodyforce_tile_zr<TXB, TYB, TM, RAW, RC>(g, D.R, wl[m].S, RAW ? D.tau : D.sig, TM ? D.E : D.Em, D.W, D.T,
                                               D.sc, bx, by, bz, zc, gz);
    else
      bodyforce_tile<TXB, TYB, TM, RAW, RC>(g, D.R, wl[m].S, RAW ? D.tau : D.sig, TM ? D.E : D.Em, D.W, D.T, D.sc,
                                            bx, by, bz, zc, gz);
    __syncthreads();  // the tile's LDS is reused by the next item
  }
}

// Compute units of the current device (cached per device, safe to call from
// several host threads; 256 if the query fails).
static int device_cus() {
  static std::atomic<int> cus[64];  // zero-initialised (static storage)
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (n <= 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

// Depth chunk of one divergence tile: each chunk re-reads the plane above
// and below it.  The register-window forms (variants 4 and 5: one workgroup
// of 64x16 tiles per CU, two of 64x8, at their register use) take the
// deepest chunk that still gives every CU a workgroup, at least 8 planes:
// at 512x512x256 one chunk of all 128 planes (256 workgroups) runs the
// divergence in 0.250 ms against 0.275 ms for chunks of 32 planes (1024
// workgroups, four rounds) and 0.268 ms for 64.
int body_depth_chunk(const Grid& g, int variant, int nm) {
  if (variant < 4) return variant >= 3 ? 32 : 16;
  const long tiles = (long)(g.n1 / 64) * (g.n2 / (variant == 5 ? 16 : 8)) * (nm > 0 ? nm : 1);
  if (tiles <= 0) return 32;  // the tiles do not fit the grid: the launcher refuses it
  const long want = (long)device_cus() * (variant == 5 ? 1 : 2);
  long gz = (want + tiles - 1) / tiles;
  const long gmax = (g.h3 + 7) / 8;
  if (gz > gmax) gz = gmax;
  if (gz < 1) gz = 1;
  return (int)((g.h3 + gz - 1) / gz);
}

int launch_bodyforce(const Grid& g, const StencilConst& S, const ModelDev* md, const ModelDev* one, int nm,
                     int rclass, int mode, hipStream_t st, int variant) {
  const int zc = body_depth_chunk(g, variant, nm);
  const int gz = (g.h3 + zc - 1) / zc;
  const bool batch = !(one && nm == 1);
  const ModelDev Dv = batch ? ModelDev{} : *one;
  if (rclass < RHEO_VISCOUS || rclass > RHEO_EXT) return RELAX_ERR_ARG;
#define LB(K, TX, TY, F, B)                                                                                 \
  {                                                                                                      \
    const dim3 grid((unsigned)((g.n1 / TX) * (g.n2 / TY) * gz * nm));                                    \
    if (mode == BODY_PRED) K<TX, TY, true, false, F, B><<<grid, TX * TY, 0, st>>>(g, Dv, md, S, zc, gz); \
    else if (mode == BODY_CORR) K<TX, TY, false, false, F, B><<<grid, TX * TY, 0, st>>>(g, Dv, md, S, zc, gz); \
    else K<TX, TY, false, true, RHEO_VISCOUS, B><<<grid, TX * TY, 0, st>>>(g, Dv, md, S, zc, gz); \
  }
#define LF(K, TX, TY, F) { if (batch) LB(K, TX, TY, F, true) else LB(K, TX, TY, F, false) }
#define L(K, TX, TY)                                   \
  {                                                 \
    if (rclass == RHEO_EXT) LF(K, TX, TY, RHEO_EXT)    \
    else if (rclass == RHEO_CREEP) LF(K, TX, TY, RHEO_CREEP) \
    else LF(K, TX, TY, RHEO_VISCOUS)                   \
  }
  // variant 0: 32x16 tiles, 1: 32x8 tiles, 2: 64x8 tiles, 3: 64x8 tiles
  // 32 planes deep (16 for the others); 4 and 5: 64x8 and 64x16 tiles with
  // the x3 differences in registers (k_bodyforce_zr), body_depth_chunk deep
  if (variant < 0 || variant >= BODY_VARIANTS) return RELAX_ERR_ARG;
  if (variant == 4 && g.n1 % 64 == 0 && g.n2 % 8 == 0) L(k_bodyforce_zr, 64, 8)
  else if (variant == 5 && g.n1 % 64 == 0 && g.n2 % 16 == 0) L(k_bodyforce_zr, 64, 16)
  else if (variant >= 4) return RELAX_ERR_SIZE;
  else if (variant == 0 && g.n1 % 32 == 0 && g.n2 % 16 == 0) L(k_bodyforce, 32, 16)
  else if (variant >= 2 && g.n1 % 64 == 0 && g.n2 % 8 == 0) L(k_bodyforce, 64, 8)
  else if (variant != 2 && g.n1 % 32 == 0 && g.n2 % 8 == 0) L(k_bodyforce, 32, 8)
  else return RELAX_ERR_SIZE;
#undef LB
#undef LF
#undef L
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

long stress_blocks(const Grid& g) { return (g.H + 255) / 256; }

int launch_stress(int mode, const Grid& g, const StencilConst& S, const ModelDev* md, const ModelDev* one, int nm,
                  int rclass, hipStream_t st) {
  const dim3 grid((unsigned)stress_blocks(g), (unsigned)nm);
  const bool batch = !(one && nm == 1);
  const ModelDev Dv = batch ? ModelDev{} : *one;
  if (rclass < RHEO_VISCOUS || rclass > RHEO_EXT) return RELAX_ERR_ARG;
#define K(MODE, F)                                                          \
  {                                                                         \
    if (batch) k_stress<MODE, F, true><<<grid, 256, 0, st>>>(g, Dv, md, S);  \
    else k_stress<MODE, F, false><<<grid, 256, 0, st>>>(g, Dv, md, S);       \
  }
#define KM(MODE)                                          \
  {                                                       \
    if (rclass == RHEO_EXT) K(MODE, RHEO_EXT)             \
    else if (rclass == RHEO_CREEP) K(MODE, RHEO_CREEP)    \
    else K(MODE, RHEO_VISCOUS)                            \
  }
  if (mode == STRESS_SET) K(STRESS_SET, RHEO_VISCOUS)
  else if (mode == STRESS_PRED) KM(STRESS_PRED)
  else KM(STRESS_CORR)
#undef K
#undef KM
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

// --------------------------------------------------- work-list launches ----
int body_variant_fits(const Grid& g, int variant) {
  switch (variant) {
    case 0: return g.n1 % 32 == 0 && g.n2 % 16 == 0;
    case 1: return g.n1 % 32 == 0 && g.n2 % 8 == 0;
    case 2: case 3: case 4: return g.n1 % 64 == 0 && g.n2 % 8 == 0;
    case 5: return g.n1 % 64 == 0 && g.n2 % 16 == 0;
    default: return 0;
  }
}

long body_items(const Grid& g, int variant) {
  static const int TX[BODY_VARIANTS] = {32, 32, 64, 64, 64, 64}, TY[BODY_VARIANTS] = {16, 8, 8, 8, 8, 16};
  const int zc = variant >= 3 ? 32 : 16;
  return (long)(g.n1 / TX[variant]) * (g.n2 / TY[variant]) * ((g.h3 + zc - 1) / zc);
}

int launch_bodyforce_list(const WorkItem* wl, const ModelDev* md, int nm, long total, long grid, int rclass,
                          int mode, hipStream_t st, int variant) {
  if (rclass < RHEO_VISCOUS || rclass > RHEO_EXT || variant < 0 || variant >= BODY_VARIANTS || nm < 1)
    return RELAX_ERR_ARG;
  if (total < 1) return RELAX_OK;
  const int zc = variant >= 3 ? 32 : 16;
  const unsigned gx = (unsigned)(grid > 0 && grid < total ? grid : total);
#define LB(TX, TY, ZR, F)                                                                                     \
  {                                                                                                           \
    if (mode == BODY_PRED) k_bodyforce_list<TX, TY, true, false, F, ZR><<<gx, TX * TY, 0, st>>>(md, wl, nm, zc); \
    else if (mode == BODY_CORR) k_bodyforce_list<TX, TY, false, false, F, ZR><<<gx, TX * TY, 0, st>>>(md, wl, nm, zc); \
    else k_bodyforce_list<TX, TY, false, true, RHEO_VISCOUS, ZR><<<gx, TX * TY, 0, st>>>(md, wl, nm, zc); \
  }
#define L(TX, TY, ZR)                                         \
  {                                                           \
    if (rclass == RHEO_EXT) LB(TX, TY, ZR, RHEO_EXT)          \
    else if (rclass == RHEO_CREEP) LB(TX, TY, ZR, RHEO_CREEP) \
    else LB(TX, TY, ZR, RHEO_VISCOUS)                         \
  }
  switch (variant) {
    case 0: L(32, 16, false) break;
    case 1: L(32, 8, false) break;
    case 2: case 3: L(64, 8, false) break;
    case 4: L(64, 8, true) break;
    default: L(64, 16, true) break;
  }
#undef LB
#undef L
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}

int launch_stress_list(int mode, const WorkItem* wl, const ModelDev* md, int nm, long total, long grid, int rclass,
                       hipStream_t st) {
  if (rclass < RHEO_VISCOUS || rclass > RHEO_EXT || nm < 1) return RELAX_ERR_ARG;
  if (total < 1) return RELAX_OK;
  const unsigned gx = (unsigned)(grid > 0 && grid < total ? grid : total);
#define K(MODE, F) k_stress_list<MODE, F><<<gx, 256, 0, st>>>(md, wl, nm);
#define KM(MODE)                                          \
  {                                                       \
    if (rclass == RHEO_EXT) K(MODE, RHEO_EXT)             \
    else if (rclass == RHEO_CREEP) K(MODE, RHEO_CREEP)    \
    else K(MODE, RHEO_VISCOUS)                            \
  }
  if (mode == STRESS_SET) K(STRESS_SET, RHEO_VISCOUS)
  else if (mode == STRESS_PRED) KM(STRESS_PRED)
  else KM(STRESS_CORR)
#undef K
#undef KM
  return hipGetLastError() == hipSuccess ? RELAX_OK : RELAX_ERR_HIP;
}
