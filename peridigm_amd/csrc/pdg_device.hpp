// Device helpers shared by the bond-pass kernels.
#pragma once

#include "pdg_internal.hpp"

namespace pdg {
namespace dev {

__device__ __forceinline__ float fsqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ float frcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ float frsq(float v) { return __builtin_amdgcn_rsqf(v); }

// Influence function (src/core/Peridigm_InfluenceFunction.hpp:69-88) of a
// bond length and horizon given in the same unit.
__device__ __forceinline__ float omega_hat(int kind, float zeta, float hhat) {
  if (kind == PDG_INFLUENCE_ONE || kind == kPairModelPmb) return 1.0f;
  if (kind == PDG_INFLUENCE_PARABOLIC_DECAY) {
    if (zeta > hhat) return 0.0f;
    float s = zeta / hhat;
    return s < 0.5f ? 1.0f : -4.0f * s * s + 4.0f * s;
  }
  float h2 = hhat * hhat * 0.16f;
  return expf(-(zeta * zeta) / h2);
}

// Cross-lane moves inside 16-lane rows (DPP): quad_perm [1,0,3,2] and
// [2,3,0,1], half-row mirror, row mirror.
enum { DPP_QXOR1 = 0xB1, DPP_QXOR2 = 0x4E, DPP_HMIRROR = 0x141, DPP_MIRROR = 0x140 };

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}

// Sum over aligned groups of G lanes (4, 8, 16, 32 or 64), result in every
// lane of the group. Within a 16-lane row it takes up to four DPP adds (quad
// pairs, quads, half rows, rows); wider groups add shuffles across rows. Call
// with all lanes active.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
  v += dpp_f<DPP_QXOR1>(v);
  v += dpp_f<DPP_QXOR2>(v);
  if (G >= 8) v += dpp_f<DPP_HMIRROR>(v);
  if (G >= 16) v += dpp_f<DPP_MIRROR>(v);
  if (G >= 32) v += __shfl_xor(v, 16);
  if (G >= 64) v += __shfl_xor(v, 32);
  return v;
}

template <int G>
__device__ __forceinline__ int group_sum_i(int v) {
  v += dpp_i<DPP_QXOR1>(v);
  v += dpp_i<DPP_QXOR2>(v);
  if (G >= 8) v += dpp_i<DPP_HMIRROR>(v);
  if (G >= 16) v += dpp_i<DPP_MIRROR>(v);
  if (G >= 32) v += __shfl_xor(v, 16);
  if (G >= 64) v += __shfl_xor(v, 32);
  return v;
}

// Wave-wide maximum of a value that is uniform inside each group of G lanes
// (read from the first lane of every group; returns a wave-uniform value).
template <int G>
__device__ __forceinline__ int groups_max(int v) {
  int m = __builtin_amdgcn_readfirstlane(v);
#pragma unroll
  for (int g = G; g < 64; g += G) m = max(m, __builtin_amdgcn_readlane(v, g));
  return m;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

struct BondGeom {
  float x0, x1, x2;  // xi
  float y0, y1, y2;  // Y = xi + eta
  float zeta, ext;   // |xi|, |Y| - |xi|
  float ry;          // 1/|Y| (when requested)
};

// Compensated bond geometry, lengths in units of the position quantum q.
// xi is the exact integer difference of the neighbours' 32-bit fixed-point
// positions (taken modulo 2^32; every bond component is below 2^31 q), so
// xi_ij = -xi_ji exactly; eta = u_j - u_i is formed in FP64 from the records
// and rounded once. Both are rounded to FP32 once, Y = xi + eta, and the
// extension uses e = eta . (xi + Y) / (|Y| + |xi|), which has no
// cancellation.
template <bool NEED_RY>
__device__ __forceinline__ BondGeom bond_geom_q(unsigned xi0, unsigned xi1, unsigned xi2,
                                                double ui0, double ui1, double ui2, unsigned xj0,
                                                unsigned xj1, unsigned xj2, double uj0, double uj1,
                                                double uj2) {
  BondGeom b;
  b.x0 = (float)(int)(xj0 - xi0);
  b.x1 = (float)(int)(xj1 - xi1);
  b.x2 = (float)(int)(xj2 - xi2);
  const float e0 = (float)(uj0 - ui0);
  const float e1 = (float)(uj1 - ui1);
  const float e2 = (float)(uj2 - ui2);
  const float z2 = fmaf(b.x2, b.x2, fmaf(b.x1, b.x1, b.x0 * b.x0));
  b.zeta = fsqrt(z2);
  if (NEED_RY) {
    // |Y|^2 - |xi|^2 = eta . (xi + Y)
    b.y0 = b.x0 + e0;
    b.y1 = b.x1 + e1;
    b.y2 = b.x2 + e2;
    const float dot = fmaf(e2, b.x2 + b.y2, fmaf(e1, b.x1 + b.y1, e0 * (b.x0 + b.y0)));
    const float yy = z2 + dot;
    b.ry = frsq(yy);
    b.ext = dot * frcp(yy * b.ry + b.zeta);
  } else {
    // |Y|^2 = |xi|^2 + 2 xi.eta + eta.eta and eta.(xi + Y) = 2 xi.eta + eta.eta
    // share the two small dot products
    const float xe = fmaf(e2, b.x2, fmaf(e1, b.x1, e0 * b.x0));
    const float ee = fmaf(e2, e2, fmaf(e1, e1, e0 * e0));
    const float dot = fmaf(2.0f, xe, ee);
    const float yy = z2 + dot;
    b.y0 = b.y1 = b.y2 = 0.0f;
    b.ry = 0.0f;
    b.ext = dot * frcp(fsqrt(yy) + b.zeta);
  }
  return b;
}

// The global inputs of one Verlet component, loadable ahead of the force.
struct VerletPre {
  double rho, v;
  int slot;  // boundary-condition slot (-1: none, or no new step starts)
};

__device__ __forceinline__ VerletPre verlet_pre(const VerletArgs& va, int node, int l) {
  VerletPre p;
  p.rho = va.rho[node];
  p.v = va.v[3 * node + l];
  p.slot = va.start_step ? va.bc_slot[node] : -1;
  return p;
}

// Velocity-Verlet epilogue for one component l of node `node` (FP64,
// src/core/Peridigm.cpp:1480-1501, :1569-1577): finishes the current step
// (v += dt/2 a) and, unless this is the last force evaluation, starts the next
// one (v += dt/2 a, prescribed-displacement velocity, u += dt v). u_scaled is
// the node's displacement / q from the current record; the new one goes to
// rec_next. `pre` holds the component's inputs from verlet_pre.
__device__ __forceinline__ void verlet_epilogue(const VerletArgs& va, int node, int l, float fc,
                                                double u_scaled, NodeRec* __restrict__ rec_next,
                                                const VerletPre& pre) {
  const int q = 3 * node + l;
  const double f = (double)fc;
  if (!isfinite(f)) *va.err = 1;
  double acc = f;
  acc += 0.0;  // external force (none)
  acc /= pre.rho;
  double v = pre.v;
  double u = u_scaled * va.uq;
  const double dt2 = va.dt / 2.0;
  if (va.end_kick) v = v + dt2 * acc;
  if (va.start_step) {
    v = v + dt2 * acc;
    const int slot = pre.slot;
    if (slot >= 0 && ((va.bcs[slot].mask >> l) & 1)) {
      const double c0 = va.bcs[slot].c0[l], c1 = va.bcs[slot].c1[l];
      const double cur = c1 * va.t_cur + c0;
      double prev = c1 * va.t_prev + c0;
      if (cur - prev == 0.0) prev = u;
      v = 0.0 * (cur - prev) + 1.0 * (cur - prev) * va.inv_dtbc;
    }
    u = u + va.dt * v;
  } else {
    va.a[q] = acc;
    va.f[q] = f;
  }
  va.v[q] = v;
  rec_next[node].u[l] = u * va.inv_uq;
}

__device__ __forceinline__ void verlet_epilogue(const VerletArgs& va, int node, int l, float fc,
                                                double u_scaled, NodeRec* __restrict__ rec_next) {
  verlet_epilogue(va, node, l, fc, u_scaled, rec_next, verlet_pre(va, node, l));
}

}  // namespace dev
}  // namespace pdg
