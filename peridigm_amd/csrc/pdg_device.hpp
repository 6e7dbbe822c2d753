// Device helpers shared by the bond-pass kernels.
#pragma once

#include "pdg_internal.hpp"

namespace pdg {
namespace dev {

__device__ __forceinline__ float fsqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ float frcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ float frsq(float v) { return __builtin_amdgcn_rsqf(v); }

// Influence function (src/core/Peridigm_InfluenceFunction.hpp:69-88) of a
// bond with squared length z2, horizon hhat in the same unit. The
// Gaussian needs no square root (exp(-zeta^2 / (0.16 delta^2))).
__device__ __forceinline__ float omega_z2(int kind, float z2, float hhat) {
  if (kind == PDG_INFLUENCE_ONE || kind == kPairModelPmb) return 1.0f;
  if (kind == PDG_INFLUENCE_PARABOLIC_DECAY) {
    const float zeta = fsqrt(z2);
    if (zeta > hhat) return 0.0f;
    float s = zeta / hhat;
    return s < 0.5f ? 1.0f : -4.0f * s * s + 4.0f * s;
  }
  float h2 = hhat * hhat * 0.16f;
  return expf(-z2 / h2);
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

// Bond geometry in units of the position quantum q. xi is the integer
// difference of the neighbours' 32-bit fixed-point positions (taken modulo
// 2^32; every bond component is below 2^27 q), rounded once to FP32, so
// xi_ij = -xi_ji exactly. eta = u_j - u_i is formed in FP64 from the records
// and rounded once. d = |Y|^2 - |xi|^2 = eta . (2 xi + eta), with
// Y = xi + eta, has no cancellation; with g = |xi| |Y| = sqrt(z2 yy) the
// stretch and the extension over the current length follow with one square
// root and one reciprocal:
//   s     = (|Y| - |xi|) / |xi| = d / (z2 + g),
//   e/|Y| = (|Y| - |xi|) / |Y|  = d / (yy + g).
// Swapping i and j negates xi and eta, which leaves z2, d, s and e/|Y|
// bit-identical, so both ends of a bond decide its break alike.
struct BondQ {
  float x0, x1, x2;  // xi
  float e0, e1, e2;  // eta
};

__device__ __forceinline__ BondQ bond_q(unsigned xi0, unsigned xi1, unsigned xi2, double ui0,
                                        double ui1, double ui2, unsigned xj0, unsigned xj1,
                                        unsigned xj2, double uj0, double uj1, double uj2) {
  BondQ b;
  b.x0 = (float)(int)(xj0 - xi0);
  b.x1 = (float)(int)(xj1 - xi1);
  b.x2 = (float)(int)(xj2 - xi2);
  b.e0 = (float)(uj0 - ui0);
  b.e1 = (float)(uj1 - ui1);
  b.e2 = (float)(uj2 - ui2);
  return b;
}

__device__ __forceinline__ float bond_z2(const BondQ& b) {
  return fmaf(b.x2, b.x2, fmaf(b.x1, b.x1, b.x0 * b.x0));
}

// Pass-1 quantities: squared length and stretch.
struct BondS {
  float z2, s;
};
__device__ __forceinline__ BondS bond_stretch(const BondQ& b) {
  BondS r;
  const float w0 = fmaf(2.0f, b.x0, b.e0), w1 = fmaf(2.0f, b.x1, b.e1),
              w2 = fmaf(2.0f, b.x2, b.e2);
  const float d = fmaf(b.e2, w2, fmaf(b.e1, w1, b.e0 * w0));
  r.z2 = bond_z2(b);
  const float yy = r.z2 + d;
  const float g = fsqrt(r.z2 * yy);
  r.s = d * frcp(r.z2 + g);
  return r;
}

// Pass-2 quantities: the current bond vector Y, squared length and
// extension over current length k = e / |Y|.
struct BondK {
  float y0, y1, y2;
  float z2, k;
};
__device__ __forceinline__ BondK bond_k(const BondQ& b) {
  BondK r;
  r.y0 = b.x0 + b.e0;
  r.y1 = b.x1 + b.e1;
  r.y2 = b.x2 + b.e2;
  const float d = fmaf(b.e2, b.x2 + r.y2, fmaf(b.e1, b.x1 + r.y1, b.e0 * (b.x0 + r.y0)));
  r.z2 = bond_z2(b);
  const float yy = r.z2 + d;
  const float g = fsqrt(r.z2 * yy);
  r.k = d * frcp(yy + g);
  return r;
}

// Pair force coefficient of a bond (both ends, f_i += t Y): linear
// peridynamic solid, src/materials/elastic.cxx:94-145 with
//   t = w V_j (zeta/|Y| P + e/|Y| A) = w V_j (P + k (A - P)),
// P = p_i + p_j (dilatation terms), A = alpha_i + alpha_j, D = A - P passed as
// D_i + D_j; bond-based, src/materials/elastic_bond_based.cxx:100-125 with
// t = c s V_j / |Y| = c k V_j / zeta.
template <int INFL>
__device__ __forceinline__ float pair_t(const BondK& b, float P, float D, float vj, float hhat,
                                        float pmb_c) {
  if (INFL == kPairModelPmb) return pmb_c * b.k * frsq(b.z2) * vj;
  const float t = fmaf(b.k, D, P) * vj;
  return INFL == PDG_INFLUENCE_ONE ? t : omega_z2(INFL, b.z2, hhat) * t;
}

// The same with the neighbour's terms premultiplied by its volume:
// V_j (P + k D) = (p_i V_j + V_j p_j) + k (d_i V_j + V_j d_j), d = alpha - p.
template <int INFL>
__device__ __forceinline__ float pair_tv(const BondK& b, float p_i, float d_i, float vj,
                                         float vp_j, float vd_j, float hhat, float pmb_c) {
  if (INFL == kPairModelPmb) return pmb_c * b.k * frsq(b.z2) * vj;
  const float t = fmaf(b.k, fmaf(d_i, vj, vd_j), fmaf(p_i, vj, vp_j));
  return INFL == PDG_INFLUENCE_ONE ? t : omega_z2(INFL, b.z2, hhat) * t;
}

// The global inputs of one Verlet component, loadable ahead of the force.
struct VerletPre {
  double inv_rho, v;  // 1 / density, velocity
  double fadd;  // added force density (contact), 0 without
  int slot;     // boundary-condition slot (-1: none, or no new step starts)
};

__device__ __forceinline__ VerletPre verlet_pre(const VerletArgs& va, int node, int l) {
  VerletPre p;
  p.inv_rho = va.inv_rho[(unsigned)node];
  p.v = va.v[3u * (unsigned)node + (unsigned)l];
  const int slot = va.bc_slot[(unsigned)node];  // loaded unconditionally, selected after
  p.slot = va.start_step ? slot : -1;
  p.fadd = va.fadd ? va.fadd[3u * (unsigned)node + (unsigned)l] : 0.0;
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
  double tot = f;
  if (va.fadd) tot += pre.fadd;  // contact force density
  if (!isfinite(tot)) *va.err = 1;
  double acc = tot;
  acc += 0.0;  // external force (none)
  acc *= pre.inv_rho;  // a = f / rho, with 1 / rho formed once at set-up
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
    va.f[q] = tot;
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
