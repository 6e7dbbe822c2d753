// Bond-pass kernels of the LPS + critical-stretch path for gfx950.
//
// Two CSR sweeps per explicit step (theta_j must be complete before any
// force on i):
//   pass 1: critical-stretch damage update + dilatation   (reference:
//           src/damage/Peridigm_CriticalStretchDamageModel.cpp:108-184,
//           src/materials/material_utilities.cxx:235-288)
//   pass 2: LPS pair force gathered from both bond ends + fused
//           velocity-Verlet epilogue (src/materials/elastic.cxx:94-145,
//           src/core/Peridigm.cpp:1447-1590)
//
// Mapping: a group of G lanes (G = 16, 32 or 64) owns one node; the lanes
// stride over the node's CSR row, so index loads are contiguous and each
// neighbour record is one 64-byte gather. Partial sums are reduced inside the
// group with cross-lane shuffles: no atomics, deterministic results.
//
// Bond arithmetic is FP32 with compensated deformed-bond vectors: the bond
// vector xi = x_j - x_i is the exact integer difference of the records'
// fixed-point positions and the displacement difference eta = u_j - u_i is
// formed in FP64 from the records; both are rounded once to FP32, and the
// extension uses the cancellation-free identity
//     e = |Y| - |xi| = eta . (xi + Y) / (|Y| + |xi|),   Y = xi + eta,
// so stretches near 1e-4 keep full FP32 relative accuracy. Sums are FP32.
//
// The reference scatters f_j -= t_ij (Y/|Y|) V_i. With a symmetric list and
// symmetric damage the net force is f_i = sum_j (t_ij + t_ji) (Y_ij/|Y_ij|) V_j,
// t_ij = w [theta_i (3K/m_i - alpha_i/3) zeta + alpha_i e] (1-d_ij); each
// directed bond is evaluated once from each end.

#include "pdg_device.hpp"

namespace pdg {

namespace {

using namespace dev;

// ---------------------------------------------------------------------------
// pass 1: damage + dilatation
// ---------------------------------------------------------------------------
template <int G, bool DAMAGE, bool NEAR>
__global__ __launch_bounds__(256) void k_pass1(BondArgs a, NodeRec* __restrict__ rec,
                                               float* __restrict__ theta,
                                               int* __restrict__ nbroken) {
  constexpr int NPW = 64 / G;
  const int lane = threadIdx.x & 63;
  const int g = lane / G;
  const int l = lane & (G - 1);
  const int node = (blockIdx.x * 4 + (threadIdx.x >> 6)) * NPW + g;
  const bool valid = node < a.n;

  int rs = 0, deg = 0, w0 = 0;
  unsigned xi[3] = {0u, 0u, 0u};
  double ui[3] = {0.0, 0.0, 0.0};
  NodeConst c{0.0f, 0.0f, 0.0f, 1.0f};
  if (valid) {
    rs = a.row[node];
    deg = a.row[node + 1] - rs;
    w0 = a.mw[node];
    const NodeRec me = rec[node];
    xi[0] = me.xq[0]; xi[1] = me.xq[1]; xi[2] = me.xq[2];
    ui[0] = me.u[0]; ui[1] = me.u[1]; ui[2] = me.u[2];
    c = a.nc[node];
  }
  const int nit = (deg + G - 1) / G;
  const int nmax = wave_max(nit);

  float acc = 0.0f;
  unsigned long long wnew = 0ull, wold = 0ull;
  int nb = 0, nnear = 0;
  for (int it = 0; it < nmax; ++it) {
    const int k = it * G + l;
    bool brk = false;
    if (k < deg) {
      const int j = a.col[rs + k];
      const NodeRec r = rec[j];
      wold = a.mask[w0 + (k >> 6)];
      brk = ((wold >> (k & 63)) & 1ull) != 0ull;
      const BondS b = bond_stretch(bond_q(xi[0], xi[1], xi[2], ui[0], ui[1], ui[2], r.xq[0],
                                          r.xq[1], r.xq[2], r.u[0], r.u[1], r.u[2]));
      if (DAMAGE) {
        // a time-dependent critical stretch replaces every damage-enabled node's
        const float scz = (a.sc_now > 0.0f && c.sc > 0.0f) ? a.sc_now : c.sc;
        if (NEAR && !brk && fabsf(b.s - scz) <= a.near_rel * scz) nnear++;
        brk = brk || (b.s > scz);
      }
      // w zeta e V = w z2 s V
      const float w = omega_z2(a.influence, b.z2, c.hhat);
      if (!brk) acc = fmaf(w * b.z2 * b.s, r.vol, acc);
    }
    if (DAMAGE) {
      const unsigned long long bal = __ballot(brk);
      const unsigned long long gb =
          G == 64 ? bal : ((bal >> (g * G)) & ((1ull << (G & 63)) - 1ull));
      wnew |= gb << ((it * G) & 63);
      const bool wend = ((((it + 1) * G) & 63) == 0) || (it + 1 == nit);
      if (it < nit && wend) {
        if (l == 0) {
          if (wnew != wold) a.mask[w0 + ((it * G) >> 6)] = wnew;
          nb += __popcll(wnew);
        }
        wnew = 0ull;
      }
    }
  }
  acc = group_sum<G>(acc);
  if (NEAR) nnear = group_sum_i<G>(nnear);
  if (valid && l == 0) {
    theta[node] = c.ctheta * acc;
    rec[node].p = c.kp * acc;
    if (DAMAGE && nbroken) nbroken[node] = nb;
    if (NEAR && nnear) atomicAdd(a.near_ctr, (unsigned long long)nnear);
  }
}

// ---------------------------------------------------------------------------
// pass 2: both-ends pair force (+ fused velocity-Verlet)
// ---------------------------------------------------------------------------
template <int G, bool VERLET>
__global__ __launch_bounds__(256) void k_pass2(BondArgs a, const NodeRec* __restrict__ rec,
                                               NodeRec* __restrict__ rec_next,
                                               double* __restrict__ fout, VerletArgs va) {
  constexpr int NPW = 64 / G;
  const int lane = threadIdx.x & 63;
  const int g = lane / G;
  const int l = lane & (G - 1);
  const int node = (blockIdx.x * 4 + (threadIdx.x >> 6)) * NPW + g;
  const bool valid = node < a.n;

  int rs = 0, deg = 0, w0 = 0;
  unsigned xi[3] = {0u, 0u, 0u};
  double ui[3] = {0.0, 0.0, 0.0};
  float alpha_i = 0.0f, p_i = 0.0f, hhat = 1.0f;
  if (valid) {
    rs = a.row[node];
    deg = a.row[node + 1] - rs;
    w0 = a.mw[node];
    const NodeRec me = rec[node];
    xi[0] = me.xq[0]; xi[1] = me.xq[1]; xi[2] = me.xq[2];
    ui[0] = me.u[0]; ui[1] = me.u[1]; ui[2] = me.u[2];
    alpha_i = me.alpha;
    p_i = me.p;
    hhat = a.nc[node].hhat;
  }
  float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
  for (int k = l; k < deg; k += G) {
    const int j = a.col[rs + k];
    const NodeRec r = rec[j];
    const unsigned long long word = a.mask[w0 + (k >> 6)];
    const bool brk = ((word >> (k & 63)) & 1ull) != 0ull;
    const BondK b = bond_k(bond_q(xi[0], xi[1], xi[2], ui[0], ui[1], ui[2], r.xq[0], r.xq[1],
                                  r.xq[2], r.u[0], r.u[1], r.u[2]));
    float t;
    if (a.influence == kPairModelPmb) {
      t = pair_t<kPairModelPmb>(b, 0.0f, 0.0f, r.vol, hhat, a.pmb_c);
    } else {
      // t = w V_j (P + k (A - P)), src/materials/elastic.cxx:94-145
      const float P = p_i + r.p, D = (alpha_i + r.alpha) - P;
      t = omega_z2(a.influence, b.z2, hhat) * fmaf(b.k, D, P) * r.vol;
    }
    t = brk ? 0.0f : t;
    f0 = fmaf(t, b.y0, f0);
    f1 = fmaf(t, b.y1, f1);
    f2 = fmaf(t, b.y2, f2);
  }
  f0 = group_sum<G>(f0);
  f1 = group_sum<G>(f1);
  f2 = group_sum<G>(f2);
  if (!valid || l >= 3) return;
  // the sums are in units of 1/q: back to force density (exact scaling)
  const float fc = (l == 0 ? f0 : (l == 1 ? f1 : f2)) * a.fq;
  if (!VERLET) {
    fout[3 * node + l] = (double)fc;
    return;
  }
  verlet_epilogue(va, node, l, fc, rec[node].u[l], rec_next);
}

// Start of a step from a stored acceleration (continuing runs).
__global__ __launch_bounds__(256) void k_verlet_prologue(int n, const NodeRec* __restrict__ rec,
                                                         NodeRec* __restrict__ rec_next,
                                                         VerletArgs va) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  const int node = q / 3, l = q - 3 * (q / 3);
  const double dt2 = va.dt / 2.0;
  double v = va.v[q];
  double u = rec[node].u[l] * va.uq;
  v = v + dt2 * va.a[q];
  const int slot = va.bc_slot[node];
  if (slot >= 0 && ((va.bcs[slot].mask >> l) & 1)) {
    const double c0 = va.bcs[slot].c0[l], c1 = va.bcs[slot].c1[l];
    const double cur = c1 * va.t_cur + c0;
    double prev = c1 * va.t_prev + c0;
    if (cur - prev == 0.0) prev = u;
    v = 0.0 * (cur - prev) + 1.0 * (cur - prev) * va.inv_dtbc;
  }
  u = u + va.dt * v;
  va.v[q] = v;
  rec_next[node].u[l] = u * va.inv_uq;
}

// Weighted volume in the reference's own summation order (one thread per
// point, src/materials/material_utilities.cxx:161-188 / :501-520). FP64 with
// no contraction (the file is built with -ffp-contract=off).
__device__ double omega_fp64(int kind, double zeta, double horizon) {
  if (kind == PDG_INFLUENCE_ONE) return 1.0;
  if (kind == PDG_INFLUENCE_PARABOLIC_DECAY) {
    if (zeta > horizon) return 0.0;
    double s = zeta / horizon;
    return s < 0.5 ? 1.0 : -4.0 * s * s + 4.0 * s;
  }
  double h2 = horizon * horizon * 0.4 * 0.4;
  return exp(-(zeta * zeta) / h2);
}

__global__ __launch_bounds__(256) void k_weighted_volume(int n, const double* __restrict__ x,
                                                         const double* __restrict__ vol,
                                                         const int* __restrict__ nlist,
                                                         const long long* __restrict__ head,
                                                         double horizon, int kind,
                                                         double* __restrict__ m) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const long long h = head[p];
  const int nn = nlist[h];
  const double X0 = x[3 * p], X1 = x[3 * p + 1], X2 = x[3 * p + 2];
  double mm = 0.0;
  for (int k = 0; k < nn; k++) {
    const int j = nlist[h + 1 + k];
    const double dx = x[3 * j] - X0;
    const double dy = x[3 * j + 1] - X1;
    const double dz = x[3 * j + 2] - X2;
    const double z2 = dx * dx + dy * dy + dz * dz;
    const double d = sqrt(z2);
    mm += omega_fp64(kind, d, horizon) * (z2) * vol[j];
  }
  m[p] = mm;
}

inline unsigned blocks_for(int n, int lanes) {
  const int nodes_per_block = 4 * (64 / lanes);
  return (unsigned)((n + nodes_per_block - 1) / nodes_per_block);
}

}  // namespace

hipError_t launch_pass1(const BondArgs& a, NodeRec* rec, float* theta, int* nbroken,
                        bool damage, bool count_near, int lanes, hipStream_t s) {
  if (lanes < 16) lanes = 16;  // the untiled kernels use 16, 32 or 64 lanes
  if (a.n <= 0) return hipSuccess;
  const dim3 grid(blocks_for(a.n, lanes)), block(256);
#define PDG_P1(G)                                                                        \
  if (damage && count_near)                                                              \
    hipLaunchKernelGGL((k_pass1<G, true, true>), grid, block, 0, s, a, rec, theta, nbroken); \
  else if (damage)                                                                       \
    hipLaunchKernelGGL((k_pass1<G, true, false>), grid, block, 0, s, a, rec, theta, nbroken); \
  else                                                                                   \
    hipLaunchKernelGGL((k_pass1<G, false, false>), grid, block, 0, s, a, rec, theta, nbroken);
  if (lanes == 16) { PDG_P1(16) }
  else if (lanes == 32) { PDG_P1(32) }
  else { PDG_P1(64) }
#undef PDG_P1
  return hipGetLastError();
}

hipError_t launch_pass2_force(const BondArgs& a, const NodeRec* rec, double* f, int lanes,
                              hipStream_t s) {
  if (lanes < 16) lanes = 16;  // the untiled kernels use 16, 32 or 64 lanes
  if (a.n <= 0) return hipSuccess;
  const dim3 grid(blocks_for(a.n, lanes)), block(256);
  VerletArgs va{};
  if (lanes == 16)
    hipLaunchKernelGGL((k_pass2<16, false>), grid, block, 0, s, a, rec, nullptr, f, va);
  else if (lanes == 32)
    hipLaunchKernelGGL((k_pass2<32, false>), grid, block, 0, s, a, rec, nullptr, f, va);
  else
    hipLaunchKernelGGL((k_pass2<64, false>), grid, block, 0, s, a, rec, nullptr, f, va);
  return hipGetLastError();
}

hipError_t launch_pass2_verlet(const BondArgs& a, const NodeRec* rec, NodeRec* rec_next,
                               const VerletArgs& va, int lanes, hipStream_t s) {
  if (lanes < 16) lanes = 16;  // the untiled kernels use 16, 32 or 64 lanes
  if (a.n <= 0) return hipSuccess;
  const dim3 grid(blocks_for(a.n, lanes)), block(256);
  if (lanes == 16)
    hipLaunchKernelGGL((k_pass2<16, true>), grid, block, 0, s, a, rec, rec_next, nullptr, va);
  else if (lanes == 32)
    hipLaunchKernelGGL((k_pass2<32, true>), grid, block, 0, s, a, rec, rec_next, nullptr, va);
  else
    hipLaunchKernelGGL((k_pass2<64, true>), grid, block, 0, s, a, rec, rec_next, nullptr, va);
  return hipGetLastError();
}

hipError_t launch_verlet_prologue(int n, NodeRec* rec, NodeRec* rec_next, const VerletArgs& va,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int tot = 3 * n;
  hipLaunchKernelGGL(k_verlet_prologue, dim3((tot + 255) / 256), dim3(256), 0, s, n, rec,
                     rec_next, va);
  return hipGetLastError();
}

hipError_t launch_weighted_volume(int n, const double* x, const double* vol, const int* nlist,
                                  const long long* head, double horizon, int influence, double* m,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_weighted_volume, dim3((n + 255) / 256), dim3(256), 0, s, n, x, vol, nlist,
                     head, horizon, influence, m);
  return hipGetLastError();
}

}  // namespace pdg
