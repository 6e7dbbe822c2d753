// LDS-staged bond passes over Hilbert tiles.
//
// One workgroup (8 waves) owns a tile: a contiguous range of at most 128
// Hilbert-ordered nodes (32 or 64 on small meshes), ordered by degree inside
// the tile. It first stages the fixed-point reference positions, FP64
// displacements and per-node scalars of the tile's halo (the union of the
// tile's nodes and all their neighbours, at most 800 nodes) in LDS with one
// 64-byte record load per halo node, then runs group-of-G-lanes bond loops
// that read neighbours from LDS through 16-bit local indices, three 16-byte
// LDS reads per neighbour. Per bond, HBM and L2 see a 2-byte index and one
// bit of the damage mask instead of a 4-byte index plus a 64-byte record
// gather.
//
// A node's damage-mask words (at most 2, degree <= 128) are held in registers
// for the whole bond loop, the lane's 16-bit neighbour indices are prefetched
// before it, and the loop is unrolled so every register index is static;
// pass 1 merges the per-iteration ballots of new breaks into the words and
// writes back only the words that changed. The influence function is a template parameter.
//
// Workgroup -> tile mapping is XCD-aware: consecutive workgroup ids are
// dispatched to the 8 XCDs in turn, so each XCD gets a contiguous run of
// tiles and the halos shared by neighbouring tiles stay in its L2.
//
// The bond arithmetic and the Verlet epilogue are those of pdg_bonds.hip
// (same helpers), so results match the untiled kernels up to FP32 summation
// order.

#include <hip/hip_cooperative_groups.h>

#include <algorithm>

#include "pdg_device.hpp"

namespace cg = cooperative_groups;

namespace pdg {

namespace {

using namespace dev;

constexpr int HMAX = kTileHaloMax;
constexpr int WG = kTileThreads;
// Minimum waves per SIMD the register allocation must allow (the second
// __launch_bounds__ argument): 4, i.e. two 512-thread workgroups per CU, so at
// most 128 VGPRs. The 39.9 KB of LDS per workgroup (halo capacity 800) would
// admit four, but three already need <= 80 VGPRs, which spills.
constexpr int WAVES_PER_SIMD = 4;
constexpr int MAXDEG = 128;
static_assert(3 * kTileNodesMax <= WG, "one Verlet component per thread");
constexpr int NWORDS = MAXDEG / 64;

__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int per = (ntiles + 7) >> 3;
  return (bid & 7) * per + (bid >> 3);
}

// Tile for the q-th work item of a grid of nb workgroups striding over
// ntiles. In a full round with nb a multiple of 8, workgroup b (dispatched to
// XCD b % 8) takes a contiguous run of tiles; a short last round maps
// identically. Both are bijections onto the round's tiles.
__device__ __forceinline__ int xcd_tile_strided(int q, int nb, int ntiles) {
  const int b = q % nb, base = q - b;
  const int span = min(nb, ntiles - base);
  if (span == nb && (nb & 7) == 0) return base + (b & 7) * (nb >> 3) + (b >> 3);
  return base + b;
}

inline unsigned tile_grid(int ntiles) { return 8u * (unsigned)((ntiles + 7) >> 3); }

template <int INFL>
__device__ __forceinline__ float omega_t(float zeta, float hhat) {
  return omega_hat(INFL, zeta, hhat);
}

// Loads a node's mask words into registers, with the bits of the bonds past
// the row's end set: such "bonds" read as broken, so the loops need no
// separate existence test. The mask array is padded by NWORDS words and the
// index array by MAXDEG entries, so the loads past a row's end stay in bounds
// and need no branch.
__device__ __forceinline__ unsigned long long valid_bits(int deg, int w) {
  const int r = deg - 64 * w;
  return r >= 64 ? ~0ull : (r <= 0 ? 0ull : ((1ull << r) - 1ull));
}

__device__ __forceinline__ void load_words(const unsigned long long* __restrict__ mask, int w0,
                                           int deg, unsigned long long (&m)[NWORDS]) {
  const int nw = (deg + 63) >> 6;
  const unsigned long long* __restrict__ p = mask + w0;
#pragma unroll
  for (int w = 0; w < NWORDS; w++) {
    const unsigned long long v = p[w];
    m[w] = (w < nw ? v : 0ull) | ~valid_bits(deg, w);
  }
}

// Prefetches a lane's local neighbour indices for the whole row: one base
// address, constant offsets. Entries past the row's end belong to later rows
// (or the zero padding), so they are byte offsets below kLdsSlot * HMAX; the
// bonds they stand for are masked out, and the LDS slots they address always
// hold finite values (staged or zero-filled), so those bonds' arithmetic
// stays finite.
template <int G>
__device__ __forceinline__ void load_lcol(const unsigned short* __restrict__ lcol, int rs, int l,
                                          int (&lj)[MAXDEG / G]) {
  const unsigned short* __restrict__ p = lcol + rs + l;
#pragma unroll
  for (int it = 0; it < MAXDEG / G; it++) lj[it] = (int)p[it * G];
}

// LDS staging arrays of one workgroup (shared by both passes), 16-byte slots
// so one byte offset (16 x the halo position) addresses every array and a
// neighbour takes three 16-byte reads:
//   sa: fixed-point position and volume, sb: displacement x and y (units of
//   q), sc: displacement z, alpha and p.
struct __align__(16) SlotC {
  double u2;
  float alpha, p;
};
struct TileLds {
  SlotC sc[HMAX];    // (u2, alpha, p); first, so its reads need no address add
  uint4 sa[HMAX];    // (xq0, xq1, xq2, bits of V/S)
  double2 sb[HMAX];  // (u0, u1)
  float sf[3 * kTileNodesMax];  // pass 2: the tile's forces, for the Verlet epilogue
};
static_assert(sizeof(SlotC) == kLdsSlot && sizeof(uint4) == kLdsSlot && sizeof(double2) == kLdsSlot,
              "16-byte LDS slots");

// One neighbour's staged data, loaded one bond ahead of its use by the
// pipelined bond loops so LDS latency overlaps the previous bond's math.
struct Nb {
  unsigned x0, x1, x2;
  float vol;
  double u0, u1, u2;
  float alpha, p;
};

// jb is the neighbour's byte offset in the slot arrays (local index times
// kLdsSlot, as stored in the tile index array).
template <typename T>
__device__ __forceinline__ T lds_at(const void* arr, int jb) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(arr) + jb);
}

template <bool FORCE>
__device__ __forceinline__ Nb nb_load(const TileLds& L, int jb) {
  Nb n;
  const uint4 a = lds_at<uint4>(L.sa, jb);
  const double2 b = lds_at<double2>(L.sb, jb);
  n.x0 = a.x; n.x1 = a.y; n.x2 = a.z; n.vol = __uint_as_float(a.w);
  n.u0 = b.x; n.u1 = b.y;
  if (FORCE) {
    // one 16-byte read: (u2, the bits of alpha and p)
    const double2 c = lds_at<double2>(L.sc, jb);
    n.u2 = c.x;
    n.alpha = __int_as_float(__double2loint(c.y));
    n.p = __int_as_float(__double2hiint(c.y));
  } else {
    n.u2 = lds_at<double>(L.sc, jb);
    n.alpha = 0.0f; n.p = 0.0f;
  }
  return n;
}

// Stages halo slot h from the node record r (zeros when r is null).
__device__ __forceinline__ void stage_slot(TileLds& L, int h, const NodeRec* r) {
  if (r) {
    L.sa[h] = make_uint4(r->xq[0], r->xq[1], r->xq[2], __float_as_uint(r->vol));
    L.sb[h] = make_double2(r->u[0], r->u[1]);
    L.sc[h] = SlotC{r->u[2], r->alpha, r->p};
  } else {
    L.sa[h] = make_uint4(0u, 0u, 0u, 0u);
    L.sb[h] = make_double2(0.0, 0.0);
    L.sc[h] = SlotC{0.0, 0.0f, 0.0f};
  }
}

// Component l of the displacement staged in slot h.
__device__ __forceinline__ double slot_u(const TileLds& L, int h, int l) {
  return l == 0 ? L.sb[h].x : (l == 1 ? L.sb[h].y : L.sc[h].u2);
}

// Break bit of bond k = it * G + l of a node (it a compile-time constant):
// for G <= 32 the bit lies in one 32-bit half of one mask word.
template <int G>
__device__ __forceinline__ bool mask_bit(const unsigned long long (&m)[NWORDS], int it, int l) {
  const int base = it * G;
  if (G == 64) return ((m[base >> 6] >> l) & 1ull) != 0ull;
  const unsigned long long w = m[base >> 6];
  const unsigned half = (base & 32) ? (unsigned)(w >> 32) : (unsigned)w;
  return ((half >> ((base & 31) + l)) & 1u) != 0u;
}

// A node group's row data: CSR range, mask words (bits past the row set),
// local indices, constants.
template <int G>
struct RowData {
  int rs, deg, w0;
  unsigned long long m[NWORDS];
  int lj[MAXDEG / G];
  NodeConst c;
};

template <int G>
__device__ __forceinline__ void row_load(const BondArgs& a, const TileArgs& t, int node,
                                         bool valid, int l, RowData<G>& r) {
  r.rs = 0;
  r.deg = 0;
  r.w0 = 0;
  r.c = NodeConst{0.0f, 0.0f, 0.0f, 1.0f};
  if (valid) {
    r.rs = a.row[node];
    r.deg = a.row[node + 1] - r.rs;
    r.w0 = a.mw[node];
    r.c = a.nc[node];
  }
  load_words(a.mask, r.w0, r.deg, r.m);
  load_lcol<G>(t.lcol, r.rs, l, r.lj);
}

// Lane masks of the wave (compare results kept in scalar registers):
// predicate codes of the compare builtins.
enum { CMP_FOGT = 2, CMP_FOLE = 5, CMP_IEQ = 32 };

// Lanes whose bond k = it * G + l exists and is not marked broken in m (the
// words carry set bits past the row's end).
template <int G>
__device__ __forceinline__ unsigned long long intact_lanes(const unsigned long long (&m)[NWORDS],
                                                           int it, int l) {
  const int base = it * G;
  unsigned v;
  if (G == 64) {
    v = (unsigned)((m[base >> 6] >> l) & 1ull);
  } else {
    const unsigned long long w = m[base >> 6];
    const unsigned half = (base & 32) ? (unsigned)(w >> 32) : (unsigned)w;
    v = half & (1u << ((base & 31) + l));
  }
  return __builtin_amdgcn_uicmp(v, 0u, CMP_IEQ);
}

// Pass 1 (damage + dilatation) of one tile; every thread of the workgroup
// calls it with the same tile.
template <int G, bool DAMAGE, bool NEAR, int INFL, bool PIPE>
__device__ __forceinline__ void tile_pass1(const BondArgs& a, const TileArgs& t,
                                           NodeRec* __restrict__ rec, float* __restrict__ theta,
                                           int* __restrict__ nbroken, int tile, TileLds& L) {
  const int n0 = t.tile_start[tile], n1 = t.tile_start[tile + 1];
  const int h0 = t.halo_off[tile], hn = min(t.halo_off[tile + 1] - h0, HMAX);
  constexpr int NG = WG / G;
  const int grp = threadIdx.x / G;
  const int g = (threadIdx.x & 63) / G;
  const int l = threadIdx.x & (G - 1);
  // the first round's row data is loaded before the halo staging, so the two
  // sets of global loads are in flight together
  RowData<G> rd;
  row_load<G>(a, t, n0 + grp, n0 + grp < n1, l, rd);
  // slots past the halo are zero-filled: masked-out bonds may address them
  for (int h = threadIdx.x; h < HMAX; h += WG)
    stage_slot(L, h, h < hn ? &rec[(unsigned)t.halo[(unsigned)(h0 + h)]] : nullptr);
  __syncthreads();
  const int own = t.own_off[tile];
  for (int base = n0; base < n1; base += NG) {
    const int node = base + grp;
    const bool valid = node < n1;
    if (base != n0) row_load<G>(a, t, node, valid, l, rd);
    const int deg = rd.deg, w0 = rd.w0;
    const unsigned long long (&mo)[NWORDS] = rd.m;
    const int (&lj)[MAXDEG / G] = rd.lj;
    const NodeConst c = rd.c;
    unsigned xi0 = 0u, xi1 = 0u, xi2 = 0u;
    double ui0 = 0.0, ui1 = 0.0, ui2 = 0.0;
    if (valid) {
      const int li = own + (node - n0);
      const uint4 sa = L.sa[li];
      xi0 = sa.x; xi1 = sa.y; xi2 = sa.z;
      ui0 = L.sb[li].x; ui1 = L.sb[li].y; ui2 = L.sc[li].u2;
    }
    unsigned long long mn[NWORDS];
#pragma unroll
    for (int w = 0; w < NWORDS; w++) mn[w] = 0ull;
    // a time-dependent critical stretch replaces every damage-enabled node's
    const float scz = (a.sc_now > 0.0f && c.sc > 0.0f) ? a.sc_now : c.sc;
    const int nmax = groups_max<G>((deg + G - 1) / G);
    float acc = 0.0f;
    int nnear = 0;               // plain loops: per lane
    unsigned nnear_w = 0;        // pipelined loops: per wave (scalar)
    if (PIPE) {
      // branch-free bond loop: every lane computes, absent and broken bonds
      // are selected away; the next bond's LDS data is in flight. The
      // per-iteration ballots of new breaks stay wave-uniform (scalar
      // registers) and are merged into the group's mask words after the loop.
      unsigned long long balv[MAXDEG / G];
      Nb nb[2];
      nb[0] = nb_load<false>(L, lj[0]);
#pragma unroll
      for (int it = 0; it < MAXDEG / G; ++it) {
        if (it < nmax) {
          if (it + 1 < MAXDEG / G && it + 1 < nmax)
            nb[(it + 1) & 1] = nb_load<false>(L, lj[(it + 1) % (MAXDEG / G)]);
          const Nb& cur = nb[it & 1];
          // lanes whose bond exists and is unbroken; the masks stay in scalar
          // registers, so combining them costs no vector instructions
          const unsigned long long okm = intact_lanes<G>(mo, it, l);
          const BondGeom b = bond_geom_q<false>(xi0, xi1, xi2, ui0, ui1, ui2, cur.x0, cur.x1,
                                                cur.x2, cur.u0, cur.u1, cur.u2);
          unsigned long long brkm = 0ull;  // lanes whose bond breaks now
          if (DAMAGE) {
            const float lim = scz * b.zeta;
            if (NEAR) {
              // intact bonds with |s - s_c| <= near_rel * s_c, counted as a
              // wave-wide lane mask (scalar population count)
              const unsigned long long nm =
                  okm & __builtin_amdgcn_fcmpf(fabsf(b.ext - lim), a.near_rel * lim, CMP_FOLE);
              nnear_w += (unsigned)__builtin_popcountll(nm);
            }
            brkm = okm & __builtin_amdgcn_fcmpf(b.ext, lim, CMP_FOGT);
          }
          const float w = omega_t<INFL>(b.zeta, c.hhat);
          const float nacc = fmaf(w * b.zeta * b.ext, cur.vol, acc);
          acc = __builtin_amdgcn_inverse_ballot_w64(okm & ~brkm) ? nacc : acc;
          if (DAMAGE) balv[it] = brkm;
        } else if (DAMAGE) {
          balv[it] = 0ull;
        }
      }
      if (DAMAGE) {
#pragma unroll
        for (int it = 0; it < MAXDEG / G; ++it) {
          const unsigned long long gb =
              G == 64 ? balv[it] : ((balv[it] >> (g * G)) & ((1ull << (G & 63)) - 1ull));
          mn[(it * G) >> 6] |= gb << ((it * G) & 63);
        }
      }
    } else {
#pragma unroll
    for (int it = 0; it < MAXDEG / G; ++it) {
      if (it < nmax) {
        const int k = it * G + l;
        const bool act = k < deg;
        bool brk = act && mask_bit<G>(mo, it, l);
        if (act) {
          const Nb r = nb_load<false>(L, lj[it]);
          const BondGeom b = bond_geom_q<false>(xi0, xi1, xi2, ui0, ui1, ui2, r.x0, r.x1, r.x2,
                                                r.u0, r.u1, r.u2);
          if (DAMAGE) {
            // s = e / zeta > sc  <=>  e > sc * zeta  (zeta > 0)
            const float lim = scz * b.zeta;
            if (NEAR && !brk && fabsf(b.ext - lim) <= a.near_rel * lim) nnear++;
            brk = brk || (b.ext > lim);
          }
          const float w = omega_t<INFL>(b.zeta, c.hhat);
          if (!brk) acc = fmaf(w * b.zeta * b.ext, r.vol, acc);
        }
        if (DAMAGE) {
          const unsigned long long bal = __ballot(brk);
          const unsigned long long gb =
              G == 64 ? bal : ((bal >> (g * G)) & ((1ull << (G & 63)) - 1ull));
          mn[(it * G) >> 6] |= gb << ((it * G) & 63);
        }
      }
    }
    }
    acc = group_sum<G>(acc);
    if (NEAR) nnear = group_sum_i<G>(nnear);
    if (valid && l == 0) {
      theta[node] = c.ctheta * acc;
      rec[node].p = c.kp * acc;
      if (DAMAGE) {
        const int nw = (deg + 63) >> 6;
        int nb = 0;
#pragma unroll
        for (int w = 0; w < NWORDS; w++) {
          const unsigned long long vb = valid_bits(deg, w);
          const unsigned long long old = mo[w] & vb;
          const unsigned long long nw_ = (mo[w] | mn[w]) & vb;
          if (w < nw && nw_ != old) a.mask[(unsigned)(w0 + w)] = nw_;
          nb += __popcll(nw_);
        }
        if (nbroken) nbroken[node] = nb;
      }
      if (NEAR && nnear) atomicAdd(a.near_ctr, (unsigned long long)nnear);
    }
    if (NEAR && PIPE && nnear_w && (threadIdx.x & 63) == 0)
      atomicAdd(a.near_ctr, (unsigned long long)nnear_w);
  }
  __syncthreads();  // LDS is reused by the next tile
}

template <int G, bool DAMAGE, bool NEAR, int INFL, bool PIPE>
__global__ __launch_bounds__(WG, WAVES_PER_SIMD) void k_pass1_tiled(BondArgs a, TileArgs t,
                                                    NodeRec* __restrict__ rec,
                                                    float* __restrict__ theta,
                                                    int* __restrict__ nbroken) {
  __shared__ TileLds L;
  const int tile = xcd_tile(blockIdx.x, t.ntiles);
  if (tile >= t.ntiles) return;
  tile_pass1<G, DAMAGE, NEAR, INFL, PIPE>(a, t, rec, theta, nbroken, tile, L);
}

// Pass 2 (both-ends force + Verlet epilogue) of one tile.
template <int G, bool VERLET, int INFL, bool PIPE>
__device__ __forceinline__ void tile_pass2(const BondArgs& a, const TileArgs& t,
                                           const NodeRec* __restrict__ rec,
                                           NodeRec* __restrict__ rec_next,
                                           double* __restrict__ fout, const VerletArgs& va,
                                           int tile, TileLds& L) {
  const int n0 = t.tile_start[tile], n1 = t.tile_start[tile + 1];
  const int h0 = t.halo_off[tile], hn = min(t.halo_off[tile + 1] - h0, HMAX);
  constexpr int NG = WG / G;
  const int grp = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  // the first round's row data and this thread's Verlet inputs (component
  // q = threadIdx.x; a tile has at most 3 * 128 <= WG components) are loaded
  // before the halo staging, so all these global loads are in flight together
  RowData<G> rd;
  row_load<G>(a, t, n0 + grp, n0 + grp < n1, l, rd);
  const int q = threadIdx.x;
  const bool has_q = q < 3 * (n1 - n0);
  VerletPre vp{0.0, 0.0, -1};
  if (VERLET && has_q) vp = verlet_pre(va, n0 + q / 3, q - 3 * (q / 3));
  // slots past the halo are zero-filled: masked-out bonds may address them
  for (int h = threadIdx.x; h < HMAX; h += WG)
    stage_slot(L, h, h < hn ? &rec[(unsigned)t.halo[(unsigned)(h0 + h)]] : nullptr);
  __syncthreads();
  const int own = t.own_off[tile];
  for (int base = n0; base < n1; base += NG) {
    const int node = base + grp;
    const bool valid = node < n1;
    if (base != n0) row_load<G>(a, t, node, valid, l, rd);
    const int deg = rd.deg;
    const unsigned long long (&mo)[NWORDS] = rd.m;
    const int (&lj)[MAXDEG / G] = rd.lj;
    const float hhat = rd.c.hhat;
    unsigned xi0 = 0u, xi1 = 0u, xi2 = 0u;
    double ui0 = 0.0, ui1 = 0.0, ui2 = 0.0;
    float alpha_i = 0.0f, p_i = 0.0f;
    if (valid) {
      const int li = own + (node - n0);
      const uint4 sa = L.sa[li];
      xi0 = sa.x; xi1 = sa.y; xi2 = sa.z;
      ui0 = L.sb[li].x; ui1 = L.sb[li].y;
      const SlotC sc = L.sc[li];
      ui2 = sc.u2; alpha_i = sc.alpha; p_i = sc.p;
    }
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    if (PIPE) {
      const int nmax = groups_max<G>((deg + G - 1) / G);
      Nb nb[2];
      nb[0] = nb_load<true>(L, lj[0]);
#pragma unroll
      for (int it = 0; it < MAXDEG / G; ++it) {
        if (it < nmax) {
          if (it + 1 < MAXDEG / G && it + 1 < nmax)
            nb[(it + 1) & 1] = nb_load<true>(L, lj[(it + 1) % (MAXDEG / G)]);
          const Nb& cur = nb[it & 1];
          const unsigned long long okm = intact_lanes<G>(mo, it, l);
          const BondGeom b = bond_geom_q<true>(xi0, xi1, xi2, ui0, ui1, ui2, cur.x0, cur.x1,
                                               cur.x2, cur.u0, cur.u1, cur.u2);
          float tt;
          if (INFL == kPairModelPmb) {
            // bond-based: (t_ij + t_ji) = c s with s = e / zeta
            tt = a.pmb_c * b.ext * frcp(b.zeta) * cur.vol * b.ry;
          } else {
            const float w = omega_t<INFL>(b.zeta, hhat);
            tt = fmaf(b.zeta, p_i + cur.p, b.ext * (alpha_i + cur.alpha));
            tt = w * tt * cur.vol * b.ry;
          }
          tt = __builtin_amdgcn_inverse_ballot_w64(okm) ? tt : 0.0f;
          f0 = fmaf(tt, b.y0, f0);
          f1 = fmaf(tt, b.y1, f1);
          f2 = fmaf(tt, b.y2, f2);
        }
      }
    } else {
#pragma unroll
    for (int it = 0; it < MAXDEG / G; ++it) {
      const int k = it * G + l;
      if (k < deg) {
        const bool brk = mask_bit<G>(mo, it, l);
        const Nb r = nb_load<true>(L, lj[it]);
        const BondGeom b = bond_geom_q<true>(xi0, xi1, xi2, ui0, ui1, ui2, r.x0, r.x1, r.x2,
                                             r.u0, r.u1, r.u2);
        float tt;
        if (INFL == kPairModelPmb) {
          tt = brk ? 0.0f : a.pmb_c * b.ext * frcp(b.zeta) * r.vol * b.ry;
        } else {
          const float w = omega_t<INFL>(b.zeta, hhat);
          tt = fmaf(b.zeta, p_i + r.p, b.ext * (alpha_i + r.alpha));
          tt = brk ? 0.0f : w * tt * r.vol * b.ry;
        }
        f0 = fmaf(tt, b.y0, f0);
        f1 = fmaf(tt, b.y1, f1);
        f2 = fmaf(tt, b.y2, f2);
      }
    }
    }
    f0 = group_sum<G>(f0);
    f1 = group_sum<G>(f1);
    f2 = group_sum<G>(f2);
    if (valid && l < 3) {
      // the sums are in units of 1/q: back to force density (exact scaling)
      const float fc = (l == 0 ? f0 : (l == 1 ? f1 : f2)) * a.fq;
      if (!VERLET) fout[3 * node + l] = (double)fc;
      else L.sf[3 * (node - n0) + l] = fc;
    }
  }
  if (VERLET) {
    // Verlet epilogue of the whole tile at once: one component per thread
    // instead of 3 of every G lanes per node round
    __syncthreads();
    if (has_q) {
      const int k = q / 3, l3 = q - 3 * k;
      verlet_epilogue(va, n0 + k, l3, L.sf[q], slot_u(L, own + k, l3), rec_next, vp);
    }
  }
  __syncthreads();  // LDS is reused by the next tile
}

template <int G, bool VERLET, int INFL, bool PIPE>
__global__ __launch_bounds__(WG, WAVES_PER_SIMD) void k_pass2_tiled(BondArgs a, TileArgs t,
                                                    const NodeRec* __restrict__ rec,
                                                    NodeRec* __restrict__ rec_next,
                                                    double* __restrict__ fout, VerletArgs va) {
  __shared__ TileLds L;
  const int tile = xcd_tile(blockIdx.x, t.ntiles);
  if (tile >= t.ntiles) return;
  tile_pass2<G, VERLET, INFL, PIPE>(a, t, rec, rec_next, fout, va, tile, L);
}

// Persistent explicit run: one cooperative launch advances `nsteps` Verlet
// steps; each step is pass 1 over all tiles, a grid barrier, pass 2 over all
// tiles, a grid barrier. Workgroups stride over tiles in XCD-contiguous
// order. Co-residency of the whole grid is guaranteed by the cooperative
// launch, so the barriers always complete.
template <int G, bool DAMAGE, int INFL>
__global__ __launch_bounds__(WG, WAVES_PER_SIMD) void k_persistent(BondArgs a, TileArgs t, NodeRec* rec0,
                                                   NodeRec* rec1, float* theta, int* nbroken,
                                                   VerletArgs va, int nsteps, double t0,
                                                   int initial_eval) {
  __shared__ TileLds L;
  cg::grid_group grid = cg::this_grid();
  NodeRec* rec[2] = {rec0, rec1};
  int cur = 0;
  const int nb = gridDim.x;
  const int first = initial_eval ? 0 : 1;
  for (int k = first; k <= nsteps; k++) {
    for (int q = blockIdx.x; q < t.ntiles; q += nb) {
      const int tile = xcd_tile_strided(q, nb, t.ntiles);
      tile_pass1<G, DAMAGE, DAMAGE, INFL, true>(a, t, rec[cur], theta, nbroken, tile, L);
    }
    grid.sync();
    VerletArgs v = va;
    v.end_kick = k > 0 ? 1 : 0;
    v.start_step = k < nsteps ? 1 : 0;
    v.t_cur = t0 + ((k + 1) * va.dt);
    v.t_prev = t0 + (k * va.dt);
    v.inv_dtbc = 1.0 / (v.t_cur - v.t_prev);
    for (int q = blockIdx.x; q < t.ntiles; q += nb) {
      const int tile = xcd_tile_strided(q, nb, t.ntiles);
      tile_pass2<G, true, INFL, true>(a, t, rec[cur], rec[cur ^ 1], nullptr, v, tile, L);
    }
    grid.sync();
    cur ^= 1;
  }
}

// pipe: the pipelined branch-free bond loops (default); the plain loops are
// kept for G = 16 as a measured alternative.
template <int G, int INFL, bool PIPE>
void p1_launch(const BondArgs& a, const TileArgs& t, NodeRec* rec, float* theta, int* nbroken,
               bool damage, bool count_near, hipStream_t s) {
  const dim3 grid(tile_grid(t.ntiles)), block(WG);
  if (damage && count_near)
    hipLaunchKernelGGL((k_pass1_tiled<G, true, true, INFL, PIPE>), grid, block, 0, s, a, t, rec, theta, nbroken);
  else if (damage)
    hipLaunchKernelGGL((k_pass1_tiled<G, true, false, INFL, PIPE>), grid, block, 0, s, a, t, rec, theta, nbroken);
  else
    hipLaunchKernelGGL((k_pass1_tiled<G, false, false, INFL, PIPE>), grid, block, 0, s, a, t, rec, theta, nbroken);
}

template <int G, bool VERLET, int INFL, bool PIPE>
void p2_launch(const BondArgs& a, const TileArgs& t, const NodeRec* rec, NodeRec* rec_next,
               double* f, const VerletArgs& va, hipStream_t s) {
  const dim3 grid(tile_grid(t.ntiles)), block(WG);
  hipLaunchKernelGGL((k_pass2_tiled<G, VERLET, INFL, PIPE>), grid, block, 0, s, a, t, rec, rec_next, f, va);
}

template <int G, int INFL>
void p1_pipe(const BondArgs& a, const TileArgs& t, NodeRec* rec, float* theta, int* nbroken,
             bool damage, bool count_near, bool pipe, hipStream_t s) {
  if constexpr (G == 16) {
    if (!pipe) return p1_launch<G, INFL, false>(a, t, rec, theta, nbroken, damage, count_near, s);
  }
  p1_launch<G, INFL, true>(a, t, rec, theta, nbroken, damage, count_near, s);
}

template <int G, bool VERLET, int INFL>
void p2_pipe(const BondArgs& a, const TileArgs& t, const NodeRec* rec, NodeRec* rec_next, double* f,
             const VerletArgs& va, bool pipe, hipStream_t s) {
  if constexpr (G == 16) {
    if (!pipe) return p2_launch<G, VERLET, INFL, false>(a, t, rec, rec_next, f, va, s);
  }
  p2_launch<G, VERLET, INFL, true>(a, t, rec, rec_next, f, va, s);
}

template <int G>
void p1_infl(const BondArgs& a, const TileArgs& t, NodeRec* rec, float* theta, int* nbroken,
             bool damage, bool count_near, bool pipe, hipStream_t s) {
  if (a.influence == PDG_INFLUENCE_ONE || a.influence == kPairModelPmb)
    p1_pipe<G, 0>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
  else if (a.influence == PDG_INFLUENCE_PARABOLIC_DECAY) p1_pipe<G, 1>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
  else p1_pipe<G, 2>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
}

template <int G, bool VERLET>
void p2_infl(const BondArgs& a, const TileArgs& t, const NodeRec* rec, NodeRec* rec_next, double* f,
             const VerletArgs& va, bool pipe, hipStream_t s) {
  if (a.influence == kPairModelPmb) p2_pipe<G, VERLET, kPairModelPmb>(a, t, rec, rec_next, f, va, pipe, s);
  else if (a.influence == PDG_INFLUENCE_ONE) p2_pipe<G, VERLET, 0>(a, t, rec, rec_next, f, va, pipe, s);
  else if (a.influence == PDG_INFLUENCE_PARABOLIC_DECAY) p2_pipe<G, VERLET, 1>(a, t, rec, rec_next, f, va, pipe, s);
  else p2_pipe<G, VERLET, 2>(a, t, rec, rec_next, f, va, pipe, s);
}

}  // namespace

namespace {
template <int G>
void* persistent_fn(int influence, bool damage) {
  if (influence == kPairModelPmb)
    return damage ? (void*)k_persistent<G, true, kPairModelPmb> : (void*)k_persistent<G, false, kPairModelPmb>;
  if (influence == PDG_INFLUENCE_ONE)
    return damage ? (void*)k_persistent<G, true, 0> : (void*)k_persistent<G, false, 0>;
  if (influence == PDG_INFLUENCE_PARABOLIC_DECAY)
    return damage ? (void*)k_persistent<G, true, 1> : (void*)k_persistent<G, false, 1>;
  return damage ? (void*)k_persistent<G, true, 2> : (void*)k_persistent<G, false, 2>;
}
}  // namespace

hipError_t launch_persistent(const BondArgs& a, const TileArgs& t, NodeRec* rec0, NodeRec* rec1,
                             float* theta, int* nbroken, const VerletArgs& va, int nsteps,
                             double t0, int initial_eval, bool damage, int lanes, int* grid_out,
                             hipStream_t s) {
  if (t.ntiles <= 0) return hipErrorInvalidValue;
  if (lanes != 4 && lanes != 8 && lanes != 16) return hipErrorInvalidValue;
  void* fn = lanes == 4   ? persistent_fn<4>(a.influence, damage)
             : lanes == 8 ? persistent_fn<8>(a.influence, damage)
                          : persistent_fn<16>(a.influence, damage);
  int dev = 0, ncu = 0, per_cu = 0, coop = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
  if (e == hipSuccess && !coop) return hipErrorNotSupported;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WG, 0);
  if (e != hipSuccess) return e;
  int nb = std::min(per_cu * ncu, t.ntiles);
  if (nb >= 8) nb &= ~7;
  if (nb <= 0) return hipErrorInvalidConfiguration;
  if (grid_out) *grid_out = nb;
  BondArgs aa = a;
  TileArgs tt = t;
  VerletArgs vv = va;
  void* args[] = {&aa, &tt, &rec0, &rec1, &theta, &nbroken, &vv, &nsteps, &t0, &initial_eval};
  return hipLaunchCooperativeKernel(fn, dim3(nb), dim3(WG), args, 0, s);
}

hipError_t launch_pass1_tiled(const BondArgs& a, const TileArgs& t, NodeRec* rec, float* theta,
                              int* nbroken, bool damage, bool count_near, int lanes, bool pipe,
                              hipStream_t s) {
  if (t.ntiles <= 0) return hipSuccess;
  if (lanes == 4) p1_infl<4>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
  else if (lanes == 8) p1_infl<8>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
  else if (lanes == 16) p1_infl<16>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
  else if (lanes == 32) p1_infl<32>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
  else p1_infl<64>(a, t, rec, theta, nbroken, damage, count_near, pipe, s);
  return hipGetLastError();
}

hipError_t launch_pass2_tiled_force(const BondArgs& a, const TileArgs& t, const NodeRec* rec,
                                    double* f, int lanes, bool pipe, hipStream_t s) {
  if (t.ntiles <= 0) return hipSuccess;
  VerletArgs va{};
  if (lanes == 4) p2_infl<4, false>(a, t, rec, nullptr, f, va, pipe, s);
  else if (lanes == 8) p2_infl<8, false>(a, t, rec, nullptr, f, va, pipe, s);
  else if (lanes == 16) p2_infl<16, false>(a, t, rec, nullptr, f, va, pipe, s);
  else if (lanes == 32) p2_infl<32, false>(a, t, rec, nullptr, f, va, pipe, s);
  else p2_infl<64, false>(a, t, rec, nullptr, f, va, pipe, s);
  return hipGetLastError();
}

hipError_t launch_pass2_tiled_verlet(const BondArgs& a, const TileArgs& t, const NodeRec* rec,
                                     NodeRec* rec_next, const VerletArgs& va, int lanes, bool pipe,
                                     hipStream_t s) {
  if (t.ntiles <= 0) return hipSuccess;
  if (lanes == 4) p2_infl<4, true>(a, t, rec, rec_next, nullptr, va, pipe, s);
  else if (lanes == 8) p2_infl<8, true>(a, t, rec, rec_next, nullptr, va, pipe, s);
  else if (lanes == 16) p2_infl<16, true>(a, t, rec, rec_next, nullptr, va, pipe, s);
  else if (lanes == 32) p2_infl<32, true>(a, t, rec, rec_next, nullptr, va, pipe, s);
  else p2_infl<64, true>(a, t, rec, rec_next, nullptr, va, pipe, s);
  return hipGetLastError();
}

}  // namespace pdg
