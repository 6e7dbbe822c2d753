// LDS-staged bond passes over Hilbert tiles.
//
// One workgroup (8 waves) owns a tile: a contiguous range of at most 128
// Hilbert-ordered nodes (32 or 64 on small meshes), ordered by degree inside
// the tile. It first stages its halo (the union of the tile's nodes and all
// their neighbours, at most 800 nodes) in LDS with one 64-byte record load
// per halo node, then runs group-of-G-lanes bond loops that read neighbours
// from LDS through 16-bit local indices, three 16-byte LDS reads per
// neighbour. Both passes stage the exact inputs (fixed-point positions, FP64
// displacements), so both ends of a bond compute the same stretch in pass 1
// and every bond vector keeps full FP32 accuracy in pass 2. Per bond, HBM and L2 see a
// 2-byte index and one bit of the damage mask instead of a 4-byte index plus
// a 64-byte record gather.
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
__device__ __forceinline__ float omega_t(float z2, float hhat) {
  return omega_z2(INFL, z2, hhat);
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
  const unsigned long long* __restrict__ p = mask + (unsigned)w0;
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
// hold staged node data; such a bond's arithmetic may still give NaN (a slot
// holding the node itself has xi = 0), so every use of a masked-out bond's
// results goes through a select, never a multiplication by zero.
template <int G>
__device__ __forceinline__ void load_lcol(const unsigned short* __restrict__ lcol, int rs, int l,
                                          int (&lj)[MAXDEG / G]) {
  const unsigned short* __restrict__ p = lcol + (unsigned)(rs + l);
#pragma unroll
  for (int it = 0; it < MAXDEG / G; it++) lj[it] = (int)p[it * G];
}

// LDS staging arrays of one workgroup: three arrays of 16-byte slots, so
// one byte offset (16 x the halo position) addresses every array. Each pass
// stages its own view of the halo:
//   pass 1 (exact, so both ends of a bond compute the same stretch):
//     c: displacement z (FP64, units of q), a: fixed-point position and
//     volume, b: displacement x and y (FP64);
//   pass 2: as pass 1, with the pair-force terms V (alpha - p) and V p
//     after the displacement z in array c.
// Each array has one slot past the capacity: staging threads without a slot
// write there, so staging needs no branch (no bond addresses it).
struct TileLds {
  uint4 raw_c[HMAX + 1];  // array c first, so its reads need no address add
  uint4 raw_a[HMAX + 1];
  uint4 raw_b[HMAX + 1];
  float sf[3 * kTileNodesMax];  // pass 2: the tile's forces, for the Verlet epilogue
};
static_assert(sizeof(uint4) == kLdsSlot, "16-byte LDS slots");

__device__ __forceinline__ char* lds_c(TileLds& L) { return reinterpret_cast<char*>(L.raw_c); }
__device__ __forceinline__ char* lds_a(TileLds& L) { return reinterpret_cast<char*>(L.raw_a); }
__device__ __forceinline__ char* lds_b(TileLds& L) { return reinterpret_cast<char*>(L.raw_b); }
__device__ __forceinline__ const char* lds_c(const TileLds& L) {
  return reinterpret_cast<const char*>(L.raw_c);
}
__device__ __forceinline__ const char* lds_a(const TileLds& L) {
  return reinterpret_cast<const char*>(L.raw_a);
}
__device__ __forceinline__ const char* lds_b(const TileLds& L) {
  return reinterpret_cast<const char*>(L.raw_b);
}

// jb is the neighbour's byte offset in the slot arrays (local index times
// kLdsSlot, as stored in the tile index array).
template <typename T>
__device__ __forceinline__ T lds_at(const char* arr, int jb) {
  return *reinterpret_cast<const T*>(arr + jb);
}
// One 16-byte slot as a single ds_read_b128. The read is volatile so that
// the compiler keeps it whole: combining two of its dwords into a double
// otherwise becomes a second, 8-byte read of the same address, issued after
// a wait for the first (both write the same registers).
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const volatile u32x4 lds_u32x4;
__device__ __forceinline__ u32x4 lds_slot(const char* arr, int jb) {
  return *(lds_u32x4*)(arr + jb);
}
// The same for an 8-byte value (kept where it is issued, not moved to its use).
typedef __attribute__((address_space(3))) const volatile double lds_f64;
__device__ __forceinline__ double lds_dbl(const char* arr, int jb) {
  return *(lds_f64*)(arr + jb);
}
__device__ __forceinline__ double dbl(unsigned lo, unsigned hi) {
  return __hiloint2double((int)hi, (int)lo);
}
template <typename T>
__device__ __forceinline__ void lds_put(char* arr, int h, const T& v) {
  *reinterpret_cast<T*>(arr + kLdsSlot * h) = v;
}

// One neighbour's staged data, loaded one bond ahead of its use by the
// pipelined bond loops so LDS latency overlaps the previous bond's math.
struct Nb {  // pass 1
  unsigned x0, x1, x2;
  float vol;
  double u0, u1, u2;
};
struct Nb2 {  // pass 2
  unsigned x0, x1, x2;
  float vol;
  double u0, u1, u2;
  float vd, vp;  // V (alpha - p), V p
};

__device__ __forceinline__ Nb nb_load(const TileLds& L, int jb) {
  Nb n;
  const u32x4 a = lds_slot(lds_a(L), jb);
  const u32x4 b = lds_slot(lds_b(L), jb);
  n.x0 = a.x; n.x1 = a.y; n.x2 = a.z; n.vol = __uint_as_float(a.w);
  n.u0 = dbl(b.x, b.y);
  n.u1 = dbl(b.z, b.w);
  n.u2 = lds_dbl(lds_c(L), jb);
  return n;
}

__device__ __forceinline__ Nb2 nb2_load(const TileLds& L, int jb) {
  Nb2 n;
  // three whole 16-byte reads (8-byte halves would conflict at this stride)
  const u32x4 a = lds_slot(lds_a(L), jb);
  const u32x4 b = lds_slot(lds_b(L), jb);
  const u32x4 c = lds_slot(lds_c(L), jb);
  n.x0 = a.x; n.x1 = a.y; n.x2 = a.z; n.vol = __uint_as_float(a.w);
  n.u0 = dbl(b.x, b.y);
  n.u1 = dbl(b.z, b.w);
  n.u2 = dbl(c.x, c.y);
  n.vd = __uint_as_float(c.z);
  n.vp = __uint_as_float(c.w);
  return n;
}

// The record fields the LDS slots hold (48 of the record's 64 bytes).
struct SlotVals {
  double u0, u1, u2;
  unsigned x0, x1, x2;
  float vol, alpha, p;
};

__device__ __forceinline__ SlotVals slot_vals(const NodeRec* __restrict__ r) {
  SlotVals v;
  v.u0 = r->u[0]; v.u1 = r->u[1]; v.u2 = r->u[2];
  v.x0 = r->xq[0]; v.x1 = r->xq[1]; v.x2 = r->xq[2];
  v.vol = r->vol; v.alpha = r->alpha; v.p = r->p;
  return v;
}

// Staging of the tile's halo into LDS: slot h < hn from the record of node
// halo[h0 + h]; slots hn..HMAX-1 repeat the first halo node's record (finite
// values: masked-out bonds may address them, and every use of a masked-out
// bond's results goes through a select). Each thread handles slots threadIdx.x + r * WG. The three phases are
// separate calls so a kernel can put every independent load of its prologue
// in flight together: the halo indices (halo_idx), then the records
// (halo_load, with the row data that depends on the node's CSR range), then
// the LDS writes (halo_write). Pass 1 stages the fixed-point position and
// volume (array a), the FP64 displacement (b: x, y; c: z); pass 2 adds
// V (alpha - p) and V p after z in array c.
constexpr int kSlotsPerThread = (HMAX + WG - 1) / WG;

// Loads are unconditional (a slot past the halo reads its first entry), so
// no branch makes the compiler wait between the phases. A tile's halo holds
// at least its own nodes (hn >= 1).
struct HaloStage {
  int hn;
  int idx[kSlotsPerThread];
  SlotVals v[kSlotsPerThread];
};

__device__ __forceinline__ void halo_idx(HaloStage& hs, const TileArgs& t, int h0, int hn) {
  hs.hn = hn;
#pragma unroll
  for (int r = 0; r < kSlotsPerThread; r++) {
    const int h = threadIdx.x + r * WG;
    hs.idx[r] = t.halo[(unsigned)(h0 + (h < hn ? h : 0))];
  }
}

__device__ __forceinline__ void halo_load(HaloStage& hs, const NodeRec* __restrict__ rec) {
#pragma unroll
  for (int r = 0; r < kSlotsPerThread; r++) hs.v[r] = slot_vals(&rec[(unsigned)hs.idx[r]]);
}

template <bool PASS2>
__device__ __forceinline__ void halo_write(TileLds& L, const HaloStage& hs) {
#pragma unroll
  for (int r = 0; r < kSlotsPerThread; r++) {
    const int hh = threadIdx.x + r * WG;
    const int h = hh < HMAX ? hh : HMAX;  // the spare slot
    const SlotVals& v = hs.v[r];
    lds_put(lds_a(L), h, make_uint4(v.x0, v.x1, v.x2, __float_as_uint(v.vol)));
    lds_put(lds_b(L), h, make_double2(v.u0, v.u1));
    if (PASS2) {
      const float dp[2] = {v.vol * (v.alpha - v.p), v.vol * v.p};
      lds_put(lds_c(L), h, make_double2(v.u2, *reinterpret_cast<const double*>(dp)));
    } else {
      lds_put(lds_c(L), h, make_double2(v.u2, 0.0));
    }
  }
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
  int rs, re, deg, w0;
  unsigned long long m[NWORDS];
  int lj[MAXDEG / G];
  NodeConst c;
};

// Row data in two phases: the node's CSR range, mask-word offset and
// constants (row_scalars), then the mask words and local indices they
// address (row_arrays).
// The loads are unconditional (an invalid group reads the valid node
// `fallback`) and nothing computes on their values until row_arrays, so a
// prologue's loads stay in flight together (no branch or early use makes the
// compiler wait for them).
template <int G>
__device__ __forceinline__ void row_scalars(const BondArgs& a, int node, bool valid, int fallback,
                                            RowData<G>& r) {
  const int nd = valid ? node : fallback;
  // unsigned indices: no sign extension in the 64-bit address arithmetic
  r.rs = a.row[(unsigned)nd];
  r.re = a.row[(unsigned)nd + 1u];
  r.w0 = a.mw[(unsigned)nd];
  r.c = a.nc[(unsigned)nd];
}

template <int G>
__device__ __forceinline__ void row_arrays(const BondArgs& a, const TileArgs& t, int l, bool valid,
                                           RowData<G>& r) {
  r.deg = valid ? r.re - r.rs : 0;
  if (!valid) r.c = NodeConst{0.0f, 0.0f, 0.0f, 1.0f};
  load_words(a.mask, r.w0, r.deg, r.m);
  load_lcol<G>(t.lcol, r.rs, l, r.lj);
}

template <int G>
__device__ __forceinline__ void row_load(const BondArgs& a, const TileArgs& t, int node,
                                         bool valid, int fallback, int l, RowData<G>& r) {
  row_scalars<G>(a, node, valid, fallback, r);
  row_arrays<G>(a, t, l, valid, r);
}

// Whether one round of node groups covers a tile. A tile holds at least one
// node and at most kTileNodesMax (the tile builder's invariant), so the
// node-group loop then runs exactly once with no guard, and the row loads
// are not duplicated for later rounds.
template <int G>
__device__ __forceinline__ constexpr bool one_round() {
  return WG / G >= kTileNodesMax;
}

// Bond-loop trip count of a wave: inside a tile nodes are ordered by
// descending degree (checked when the tiles are built), so the wave's first
// node group has the largest degree.
template <int G>
__device__ __forceinline__ int wave_trips(int deg) {
  return (__builtin_amdgcn_readfirstlane(deg) + G - 1) / G;
}

// Lane masks of the wave (compare results kept in scalar registers):
// predicate codes of the compare builtins.
enum { CMP_FOGT = 2, CMP_FOLE = 5, CMP_INE = 33 };

// A lane's intact bits: the complement of the mask words m (which carry set
// bits past the row's end) shifted down by the lane's l, as 32-bit halves.
// Bond k = it * G + l then sits at bit (it * G) & 31 of half (it * G) >> 5:
// a compile-time position, so testing it takes no per-lane shift.
template <int G>
__device__ __forceinline__ void lane_intact(const unsigned long long (&m)[NWORDS], int l,
                                            unsigned (&ih)[2 * NWORDS]) {
#pragma unroll
  for (int w = 0; w < NWORDS; w++) {
    const unsigned long long v = ~m[w] >> l;
    ih[2 * w] = (unsigned)v;
    ih[2 * w + 1] = (unsigned)(v >> 32);
  }
}

// Lanes whose bond it * G + l exists and is not marked broken.
template <int G>
__device__ __forceinline__ unsigned long long intact_lanes(const unsigned (&ih)[2 * NWORDS],
                                                           int it) {
  const int base = it * G;
  return __builtin_amdgcn_uicmp(ih[base >> 5] & (1u << (base & 31)), 0u, CMP_INE);
}

// All ones when the lane's bond it * G + l exists and is not marked broken,
// zero otherwise: one sign-filled bit-field extract. Written as an asm
// statement (an ALU instruction, no memory access) because the compiler turns
// the extract-and-mask into a compare and a select, one instruction more.
template <int G>
__device__ __forceinline__ unsigned intact_fill(const unsigned (&ih)[2 * NWORDS], int it) {
  const int base = it * G;
  unsigned m;
  asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"(ih[base >> 5]), "i"(base & 31));
  return m;
}

// Pass 1 (damage + dilatation) of one tile; every thread of the workgroup
// calls it with the same tile.
// Pass 1's pipelined bond loop as a chain of template instances, as
// P2Loop: step<IT> reads the slots of bond IT + 2, computes bond IT's
// stretch, break decision, near-critical count and dilatation term, and
// returns at the wave's trip count.
template <int G, bool DAMAGE, bool NEAR, int INFL>
struct P1Loop {
  const TileLds& L;
  const int (&lj)[MAXDEG / G];
  const unsigned (&ih)[2 * NWORDS];
  int nmax;
  unsigned xi0, xi1, xi2;
  double ui0, ui1, ui2;
  float scz, nlim, hhat;
  Nb nb[3];
  unsigned long long balv[MAXDEG / G];
  float acc = 0.0f;
  unsigned nnear_w = 0;

  template <int IT>
  __device__ __forceinline__ void step() {
    if constexpr (IT < MAXDEG / G) {
      if (IT >= nmax) return;
      if constexpr (IT + 2 < MAXDEG / G) nb[(IT + 2) % 3] = nb_load(L, lj[IT + 2]);
      const Nb& cur = nb[IT % 3];
      // lanes whose bond exists and is unbroken; the masks stay in scalar
      // registers, so combining them costs no vector instructions
      const unsigned long long okm = intact_lanes<G>(ih, IT);
      const BondS b = bond_stretch(bond_q(xi0, xi1, xi2, ui0, ui1, ui2, cur.x0, cur.x1, cur.x2,
                                          cur.u0, cur.u1, cur.u2));
      unsigned long long brkm = 0ull;  // lanes whose bond breaks now
      if (DAMAGE) {
        if (NEAR) {
          // intact bonds with |s - s_c| <= near_rel * s_c, counted as a
          // wave-wide lane mask (scalar population count)
          const unsigned long long nm =
              okm & __builtin_amdgcn_fcmpf(fabsf(b.s - scz), nlim, CMP_FOLE);
          nnear_w += (unsigned)__builtin_popcountll(nm);
        }
        brkm = okm & __builtin_amdgcn_fcmpf(b.s, scz, CMP_FOGT);
      }
      // w zeta e V = w z2 s V
      float term = b.z2 * b.s;
      if (INFL != PDG_INFLUENCE_ONE) term *= omega_t<INFL>(b.z2, hhat);
      const float nacc = fmaf(term, cur.vol, acc);
      acc = __builtin_amdgcn_inverse_ballot_w64(okm & ~brkm) ? nacc : acc;
      if (DAMAGE) balv[IT] = brkm;
      step<IT + 1>();
    }
  }
};

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
  // independent loads first (halo indices, CSR ranges), then what they
  // address (records, mask words, local indices), then the LDS writes
  HaloStage hs;
  halo_idx(hs, t, h0, hn);
  RowData<G> rd;
  row_scalars<G>(a, n0 + grp, n0 + grp < n1, n0, rd);
  halo_load(hs, rec);
  row_arrays<G>(a, t, l, n0 + grp < n1, rd);
  halo_write<false>(L, hs);
  __syncthreads();
  const int own = t.own_off[tile];
  for (int base = n0; one_round<G>() ? base == n0 : base < n1; base += NG) {
    const int node = base + grp;
    const bool valid = node < n1;
    if (base != n0) row_load<G>(a, t, node, valid, n0, l, rd);
    const int deg = rd.deg, w0 = rd.w0;
    const unsigned long long (&mo)[NWORDS] = rd.m;
    const int (&lj)[MAXDEG / G] = rd.lj;
    const NodeConst c = rd.c;
    unsigned xi0 = 0u, xi1 = 0u, xi2 = 0u;
    double ui0 = 0.0, ui1 = 0.0, ui2 = 0.0;
    if (valid) {
      const int li = own + (node - n0);
      const Nb me = nb_load(L, kLdsSlot * li);
      xi0 = me.x0; xi1 = me.x1; xi2 = me.x2;
      ui0 = me.u0; ui1 = me.u1; ui2 = me.u2;
    }
    unsigned long long mn[NWORDS];
#pragma unroll
    for (int w = 0; w < NWORDS; w++) mn[w] = 0ull;
    // a time-dependent critical stretch replaces every damage-enabled node's
    const float scz = (a.sc_now > 0.0f && c.sc > 0.0f) ? a.sc_now : c.sc;
    const float nlim = a.near_rel * scz;
    const int nmax = wave_trips<G>(deg);
    float acc = 0.0f;
    int nnear = 0;               // plain loops: per lane
    unsigned nnear_w = 0;        // pipelined loops: per wave (scalar)
    if (PIPE) {
      // branch-free bond loop: every lane computes, absent and broken bonds
      // are selected away; the next bond's LDS data is in flight. The
      // per-iteration ballots of new breaks stay wave-uniform (scalar
      // registers) and are merged into the group's mask words after the loop.
      // the slots of the bond after next are read whenever a bond is
      // computed, as in pass 2
      unsigned ih[2 * NWORDS];
      lane_intact<G>(mo, l, ih);
      P1Loop<G, DAMAGE, NEAR, INFL> lp{L, lj, ih, nmax, xi0, xi1, xi2, ui0, ui1, ui2,
                                       scz, nlim, c.hhat, {}, {}};
#pragma unroll
      for (int it = 0; it < MAXDEG / G; ++it) lp.balv[it] = 0ull;
      lp.nb[0] = nb_load(L, lj[0]);
      lp.nb[1] = nb_load(L, lj[1]);
      lp.template step<0>();
      acc = lp.acc;
      nnear_w = lp.nnear_w;
      const unsigned long long (&balv)[MAXDEG / G] = lp.balv;
      if (DAMAGE) {
        // new breaks are rare: a wave-uniform test skips the iterations
        // without any
#pragma unroll
        for (int it = 0; it < MAXDEG / G; ++it) {
          if (balv[it] == 0ull) continue;
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
          const Nb r = nb_load(L, lj[it]);
          const BondS b = bond_stretch(bond_q(xi0, xi1, xi2, ui0, ui1, ui2, r.x0, r.x1, r.x2,
                                              r.u0, r.u1, r.u2));
          if (DAMAGE) {
            if (NEAR && !brk && fabsf(b.s - scz) <= nlim) nnear++;
            brk = brk || (b.s > scz);
          }
          const float w = omega_t<INFL>(b.z2, c.hhat);
          if (!brk) acc = fmaf(w * b.z2 * b.s, r.vol, acc);
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
      rec[(unsigned)node].p = c.kp * acc;
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
// Pass 2's pipelined bond loop as a chain of template instances, bond IT
// computed by step<IT>: each instance reads the slots of bond IT + 2 and
// returns at the wave's trip count, so bond IT + 1 is reached only from bond
// IT (a straight chain: each bond waits only for its own reads).
template <int G, int INFL>
struct P2Loop {
  const TileLds& L;
  const int (&lj)[MAXDEG / G];
  const unsigned (&ih)[2 * NWORDS];
  int nmax;
  unsigned xi0, xi1, xi2;
  double ui0, ui1, ui2;
  float p_i, d_i, hhat, pmb_c;
  Nb2 nb[3];
  float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;

  template <int IT>
  __device__ __forceinline__ void step() {
    if constexpr (IT < MAXDEG / G) {
      if (IT >= nmax) return;
      if constexpr (IT + 2 < MAXDEG / G) nb[(IT + 2) % 3] = nb2_load(L, lj[IT + 2]);
      const Nb2& cur = nb[IT % 3];
      const BondK b = bond_k(bond_q(xi0, xi1, xi2, ui0, ui1, ui2, cur.x0, cur.x1, cur.x2,
                                    cur.u0, cur.u1, cur.u2));
      const float tv = pair_tv<INFL>(b, p_i, d_i, cur.vol, cur.vp, cur.vd, hhat, pmb_c);
      // a bitwise select: a masked-out bond's NaN or infinity becomes +0
      const float tt = __uint_as_float(__float_as_uint(tv) & intact_fill<G>(ih, IT));
      f0 = fmaf(tt, b.y0, f0);
      f1 = fmaf(tt, b.y1, f1);
      f2 = fmaf(tt, b.y2, f2);
      step<IT + 1>();
    }
  }
};

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
  // independent loads first (halo indices, CSR ranges, this thread's Verlet
  // inputs: component q = threadIdx.x, a tile has at most 3 * 128 <= WG
  // components), then what they address (records, mask words, local
  // indices), then the LDS writes
  HaloStage hs;
  halo_idx(hs, t, h0, hn);
  RowData<G> rd;
  row_scalars<G>(a, n0 + grp, n0 + grp < n1, n0, rd);
  const int q = threadIdx.x;
  const bool has_q = q < 3 * (n1 - n0);
  VerletPre vp{0.0, 0.0, 0.0, -1};
  double u_own = 0.0;  // this thread's component of the FP64 displacement (units of q)
  if (VERLET) {
    // unconditional (a thread past the tile's components reads component 0)
    const int qq = has_q ? q : 0;
    vp = verlet_pre(va, n0 + qq / 3, qq - 3 * (qq / 3));
    u_own = rec[(unsigned)(n0 + qq / 3)].u[qq - 3 * (qq / 3)];
  }
  halo_load(hs, rec);
  row_arrays<G>(a, t, l, n0 + grp < n1, rd);
  halo_write<true>(L, hs);
  __syncthreads();
  const int own = t.own_off[tile];
  for (int base = n0; one_round<G>() ? base == n0 : base < n1; base += NG) {
    const int node = base + grp;
    const bool valid = node < n1;
    if (base != n0) row_load<G>(a, t, node, valid, n0, l, rd);
    const int deg = rd.deg;
    const unsigned long long (&mo)[NWORDS] = rd.m;
    const int (&lj)[MAXDEG / G] = rd.lj;
    const float hhat = rd.c.hhat;
    Nb2 me{};
    float d_i = 0.0f, p_i = 0.0f;
    if (valid) {
      // the node's own alpha - p and p from its staged V (alpha - p), V p
      me = nb2_load(L, kLdsSlot * (own + (node - n0)));
      const float rv = me.vol > 0.0f ? __builtin_amdgcn_rcpf(me.vol) : 0.0f;
      d_i = me.vd * rv;
      p_i = me.vp * rv;
    }
    const unsigned xi0 = me.x0, xi1 = me.x1, xi2 = me.x2;
    const double ui0 = me.u0, ui1 = me.u1, ui2 = me.u2;
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    if (PIPE) {
      // the slots of the bond after next are read whenever a bond is
      // computed (every lj entry, past a row's end too, is a slot of the
      // staging arrays)
      unsigned ih[2 * NWORDS];
      lane_intact<G>(mo, l, ih);
      P2Loop<G, INFL> lp{L, lj, ih, wave_trips<G>(deg), xi0, xi1, xi2, ui0, ui1, ui2,
                         p_i, d_i, hhat, a.pmb_c, {}};
      lp.nb[0] = nb2_load(L, lj[0]);
      lp.nb[1] = nb2_load(L, lj[1]);
      lp.template step<0>();
      f0 = lp.f0; f1 = lp.f1; f2 = lp.f2;
    } else {
#pragma unroll
    for (int it = 0; it < MAXDEG / G; ++it) {
      const int k = it * G + l;
      if (k < deg) {
        const bool brk = mask_bit<G>(mo, it, l);
        const Nb2 r = nb2_load(L, lj[it]);
        const BondK b = bond_k(bond_q(xi0, xi1, xi2, ui0, ui1, ui2, r.x0, r.x1, r.x2, r.u0,
                                      r.u1, r.u2));
        float tt = pair_tv<INFL>(b, p_i, d_i, r.vol, r.vp, r.vd, hhat, a.pmb_c);
        tt = brk ? 0.0f : tt;
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
      verlet_epilogue(va, n0 + k, l3, L.sf[q], u_own, rec_next, vp);
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
