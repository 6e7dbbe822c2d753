// Internal declarations shared by the pdgpu HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pdgpu.h"

namespace pdg {

// One node's bond-gather record, 64 bytes (half a 128-byte cache line):
// everything a neighbour needs from node j in either bond pass, so a gather
// touches exactly one aligned 64-byte segment.
//   u: displacement / q (FP64), q = L * 2^-qexp the context's position
//      quantum (a power of two, so the scaling is exact);
//   xq: the low 32 bits of round(x / q), the reference position in fixed
//      point: for any bond the difference xq_j - xq_i, taken modulo 2^32, is
//      the exact integer (x_j - x_i) / q of the rounded positions, because
//      every bond component is shorter than 2^27 q; the same difference with
//      i and j swapped is its exact negative;
//   vol: V / S, alpha: 15 mu / m * L * S, p: theta * (3K/m - alpha/3) * L * S.
// L and S are powers of two, so the scalings are exact.
struct __align__(64) NodeRec {
  double u[3];  // first: the per-step displacement write fills one 32-byte sector
  unsigned xq[3];
  float vol;
  float alpha;
  float p;
  float pad[4];
};
static_assert(sizeof(NodeRec) == 64, "NodeRec must be 64 bytes");

// Per-node constants of the own-node side of both passes:
//   ctheta = 3 L^2 S / m      (theta = ctheta * sum_j w (1-d) zeta e V)
//   kp     = ctheta * (3K/m - alpha/3) * L * S   (p = kp * same sum)
//   sc     = critical stretch (0: damage disabled for this node)
// The bond kernels measure lengths in units of q (= L * 2^-qexp): their sums
// are 2^(2 qexp) times those in units of L, so ctheta and kp carry the factor
// 2^(-2 qexp).
struct __align__(16) NodeConst {
  float ctheta;
  float kp;
  float sc;
  float hhat;  // horizon / q (influence functions)
};

// Boundary-condition slot of a node: u_c(t) = c1[c] * t + c0[c] for the
// coordinates whose bit is set in `mask`.
struct BcSlot {
  double c0[3];
  double c1[3];
  int mask;
  int pad;
};

// Arguments shared by the two bond passes.
struct BondArgs {
  int n;                       // nodes
  const int* row;              // CSR row offsets, n+1
  const int* col;              // CSR neighbour ids (internal order), B
  const int* mw;               // first 64-bit mask word of each node, n+1
  unsigned long long* mask;    // bond-broken bits
  const NodeConst* nc;         // per-node constants
  float near_rel;              // near-critical band (relative)
  unsigned long long* near_ctr;
  int influence;               // PDG_INFLUENCE_*, or kPairModelPmb for the bond-based model
  float fq;                    // 2^-qexp: pass-2 FP32 force sums are in units of 1/q
  float pmb_c;                 // bond-based model: c * S * 2^qexp (c = 18K/(pi delta^4) in 3D)
  float sc_now;                // > 0: the critical stretch of every damage-enabled node
};

// The bond kernels' influence-function slot doubles as the pair-model switch:
// this value selects the elastic bond-based (PMB) force, whose bonds carry no
// influence function and no dilatation term.
constexpr int kPairModelPmb = 3;

struct VerletArgs {
  const double* inv_rho;  // n: 1 / density
  double* v;          // 3n, FP64
  double* a;          // 3n, FP64 (written on the last step only)
  double* f;          // 3n, FP64 force density (written on the last step only)
  const double* fadd; // 3n, FP64 force density added to the internal one (contact), or null
  const int* bc_slot; // n, -1 if none
  const BcSlot* bcs;
  double dt, t_cur, t_prev;
  double uq;          // record displacement unit q (u = rec.u * uq)
  double inv_uq;      // 1.0 / uq (exact: a power of two)
  double inv_dtbc;    // 1.0 / (t_cur - t_prev), rounded once like the reference's division
  int end_kick;       // finish the current step: v += dt/2 a
  int start_step;     // begin the next step: v += dt/2 a, BC, u += dt v
  int* err;           // set to 1 on a non-finite force
};

// Hilbert tiles of the LDS-staged kernels: a tile is a contiguous range of
// internal nodes [tile_start[t], tile_start[t+1]); its halo (sorted internal
// ids, own nodes included) is staged in LDS; lcol[b] is kLdsSlot times the
// position of bond b's neighbour inside its tile's halo (its byte offset in
// the 16-byte-slot LDS arrays). Tiling needs every node degree <= 128.
constexpr int kTileHaloMax = 800;
constexpr int kLdsSlot = 16;  // bytes per halo slot in each LDS array
static_assert(kLdsSlot * kTileHaloMax <= 65536, "halo byte offsets must fit 16 bits");
constexpr int kTileNodesMax = 128;
constexpr int kTileThreads = 512;  // threads per tiled-pass workgroup
constexpr int kMaskPad = 2;    // >= mask words per node in the tiled kernels (degree <= 128)
constexpr int kLcolPad = 128;  // >= the tiled kernels' maximum degree

struct TileArgs {
  int ntiles;
  const int* tile_start;      // ntiles + 1
  const int* halo_off;        // ntiles + 1
  const int* halo;            // halo node ids
  const int* own_off;         // position of tile_start[t] in its halo
  const unsigned short* lcol; // B
};

hipError_t launch_pass1_tiled(const BondArgs& a, const TileArgs& t, NodeRec* rec, float* theta,
                              int* nbroken, bool damage, bool count_near, int lanes, bool pipe,
                              hipStream_t s);
hipError_t launch_pass2_tiled_force(const BondArgs& a, const TileArgs& t, const NodeRec* rec,
                                    double* f, int lanes, bool pipe, hipStream_t s);
hipError_t launch_pass2_tiled_verlet(const BondArgs& a, const TileArgs& t, const NodeRec* rec,
                                     NodeRec* rec_next, const VerletArgs& va, int lanes, bool pipe,
                                     hipStream_t s);

// Persistent cooperative explicit run (pdg_tiles.hip): steps 1..nsteps
// (plus the initial evaluation if initial_eval) in one launch. Ends with the
// state in rec[(evaluations) % 2].
hipError_t launch_persistent(const BondArgs& a, const TileArgs& t, NodeRec* rec0, NodeRec* rec1,
                             float* theta, int* nbroken, const VerletArgs& va, int nsteps,
                             double t0, int initial_eval, bool damage, int lanes, int* grid_out,
                             hipStream_t s);

// Launchers (pdg_bonds.hip). `lanes` = lanes per node (16, 32 or 64; smaller
// values run as 16).
hipError_t launch_pass1(const BondArgs& a, NodeRec* rec, float* theta, int* nbroken,
                        bool damage, bool count_near, int lanes, hipStream_t s);
hipError_t launch_pass2_force(const BondArgs& a, const NodeRec* rec, double* f, int lanes,
                              hipStream_t s);
hipError_t launch_pass2_verlet(const BondArgs& a, const NodeRec* rec, NodeRec* rec_next,
                               const VerletArgs& va, int lanes, hipStream_t s);
hipError_t launch_verlet_prologue(int n, NodeRec* rec, NodeRec* rec_next, const VerletArgs& va,
                                  hipStream_t s);
hipError_t launch_weighted_volume(int n, const double* x, const double* vol, const int* nlist,
                                  const long long* head, double horizon, int influence, double* m,
                                  hipStream_t s);

// Looks up this device's address of the scratch array behind the
// allocation-free entry point (pdg_dt.hip); the device must be current.
hipError_t dt_scratch_init(int device);

}  // namespace pdg
