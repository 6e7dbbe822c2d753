// The context behind the opaque pdg_ctx handle.
#pragma once

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "pdg_internal.hpp"
#include "pdg_sched.hpp"

namespace pdg {

// Runs f(lo, hi) over [0, n) in contiguous chunks on up to 16 host threads
// (one chunk in the calling thread when n < min_items).
template <class F>
inline void parallel_for(long long n, F&& f, long long min_items = 65536) {
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned nt = std::min(16u, std::max(1u, hw));
  if (n < min_items || nt == 1) { f(0LL, n); return; }
  const long long chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  th.reserve(nt);
  long long done = 0;  // chunks below `done` are handed to threads
  for (unsigned t = 0; t < nt; t++) {
    const long long lo = (long long)t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    try {
      th.emplace_back([&f, lo, hi] { f(lo, hi); });
    } catch (...) {  // no thread available: the caller runs the rest
      break;
    }
    done = hi;
  }
  if (done < n) f(done, n);
  for (auto& x : th) x.join();
}

struct QsState;   // quasi-static solver workspace (pdg_qs.hip)
struct GnnState;  // learned preconditioner (pdg_gnn.hip)
struct ContactState;  // short-range contact (pdg_contact.hip)

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::string err;
  int lanes = 4;
  bool per_pass_timing = false;
  double t_total = 0.0, t_pass1 = 0.0, t_pass2 = 0.0;
  int n_pass = 0;                 // force evaluations timed in the last run
  std::vector<hipEvent_t> tev;    // per-pass timing events

  // topology (host, reference order)
  bool have_topo = false, have_consts = false;
  int N = 0;
  long long B = 0;
  std::vector<int> nlist;
  std::vector<long long> head;

  // Hilbert order and internal CSR
  std::vector<int> perm, iperm;  // perm[k]: original id of internal node k
  std::vector<int> row, col;     // CSR over internal ids (tiled: rows in LDS-schedule order)
  std::vector<int> src;          // internal bond -> list-order bond index
  std::vector<int> rev;          // internal bond (i->j) -> internal bond (j->i)
  std::vector<int> mw;           // first mask word of each node

  // constants
  double K = 0, mu = 0, horizon = 0, sc = 0, L = 1, S = 1;
  // material model (PDG_MATERIAL_*), planar variant of the bond-based model
  // (0: 3D, 1: plane strain, 2: plane stress) and its thickness
  int material = PDG_MATERIAL_ELASTIC, planar = 0;
  double height = 0.0;
  float pmb_c = 0.0f;
  // critical stretch applied to every damage-enabled node (> 0), and its
  // piecewise-linear time function for explicit runs (empty: none)
  double sc_value = 0.0, sc_cur = 0.0;
  std::vector<double> sc_t, sc_v;
  double sc_at(double t) const;
  // position quantum of the bond kernels: q = L * 2^-qexp (records hold
  // u / q and the fixed-point x / q)
  double q = 1;
  double bond_cmax = 0.0;  // largest bond component |x_j - x_i| (any axis), from set-up
  int qexp = 0;
  int infl = 0;
  std::vector<double> x, vol, m;  // reference order

  // device state
  int cur = 0;
  int* d_row = nullptr;
  int* d_col = nullptr;
  int* d_mw = nullptr;
  unsigned long long* d_mask = nullptr;
  NodeConst* d_nc = nullptr;
  NodeRec* d_rec[2] = {nullptr, nullptr};
  float* d_theta = nullptr;
  int* d_nbroken = nullptr;
  double* d_v = nullptr;
  double* d_a = nullptr;
  double* d_f = nullptr;
  double* d_fout = nullptr;
  double* d_y = nullptr;
  double* d_inv_rho = nullptr;  // n: 1 / density (internal order)
  int* d_bcslot = nullptr;
  BcSlot* d_bcs = nullptr;
  int* d_err = nullptr;
  unsigned long long* d_near = nullptr;
  int* d_perm = nullptr;
  double* d_vol = nullptr;  // FP64 volumes (internal order)
  double* d_x = nullptr;    // FP64 reference positions (internal order, xyz)

  // LDS-staged Hilbert tiles (empty if a node's neighbourhood exceeds the
  // halo capacity; the untiled kernels are used then)
  int variant = 0;  // 0: tiled (pipelined loops), 1: untiled, 2: tiled (plain loops)
  int tile_nodes_req = 0;           // requested tile size (0: chosen from the mesh size)
  int tile_nodes = kTileNodesMax;   // tile size of the current tiling
  int ntiles = 0;
  std::vector<int> tile_start, halo_off, halo, own_off;
  std::vector<unsigned short> lcol;
  int* d_tile_start = nullptr;
  int* d_halo_off = nullptr;
  int* d_halo = nullptr;
  int* d_own_off = nullptr;
  unsigned short* d_lcol = nullptr;
  bool tiled() const { return variant != 1 && ntiles > 0; }
  bool pipe() const { return variant != 2; }
  // lanes per node of the tiled passes: at least enough that one round of a
  // workgroup covers a whole tile
  int tile_lanes() const { return std::max(lanes, kTileThreads / tile_nodes); }
  TileArgs tile_args() const;
  // pinned host staging for the drop-in calls' transfers (grown on demand)
  char* h_stage = nullptr;
  size_t h_stage_bytes = 0;
  char* stage(size_t bytes);
  int upload_y(const double* y);
  int build_csr();
  int tile_bounds(std::vector<int>& bounds) const;
  int build_tiles(const std::vector<int>& bounds);
  void order_bonds();
  SchedStats sched;  // LDS service groups and their conflict cost, before and after order_bonds

  // thermal strains ("Thermal Expansion Coefficient" and the Temperature_Change
  // field dT(t) = dT1 t + dT0, reference order; the quasi-static kernels
  // evaluate it at temp_time)
  double alpha_t = 0.0;
  std::vector<double> dT0, dT1;
  double temp_time = 0.0;
  unsigned thermal_epoch = 0;
  bool thermal() const { return alpha_t != 0.0 && !dT0.empty(); }

  QsState* qs = nullptr;
  GnnState* gnn = nullptr;
  ContactState* contact = nullptr;  // null: no contact model
  unsigned gnn_epoch = 0;  // bumped whenever the network's buffers are (re)allocated
  bool qs_abs_tol = false;         // Newton tolerance is absolute (no reaction scaling)
  double qs_tol_multiplier = 1.0;  // reaction norm used by the last pdg_qs_newton

  void free_gnn();
  void free_contact();
  double* qs_dinv();
  const signed char* qs_fixed();
  int fail(int code, const std::string& msg);
  int hip(hipError_t e, const char* what);
  void free_device();
  void free_extra();
  BondArgs bond_args() const;
  hipError_t pass1(NodeRec* rec, int* nbroken, bool damage, bool count_near);
  int upload_mask(const double* bd, bool check_symmetric);
  int download_mask(double* bd);
  int download_vec3(const double* dsrc, double* out);
  int upload_vec3(const double* in, double* ddst);
};

// short-range contact (pdg_contact.hip): whether a contact model is set, its
// force density of the last explicit evaluation (device, internal order), and
// evaluation k (0: initial) of an explicit run with step dt: a search in the
// configuration of rec_prev (the step's start) when one is due, then the
// force at y = x + u_prev + dt v with the context's half-step velocities v
// (at k = 0: search and force at rec).
bool contact_on(const Ctx* c);
const double* contact_force_ptr(const Ctx* c);
int contact_step(Ctx* c, int k, double dt, const NodeRec* rec_prev, const NodeRec* rec);

// quasi-static solver (pdg_qs.hip) and learned preconditioner (pdg_gnn.hip)
int qs_solve(Ctx* c, double* x, const double* b, int precond, double tol, int maxit, int* iters,
             double* rel);
int gnn_apply(Ctx* c, const double* r, double* z);
// K w at the prepared tangent state (device vectors, internal order; kinematic
// rows return w) and a device dot product read back to the host.
int qs_tangent_apply(Ctx* c, const double* w, double* out);
int qs_dot(Ctx* c, const double* x, const double* y, double* out);
// uploads u (reference order) and forms the block-Jacobi inverse at u
int pdg_qs_prepare(Ctx* c, const double* u);

}  // namespace pdg
