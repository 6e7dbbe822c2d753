// C-ABI of the MI355X peridynamic force path: context, topology, constants,
// drop-in computeForce / computeDamage, and the device-resident explicit
// velocity-Verlet driver.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "pdg_internal.hpp"
#include "pdg_ctx.hpp"

using namespace pdg;

namespace {

// Records' u = (y - x) / uq from current positions y (reference order);
// x is the FP64 reference position in internal order.
__global__ void k_set_u(int n, const double* __restrict__ y, const int* __restrict__ perm,
                        const double* __restrict__ x, NodeRec* __restrict__ rec, double uq) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int k = t / 3, c = t - 3 * (t / 3);
  const int p = perm[k];
  rec[k].u[c] = (y[3 * p + c] - x[t]) * (1.0 / uq);
}

__global__ void k_set_u_disp(int n, const double* __restrict__ u, const int* __restrict__ perm,
                             NodeRec* __restrict__ rec, double uq) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int k = t / 3, c = t - 3 * (t / 3);
  rec[k].u[c] = u[3 * perm[k] + c] * (1.0 / uq);
}

// Hilbert index of a 21-bit cell triple (Skilling's transpose algorithm:
// undo excess work, Gray-encode, then interleave with axis 0 most significant).
uint64_t spread3(uint64_t v);
uint64_t hilbert3(uint64_t a, uint64_t b, uint64_t c) {
  uint32_t X[3] = {(uint32_t)a, (uint32_t)b, (uint32_t)c};
  const uint32_t M = 1u << 20;
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  return spread3(X[2]) | (spread3(X[1]) << 1) | (spread3(X[0]) << 2);
}


uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

}  // namespace

namespace pdg {

int Ctx::fail(int code, const std::string& msg) {
  err = msg;
  return code;
}

int Ctx::hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return PDG_OK;
  err = std::string(what) + ": " + hipGetErrorString(e);
  return PDG_E_HIP;
}

void Ctx::free_device() {
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  F(d_row); F(d_col); F(d_mw); F(d_mask); F(d_nc); F(d_rec[0]); F(d_rec[1]); F(d_theta);
  F(d_nbroken); F(d_v); F(d_a); F(d_f); F(d_inv_rho); F(d_bcslot); F(d_bcs); F(d_err); F(d_near);
  F(d_perm); F(d_fout); F(d_y); F(d_vol); F(d_x);
  F(d_tile_start); F(d_halo_off); F(d_halo); F(d_own_off); F(d_lcol);
  d_tile_start = d_halo_off = d_halo = d_own_off = nullptr;
  d_lcol = nullptr;
  d_row = d_col = d_mw = d_nbroken = d_bcslot = d_err = d_perm = nullptr;
  d_mask = d_near = nullptr;
  d_nc = nullptr; d_rec[0] = d_rec[1] = nullptr; d_theta = nullptr;
  d_v = d_a = d_f = d_inv_rho = d_fout = d_y = d_vol = d_x = nullptr;
  d_bcs = nullptr;
}

hipError_t Ctx::pass1(NodeRec* rec, int* nbroken, bool damage, bool count_near) {
  if (tiled())
    return launch_pass1_tiled(bond_args(), tile_args(), rec, d_theta, nbroken, damage, count_near,
                              tile_lanes(), pipe(), stream);
  return launch_pass1(bond_args(), rec, d_theta, nbroken, damage, count_near, lanes, stream);
}

BondArgs Ctx::bond_args() const {
  BondArgs a;
  a.n = N; a.row = d_row; a.col = d_col; a.mw = d_mw; a.mask = d_mask; a.nc = d_nc;
  a.near_rel = 1e-6f; a.near_ctr = d_near;
  a.influence = material == PDG_MATERIAL_ELASTIC_BOND_BASED ? kPairModelPmb : infl;
  a.fq = std::ldexp(1.0f, -qexp);
  a.pmb_c = pmb_c;
  a.sc_now = sc_cur > 0.0 ? (float)sc_cur : 0.0f;
  return a;
}

// The critical-stretch function at time t: linear between its points,
// constant beyond the ends; an exact point time gives its value.
double Ctx::sc_at(double t) const {
  const size_t n = sc_t.size();
  if (n == 0) return sc_value;
  if (t <= sc_t[0]) return sc_v[0];
  if (t >= sc_t[n - 1]) return sc_v[n - 1];
  const size_t k = (size_t)(std::upper_bound(sc_t.begin(), sc_t.end(), t) - sc_t.begin());
  const double t0 = sc_t[k - 1], t1 = sc_t[k];
  if (t == t0) return sc_v[k - 1];
  return sc_v[k - 1] + (sc_v[k] - sc_v[k - 1]) * ((t - t0) / (t1 - t0));
}

TileArgs Ctx::tile_args() const {
  TileArgs t;
  t.ntiles = ntiles; t.tile_start = d_tile_start; t.halo_off = d_halo_off; t.halo = d_halo;
  t.own_off = d_own_off; t.lcol = d_lcol;
  return t;
}

// Internal CSR from perm: rows sorted by internal neighbour id, src = list-order
// bond index, rev = reverse bond, mw = first mask word of each row.
int Ctx::build_csr() {
  iperm.resize(N);
  for (int k = 0; k < N; k++) iperm[perm[k]] = k;

  // Internal CSR, rows sorted by internal neighbour id; src = list-order bond index.
  row.assign(N + 1, 0);
  for (int k = 0; k < N; k++) row[k + 1] = row[k] + nlist[head[perm[k]]];
  col.resize(B);
  src.resize(B);
  std::vector<long long> bond0(N);  // first list-order bond index of each original node
  {
    long long b = 0;
    for (int p = 0; p < N; p++) { bond0[p] = b; b += nlist[head[p]]; }
  }
  std::atomic<int> dup{0}, asym{0};
  parallel_for(N, [&](long long k0, long long k1) {
    std::vector<std::pair<int, int>> tmp;
    for (int k = (int)k0; k < (int)k1; k++) {
      const int p = perm[k];
      const long long h0 = head[p];
      const int nn = nlist[h0];
      tmp.resize(nn);
      for (int q = 0; q < nn; q++) tmp[q] = {iperm[nlist[h0 + 1 + q]], (int)(bond0[p] + q)};
      std::sort(tmp.begin(), tmp.end());
      for (int q = 0; q < nn; q++) {
        if (q > 0 && tmp[q].first == tmp[q - 1].first) dup = 1;
        col[row[k] + q] = tmp[q].first;
        src[row[k] + q] = tmp[q].second;
      }
    }
  });
  if (dup) return fail(PDG_E_UNSUPPORTED, "duplicate neighbour in neighbourhood list");
  // Reverse bonds; the gather-only force needs a symmetric list.
  rev.resize(B);
  parallel_for(N, [&](long long k0, long long k1) {
    for (int k = (int)k0; k < (int)k1; k++)
      for (int b = row[k]; b < row[k + 1]; b++) {
        const int j = col[b];
        const int* s = col.data() + row[j];
        const int* e = col.data() + row[j + 1];
        const int* it = std::lower_bound(s, e, k);
        if (it == e || *it != k) { asym = 1; rev[b] = b; continue; }
        rev[b] = (int)(it - col.data());
      }
  });
  if (asym) return fail(PDG_E_UNSUPPORTED, "neighbourhood list is not symmetric");
  mw.assign(N + 1, 0);
  for (int k = 0; k < N; k++) mw[k + 1] = mw[k] + (row[k + 1] - row[k] + 63) / 64;
  return 0;
}

// Greedy tile boundaries along the Hilbert order (positions in perm,
// neighbours by reference id): extend a tile node by node while its halo (own
// nodes plus all their neighbours) stays within kTileHaloMax and it holds at
// most tile_nodes nodes. Returns 1 (no tiles) when a node has more than 128
// neighbours or one node's halo alone does not fit.
int Ctx::tile_bounds(std::vector<int>& bounds) const {
  bounds.assign(1, 0);
  for (int p = 0; p < N; p++)
    if (nlist[head[p]] > 128) { bounds.assign(1, 0); return 1; }
  std::vector<int> mark(N, -1);
  int n0 = 0, tid = 0;
  while (n0 < N) {
    int i = n0, hs = 0;
    while (i < N && i - n0 < tile_nodes) {
      const int p = perm[i];
      const long long h0 = head[p];
      const int nn = nlist[h0];
      int add = mark[p] != tid ? 1 : 0;
      for (int q = 0; q < nn; q++) add += mark[nlist[h0 + 1 + q]] != tid ? 1 : 0;
      if (hs + add > kTileHaloMax) {
        if (i == n0) { bounds.assign(1, 0); return 1; }
        break;
      }
      mark[p] = tid;
      for (int q = 0; q < nn; q++) mark[nlist[h0 + 1 + q]] = tid;
      hs += add;
      i++;
    }
    bounds.push_back(i);
    n0 = i;
    tid++;
  }
  return 0;
}

// Tile arrays over the internal CSR for the node ranges in `bounds`: each
// tile's halo (own nodes plus all their neighbours, sorted), the offset of its
// own nodes in it, and every bond's byte offset into the LDS arrays. An empty
// `bounds` (one entry) leaves the context untiled. Returns 1 if a halo does
// not fit.
int Ctx::build_tiles(const std::vector<int>& bounds) {
  ntiles = 0;
  tile_start.assign(1, 0);
  halo_off.assign(1, 0);
  halo.clear();
  own_off.clear();
  lcol.assign(B, 0);
  if (bounds.size() < 2) return 1;
  const int nt = (int)bounds.size() - 1;
  std::vector<std::vector<int>> Hs(nt);
  std::vector<int> own(nt);
  std::atomic<int> over{0};
  parallel_for(nt, [&](long long t0, long long t1) {
    for (int t = (int)t0; t < (int)t1; t++) {
      const int n0 = bounds[t], n1 = bounds[t + 1];
      std::vector<int>& H = Hs[t];
      H.clear();
      for (int i = n0; i < n1; i++) {
        H.push_back(i);
        H.insert(H.end(), col.begin() + row[i], col.begin() + row[i + 1]);
      }
      std::sort(H.begin(), H.end());
      H.erase(std::unique(H.begin(), H.end()), H.end());
      if ((int)H.size() > kTileHaloMax) { over = 1; continue; }
      // the kernels take a wave's trip count from its first node
      for (int i = n0 + 1; i < n1; i++)
        if (row[i + 1] - row[i] > row[i] - row[i - 1]) { over = 1; break; }
      own[t] = (int)(std::lower_bound(H.begin(), H.end(), n0) - H.begin());
      for (int k = n0; k < n1; k++)
        for (int b = row[k]; b < row[k + 1]; b++)
          lcol[b] = (unsigned short)(kLdsSlot * (std::lower_bound(H.begin(), H.end(), col[b]) - H.begin()));
    }
  }, 32);
  if (over) return 1;
  for (int t = 0; t < nt; t++) {
    own_off.push_back(own[t]);
    halo.insert(halo.end(), Hs[t].begin(), Hs[t].end());
    halo_off.push_back((int)halo.size());
    tile_start.push_back(bounds[t + 1]);
  }
  ntiles = nt;
  return 0;
}

// Bond order inside the rows of the tiled CSR for conflict-free LDS reads
// (pdg_sched.hpp), for the current lanes per node; col, src and lcol are
// permuted together and the reverse-bond index rebuilt.
void Ctx::order_bonds() {
  std::vector<int> slot(B), order(B);
  for (long long b = 0; b < B; b++) slot[b] = lcol[b] / kLdsSlot;
  std::vector<SchedStats> part(ntiles);
  const int G = tile_lanes();
  parallel_for(ntiles, [&](long long t0, long long t1) {
    for (long long t = t0; t < t1; t++)
      part[t] = schedule_bonds(G, kTileThreads, 128, tile_start, row, slot, order, t, t + 1);
  }, 16);
  sched = SchedStats{};
  for (const SchedStats& p : part) {
    sched.groups += p.groups;
    sched.cost_before += p.cost_before;
    sched.cost_after += p.cost_after;
  }
  {
    std::vector<int> c2(B), s2(B);
    std::vector<unsigned short> l2(B);
    parallel_for(B, [&](long long b0, long long b1) {
      for (long long b = b0; b < b1; b++) {
        c2[b] = col[order[b]];
        s2[b] = src[order[b]];
        l2[b] = lcol[order[b]];
      }
    });
    col.swap(c2);
    src.swap(s2);
    lcol.swap(l2);
  }
  // reverse bonds: rows are no longer sorted, so look up through a sorted
  // copy of each row
  std::vector<std::pair<int, int>> srt(B);
  parallel_for(N, [&](long long k0, long long k1) {
    for (int k = (int)k0; k < (int)k1; k++) {
      for (int b = row[k]; b < row[k + 1]; b++) srt[b] = {col[b], b};
      std::sort(srt.begin() + row[k], srt.begin() + row[k + 1]);
    }
  });
  parallel_for(N, [&](long long k0, long long k1) {
    for (int k = (int)k0; k < (int)k1; k++)
      for (int b = row[k]; b < row[k + 1]; b++) {
        const int j = col[b];
        auto it = std::lower_bound(srt.begin() + row[j], srt.begin() + row[j + 1],
                                   std::make_pair(k, -1));
        rev[b] = it->second;
      }
  });
}

// Pinned host buffer of at least `bytes` (hipHostMalloc), kept for later
// calls; nullptr if it cannot be allocated (callers then copy from the
// caller's memory directly).
char* Ctx::stage(size_t bytes) {
  if (bytes <= h_stage_bytes) return h_stage;
  // a copy queued by an earlier call (one that failed before its own
  // synchronisation) may still read the old buffer
  (void)hipStreamSynchronize(stream);
  if (h_stage) (void)hipHostFree(h_stage);
  h_stage = nullptr;
  h_stage_bytes = 0;
  if (hipHostMalloc((void**)&h_stage, bytes, hipHostMallocDefault) != hipSuccess) {
    h_stage = nullptr;
    return nullptr;
  }
  h_stage_bytes = bytes;
  return h_stage;
}

// Current positions y (reference order, host) -> the records' u = y - x,
// through the pinned staging buffer when it is available. The stream is
// synchronised first: a call that failed after its upload (a rejected bond
// field) returns without waiting, and its copy may still read the buffer.
// Such a call leaves u set to its y.
int Ctx::upload_y(const double* y) {
  if (N <= 0) return PDG_OK;
  const size_t vbytes = 24 * (size_t)N;
  int rc = hip(hipStreamSynchronize(stream), "sync");
  if (rc) return rc;
  char* st = stage(vbytes + 4 * (size_t)N);
  if (st) {
    parallel_for(3 * (long long)N, [&](long long q0, long long q1) {
      std::memcpy(st + 8 * q0, y + q0, 8 * (size_t)(q1 - q0));
    });
    rc = hip(hipMemcpyAsync(d_y, st, vbytes, hipMemcpyHostToDevice, stream), "h2d y");
  } else {
    rc = hip(hipMemcpyAsync(d_y, y, vbytes, hipMemcpyHostToDevice, stream), "h2d y");
  }
  if (!rc)
    hipLaunchKernelGGL(k_set_u, dim3((3 * N + 255) / 256), dim3(256), 0, stream, N, d_y, d_perm,
                       d_x, d_rec[cur], q);
  return rc;
}

// Host list (reference order) -> device mask words (internal order).
int Ctx::upload_mask(const double* bd, bool check_symmetric) {
  if (!bd)  // no field: every bond intact
    return hip(hipMemsetAsync(d_mask, 0, (size_t)mw.back() * 8, stream), "zero mask");
  std::vector<unsigned long long> words(mw.back(), 0ull);
  std::atomic<int> bad_value{0}, asym{0};
  // each node writes only its own words; a node's bonds are one contiguous
  // run of the list-order field
  parallel_for(N, [&](long long i0, long long i1) {
    for (int i = (int)i0; i < (int)i1; i++) {
      const int rs = row[i], deg = row[i + 1] - rs;
      for (int k = 0; k < deg; k++) {
        const double d = bd[src[rs + k]];
        if (d != 0.0 && d != 1.0) bad_value = 1;
        if (d == 1.0) words[mw[i] + (k >> 6)] |= 1ull << (k & 63);
      }
    }
  });
  if (check_symmetric && !bad_value) {
    // d_ij == d_ji, checked on the bit words (a few MB) rather than on the
    // list-order field
    auto bit = [&](int node, int k) {
      return (words[mw[node] + (k >> 6)] >> (k & 63)) & 1ull;
    };
    parallel_for(N, [&](long long i0, long long i1) {
      for (int i = (int)i0; i < (int)i1; i++)
        for (int b = row[i]; b < row[i + 1]; b++) {
          const int j = col[b];
          if (bit(i, b - row[i]) != bit(j, rev[b] - row[j])) { asym = 1; return; }
        }
    });
  }
  if (bad_value) return fail(PDG_E_UNSUPPORTED, "bond damage values must be 0 or 1");
  if (asym) return fail(PDG_E_UNSUPPORTED, "bond damage must be symmetric (d_ij == d_ji)");
  int rc = hip(hipMemcpyAsync(d_mask, words.data(), words.size() * 8, hipMemcpyHostToDevice, stream),
               "upload mask");
  if (!rc) rc = hip(hipStreamSynchronize(stream), "sync");
  return rc;
}

int Ctx::download_mask(double* bd) {
  std::vector<unsigned long long> words(mw.back());
  int rc = hip(hipMemcpyAsync(words.data(), d_mask, words.size() * 8, hipMemcpyDeviceToHost, stream),
               "download mask");
  if (rc) return rc;
  rc = hip(hipStreamSynchronize(stream), "sync");
  if (rc) return rc;
  parallel_for(N, [&](long long i0, long long i1) {
    for (int i = (int)i0; i < (int)i1; i++) {
      const int rs = row[i], deg = row[i + 1] - rs;
      for (int k = 0; k < deg; k++)
        bd[src[rs + k]] = ((words[mw[i] + (k >> 6)] >> (k & 63)) & 1ull) ? 1.0 : 0.0;
    }
  });
  return PDG_OK;
}

// internal-order 3-vectors (device) -> reference order (host)
int Ctx::download_vec3(const double* dsrc, double* out) {
  std::vector<double> tmp(3 * (size_t)N);
  int rc = hip(hipMemcpyAsync(tmp.data(), dsrc, tmp.size() * 8, hipMemcpyDeviceToHost, stream), "d2h");
  if (rc) return rc;
  rc = hip(hipStreamSynchronize(stream), "sync");
  if (rc) return rc;
  parallel_for(N, [&](long long k0, long long k1) {
    for (int k = (int)k0; k < (int)k1; k++)
      for (int c = 0; c < 3; c++) out[3 * (size_t)perm[k] + c] = tmp[3 * (size_t)k + c];
  });
  return PDG_OK;
}

int Ctx::upload_vec3(const double* in, double* ddst) {
  std::vector<double> tmp(3 * (size_t)N);
  parallel_for(N, [&](long long k0, long long k1) {
    for (int k = (int)k0; k < (int)k1; k++)
      for (int c = 0; c < 3; c++) tmp[3 * (size_t)k + c] = in ? in[3 * (size_t)perm[k] + c] : 0.0;
  });
  int rc = hip(hipMemcpyAsync(ddst, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, stream), "h2d");
  if (rc) return rc;
  return hip(hipStreamSynchronize(stream), "sync");
}

}  // namespace pdg

extern "C" {

int pdg_ctx_create(int device, pdg_ctx** out) {
  if (!out) return PDG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PDG_E_HIP;
  Ctx* c = new Ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return PDG_E_HIP;
  }
  for (int i = 0; i < 4; i++) (void)hipEventCreate(&c->ev[i]);
  // this device's address of the allocation-free entry point's scratch array
  if (dt_scratch_init(device) != hipSuccess) {
    for (int i = 0; i < 4; i++) if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PDG_E_HIP;
  }
  *out = reinterpret_cast<pdg_ctx*>(c);
  return PDG_OK;
}

int pdg_ctx_destroy(pdg_ctx* h) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->free_device();
  c->free_extra();
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  for (int i = 0; i < 4; i++) if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return PDG_OK;
}

const char* pdg_last_error(const pdg_ctx* h) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  return c ? c->err.c_str() : "null context";
}

int pdg_set_lanes_per_node(pdg_ctx* h, int lanes) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
    return c->fail(PDG_E_ARG, "lanes per node must be 4, 8, 16, 32 or 64");
  c->lanes = lanes;
  return PDG_OK;
}

int pdg_set_tile_nodes(pdg_ctx* h, int nodes) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (nodes != 0 && nodes != 32 && nodes != 64 && nodes != 128)
    return c->fail(PDG_E_ARG, "tile size must be 0 (chosen from the mesh), 32, 64 or 128 nodes");
  c->tile_nodes_req = nodes;
  return PDG_OK;
}

int pdg_tile_config(const pdg_ctx* h, int* tileNodes, int* lanes) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (tileNodes) *tileNodes = c->tiled() ? c->tile_nodes : 0;
  if (lanes) *lanes = c->tiled() ? c->tile_lanes() : std::max(c->lanes, 16);
  return PDG_OK;
}

int pdg_lds_schedule(const pdg_ctx* h, long long* groups, long long* before, long long* after) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (groups) *groups = c->sched.groups;
  if (before) *before = c->sched.cost_before;
  if (after) *after = c->sched.cost_after;
  return PDG_OK;
}

int pdg_set_kernel_variant(pdg_ctx* h, int variant) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (variant < 0 || variant > 2)
    return c->fail(PDG_E_ARG, "variant must be 0 (tiled), 1 (untiled) or 2 (tiled, plain loops)");
  c->variant = variant;
  return PDG_OK;
}

int pdg_get_tiling(const pdg_ctx* h, int* ntiles, long long* haloTotal) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !c->have_consts) return PDG_E_ARG;
  if (ntiles) *ntiles = c->ntiles;
  if (haloTotal) *haloTotal = (long long)c->halo.size();
  return PDG_OK;
}

int pdg_set_topology(pdg_ctx* h, int numOwned, const int* ownedIDs, const int* nlist,
                     long long listSize) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (numOwned < 0 || (numOwned > 0 && (!nlist || listSize < numOwned)))
    return c->fail(PDG_E_ARG, "invalid neighbourhood list");
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->free_extra();  // quasi-static and network state are sized by the old topology
  c->have_topo = false;
  c->have_consts = false;
  c->N = numOwned;
  if (ownedIDs)
    for (int i = 0; i < numOwned; i++)
      if (ownedIDs[i] != i)
        return c->fail(PDG_E_UNSUPPORTED, "ownedIDs must be 0..numOwnedPoints-1 (single device)");
  c->nlist.assign(nlist, nlist + listSize);
  c->head.assign(numOwned, 0);
  long long li = 0, B = 0;
  for (int i = 0; i < numOwned; i++) {
    if (li >= listSize) return c->fail(PDG_E_ARG, "neighbourhood list too short");
    c->head[i] = li;
    const int nn = nlist[li];
    if (nn < 0 || li + 1 + nn > listSize) return c->fail(PDG_E_ARG, "neighbourhood list truncated");
    for (int k = 0; k < nn; k++) {
      const int j = nlist[li + 1 + k];
      if (j < 0 || j >= numOwned)
        return c->fail(PDG_E_UNSUPPORTED, "neighbour id outside the owned points (ghosts unsupported)");
      if (j == i) return c->fail(PDG_E_UNSUPPORTED, "self bond in neighbourhood list");
    }
    li += nn + 1;
    B += nn;
  }
  if (B >= (1ll << 31)) return c->fail(PDG_E_UNSUPPORTED, "more than 2^31-1 bonds");
  c->B = B;
  c->have_topo = true;
  return PDG_OK;
}

int pdg_get_sizes(const pdg_ctx* h, int* n, long long* b) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !c->have_topo) return PDG_E_ARG;
  if (n) *n = c->N;
  if (b) *b = c->B;
  return PDG_OK;
}

int pdg_set_constants(pdg_ctx* h, const double* x, const double* vol, double K, double mu,
                      double horizon, int infl, double sc) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (!c->have_topo) return c->fail(PDG_E_ARG, "pdg_set_topology must be called first");
  // an empty block (no owned points) may pass null arrays
  if ((c->N > 0 && (!x || !vol)) || !(horizon > 0.0) || infl < 0 || infl > 2)
    return c->fail(PDG_E_ARG, "invalid constants");
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->free_extra();  // quasi-static and network state hold the old constants
  if (hipSetDevice(c->device) != hipSuccess) return c->fail(PDG_E_HIP, "hipSetDevice");
  const int N = c->N;
  c->K = K; c->mu = mu; c->horizon = horizon; c->infl = infl; c->sc = sc;
  c->x.assign(x, x + 3 * (size_t)N);
  c->vol.assign(vol, vol + N);

  // Hilbert-curve order of the nodes (cells of horizon/2), ties by original id.
  double lo[3] = {1e300, 1e300, 1e300};
  for (int i = 0; i < N; i++)
    for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], x[3 * (size_t)i + d]);
  const double cell = horizon * 0.5;
  std::vector<uint64_t> key(N);
  for (int i = 0; i < N; i++) {
    uint64_t q[3];
    for (int d = 0; d < 3; d++) {
      double t = std::floor((x[3 * (size_t)i + d] - lo[d]) / cell);
      q[d] = (uint64_t)std::min(std::max(t, 0.0), 2097151.0);
    }
    key[i] = hilbert3(q[0], q[1], q[2]);
  }
  c->perm.resize(N);
  std::iota(c->perm.begin(), c->perm.end(), 0);
  std::sort(c->perm.begin(), c->perm.end(),
            [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
  // Tiles along the curve; inside each tile, nodes by descending degree
  // (ties in curve order): a wave's node groups then share one trip count,
  // which cuts the lanes idle in the bond loops (plate: 84% -> 94% busy). The
  // tiles keep their node sets, so their halos are unchanged.
  // Tile size: 128 nodes, halved (down to 32) while the mesh gives fewer
  // than two workgroups per CU, so small meshes still fill the GPU.
  if (c->tile_nodes_req) {
    c->tile_nodes = c->tile_nodes_req;
  } else {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        ncu <= 0)
      ncu = 256;
    int t = kTileNodesMax;
    while (t > 32 && (N + t - 1) / t < 2 * ncu) t /= 2;
    c->tile_nodes = t;
  }
  std::vector<int> bounds;
  if (c->tile_bounds(bounds) == 0) {
    for (size_t t = 0; t + 1 < bounds.size(); t++)
      std::stable_sort(c->perm.begin() + bounds[t], c->perm.begin() + bounds[t + 1],
                       [&](int a, int b) { return c->nlist[c->head[a]] > c->nlist[c->head[b]]; });
  }
  int rc0 = c->build_csr();
  if (rc0) return rc0;
  // an untileable mesh (bounds = {0}) leaves the context untiled
  if (c->build_tiles(bounds) != 0 && bounds.size() > 1)
    return c->fail(PDG_E_UNSUPPORTED, "a tile halo outgrew its capacity under the in-tile node order");
  if (c->ntiles > 0) c->order_bonds();

  // Scales: L and S powers of two near the horizon and the mean volume.
  double vsum = 0.0;
  for (int i = 0; i < N; i++) vsum += vol[i];
  const double vmean = N ? vsum / N : 1.0;
  c->L = std::ldexp(1.0, (int)std::lround(std::log2(horizon)));
  c->S = std::ldexp(1.0, vmean > 0 ? (int)std::lround(std::log2(vmean)) : 0);
  // Position quantum q = L * 2^-qexp: the longest bond component stays below
  // 2^27 q, so bond differences of the 32-bit fixed-point positions are exact
  // integers (rounded once to FP32), rounding the positions to q moves them by
  // at most 4e-9 of the longest bond component, and the bond kernels' product
  // |xi|^2 |Y|^2 (below 2^110 (1 + s)^2 in units of q^4) stays finite in FP32
  // for every stretch s below 300.
  {
    std::vector<double> mx(std::max(1, N), 0.0);
    parallel_for(N, [&](long long i0, long long i1) {
      for (int i = (int)i0; i < (int)i1; i++) {
        double m = 0.0;
        for (int b = c->row[i]; b < c->row[i + 1]; b++) {
          const int j = c->perm[c->col[b]], pi = c->perm[i];
          for (int d = 0; d < 3; d++)
            m = std::max(m, std::fabs(x[3 * (size_t)j + d] - x[3 * (size_t)pi + d]));
        }
        mx[i] = m;
      }
    });
    const double maxlen = *std::max_element(mx.begin(), mx.end()) / c->L;
    c->bond_cmax = maxlen * c->L;
    int e = 0;
    if (maxlen > 0.0) (void)std::frexp(maxlen, &e);  // maxlen < 2^e
    c->qexp = 27 - e;
    c->q = std::ldexp(c->L, -c->qexp);
    double xmax = 0.0;
    for (size_t t = 0; t < 3 * (size_t)N; t++) xmax = std::max(xmax, std::fabs(x[t]));
    if (!(xmax / c->q < 4.0e18) || c->qexp > 120 || c->qexp < -120)
      return c->fail(PDG_E_UNSUPPORTED, "coordinates too large relative to the bond lengths "
                                        "for 64-bit fixed point");
  }

  c->free_device();
  auto A = [&](void** p, size_t bytes, const char* what) {
    return c->hip(hipMalloc(p, std::max<size_t>(bytes, 64)), what);
  };
  int rc = 0;
  if ((rc = A((void**)&c->d_row, (N + 1) * 4, "alloc row"))) return rc;
  if ((rc = A((void**)&c->d_col, c->B * 4, "alloc col"))) return rc;
  if ((rc = A((void**)&c->d_mw, (N + 1) * 4, "alloc mw"))) return rc;
  // mask words and local indices carry kMaskPad / kLcolPad zeroed entries so
  // the bond kernels may load past a row's end without a branch
  if ((rc = A((void**)&c->d_mask, ((size_t)c->mw.back() + kMaskPad) * 8, "alloc mask"))) return rc;
  if ((rc = c->hip(hipMemsetAsync(c->d_mask, 0, ((size_t)c->mw.back() + kMaskPad) * 8, c->stream), "zero mask"))) return rc;
  if ((rc = A((void**)&c->d_nc, (size_t)N * sizeof(NodeConst), "alloc nc"))) return rc;
  if ((rc = A((void**)&c->d_rec[0], (size_t)N * sizeof(NodeRec), "alloc rec0"))) return rc;
  if ((rc = A((void**)&c->d_rec[1], (size_t)N * sizeof(NodeRec), "alloc rec1"))) return rc;
  if ((rc = A((void**)&c->d_theta, (size_t)N * 4, "alloc theta"))) return rc;
  if ((rc = A((void**)&c->d_nbroken, (size_t)N * 4, "alloc nbroken"))) return rc;
  if ((rc = A((void**)&c->d_v, (size_t)N * 24, "alloc v"))) return rc;
  if ((rc = A((void**)&c->d_a, (size_t)N * 24, "alloc a"))) return rc;
  if ((rc = A((void**)&c->d_f, (size_t)N * 24, "alloc f"))) return rc;
  if ((rc = A((void**)&c->d_fout, (size_t)N * 24, "alloc fout"))) return rc;
  if ((rc = A((void**)&c->d_y, (size_t)N * 24, "alloc y"))) return rc;
  if ((rc = A((void**)&c->d_inv_rho, (size_t)N * 8, "alloc rho"))) return rc;
  if ((rc = A((void**)&c->d_bcslot, (size_t)N * 4, "alloc bcslot"))) return rc;
  if ((rc = A((void**)&c->d_bcs, sizeof(BcSlot), "alloc bcs"))) return rc;
  if ((rc = A((void**)&c->d_err, 4, "alloc err"))) return rc;
  if ((rc = A((void**)&c->d_near, 8, "alloc near"))) return rc;
  if ((rc = A((void**)&c->d_perm, (size_t)N * 4, "alloc perm"))) return rc;
  if ((rc = A((void**)&c->d_vol, (size_t)N * 8, "alloc vol"))) return rc;
  if ((rc = A((void**)&c->d_x, (size_t)N * 24, "alloc x"))) return rc;
  if (c->ntiles > 0) {
    if ((rc = A((void**)&c->d_tile_start, 4 * c->tile_start.size(), "alloc tiles"))) return rc;
    if ((rc = A((void**)&c->d_halo_off, 4 * c->halo_off.size(), "alloc halo off"))) return rc;
    if ((rc = A((void**)&c->d_halo, 4 * c->halo.size(), "alloc halo"))) return rc;
    if ((rc = A((void**)&c->d_own_off, 4 * c->own_off.size(), "alloc own off"))) return rc;
    if ((rc = A((void**)&c->d_lcol, 2 * ((size_t)c->B + kLcolPad), "alloc lcol"))) return rc;
    if ((rc = c->hip(hipMemsetAsync(c->d_lcol, 0, 2 * ((size_t)c->B + kLcolPad), c->stream), "zero lcol"))) return rc;
  }

  hipStream_t s = c->stream;
  // Weighted volume on the device in the reference's summation order.
  {
    double *dx = nullptr, *dv = nullptr, *dm = nullptr;
    int* dl = nullptr;
    long long* dh = nullptr;
    rc = A((void**)&dx, 24 * (size_t)N, "alloc x");
    if (!rc) rc = A((void**)&dv, 8 * (size_t)N, "alloc vol");
    if (!rc) rc = A((void**)&dm, 8 * (size_t)N, "alloc m");
    if (!rc) rc = A((void**)&dl, 4 * c->nlist.size(), "alloc list");
    if (!rc) rc = A((void**)&dh, 8 * (size_t)N, "alloc head");
    if (!rc) rc = c->hip(hipMemcpyAsync(dx, x, 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d x");
    if (!rc) rc = c->hip(hipMemcpyAsync(dv, vol, 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d vol");
    if (!rc) rc = c->hip(hipMemcpyAsync(dl, c->nlist.data(), 4 * c->nlist.size(), hipMemcpyHostToDevice, s), "h2d list");
    if (!rc) rc = c->hip(hipMemcpyAsync(dh, c->head.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d head");
    if (!rc) rc = c->hip(launch_weighted_volume(N, dx, dv, dl, dh, horizon, infl, dm, s), "weighted volume");
    c->m.assign(N, 0.0);
    if (!rc) rc = c->hip(hipMemcpyAsync(c->m.data(), dm, 8 * (size_t)N, hipMemcpyDeviceToHost, s), "d2h m");
    if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
    (void)hipFree(dx); (void)hipFree(dv); (void)hipFree(dm); (void)hipFree(dl); (void)hipFree(dh);
    if (rc) return rc;
  }

  // Node records and constants (internal order).
  std::vector<NodeRec> rec(N);
  std::vector<NodeConst> nc(N);
  const double L = c->L, S = c->S, invq = 1.0 / c->q;
  const double s2 = std::ldexp(1.0, -2 * c->qexp);  // sums in q units -> L units
  std::vector<double> x_int(3 * (size_t)N);
  for (int k = 0; k < N; k++) {
    const int p = c->perm[k];
    const double m = c->m[p];
    NodeRec& r = rec[k];
    for (int d = 0; d < 3; d++) {
      const double xd = x[3 * (size_t)p + d];
      x_int[3 * (size_t)k + d] = xd;
      r.xq[d] = (unsigned)(unsigned long long)std::llrint(xd * invq);
      r.u[d] = 0.0;
    }
    r.vol = (float)(vol[p] / S);
    r.p = 0.0f;
    for (float& z : r.pad) z = 0.0f;
    NodeConst& q = nc[k];
    q.sc = sc > 0.0 ? (float)sc : 0.0f;
    q.hhat = (float)(horizon * invq);
    if (c->material == PDG_MATERIAL_ELASTIC_BOND_BASED) {
      // no dilatation term: theta and p stay zero
      r.alpha = 0.0f; q.ctheta = 0.0f; q.kp = 0.0f;
    } else if (m > 0.0) {
      const double alpha = 15.0 * mu / m;
      const double beta = 3.0 * K / m - alpha / 3.0;
      const double ct = 3.0 * L * L * S / m;
      r.alpha = (float)(alpha * L * S);
      q.ctheta = (float)(ct * s2);
      q.kp = (float)(ct * beta * L * S * s2);
    } else {
      r.alpha = 0.0f; q.ctheta = 0.0f; q.kp = 0.0f;
    }
  }
  // bond-based pair constant (src/materials/elastic_bond_based.cxx:80-84) in
  // the kernels' units: volumes scaled by S, lengths by q
  c->pmb_c = 0.0f;
  if (c->material == PDG_MATERIAL_ELASTIC_BOND_BASED) {
    const double pi = 3.14159265358979323846;
    const double cc = c->planar ? 12.0 * K / (pi * c->height * horizon * horizon * horizon)
                                : 18.0 * K / (pi * horizon * horizon * horizon * horizon);
    c->pmb_c = (float)std::ldexp(cc * S, c->qexp);
  }
  std::vector<int> noslot(N, -1);
  std::vector<double> rho1(N, 1.0);
  std::vector<double> vol_int(N);
  for (int k = 0; k < N; k++) vol_int[k] = vol[c->perm[k]];
  rc = c->hip(hipMemcpyAsync(c->d_row, c->row.data(), 4 * (N + 1), hipMemcpyHostToDevice, s), "h2d row");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_col, c->col.data(), 4 * c->B, hipMemcpyHostToDevice, s), "h2d col");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_mw, c->mw.data(), 4 * (N + 1), hipMemcpyHostToDevice, s), "h2d mw");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_mask, 0, (size_t)c->mw.back() * 8, s), "zero mask");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_nc, nc.data(), N * sizeof(NodeConst), hipMemcpyHostToDevice, s), "h2d nc");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_rec[0], rec.data(), N * sizeof(NodeRec), hipMemcpyHostToDevice, s), "h2d rec0");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_rec[1], rec.data(), N * sizeof(NodeRec), hipMemcpyHostToDevice, s), "h2d rec1");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_v, 0, 24 * (size_t)N, s), "zero v");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_a, 0, 24 * (size_t)N, s), "zero a");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_f, 0, 24 * (size_t)N, s), "zero f");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_theta, 0, 4 * (size_t)N, s), "zero theta");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_nbroken, 0, 4 * (size_t)N, s), "zero nbroken");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_bcslot, noslot.data(), 4 * (size_t)N, hipMemcpyHostToDevice, s), "h2d bcslot");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_inv_rho, rho1.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d rho");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_perm, c->perm.data(), 4 * (size_t)N, hipMemcpyHostToDevice, s), "h2d perm");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_vol, vol_int.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d vol");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_x, x_int.data(), 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d x");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_err, 0, 4, s), "zero err");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_near, 0, 8, s), "zero near");
  if (!rc && c->ntiles > 0) {
    rc = c->hip(hipMemcpyAsync(c->d_tile_start, c->tile_start.data(), 4 * c->tile_start.size(), hipMemcpyHostToDevice, s), "h2d tiles");
    if (!rc) rc = c->hip(hipMemcpyAsync(c->d_halo_off, c->halo_off.data(), 4 * c->halo_off.size(), hipMemcpyHostToDevice, s), "h2d halo off");
    if (!rc) rc = c->hip(hipMemcpyAsync(c->d_halo, c->halo.data(), 4 * c->halo.size(), hipMemcpyHostToDevice, s), "h2d halo");
    if (!rc) rc = c->hip(hipMemcpyAsync(c->d_own_off, c->own_off.data(), 4 * c->own_off.size(), hipMemcpyHostToDevice, s), "h2d own off");
    if (!rc) rc = c->hip(hipMemcpyAsync(c->d_lcol, c->lcol.data(), 2 * (size_t)c->B, hipMemcpyHostToDevice, s), "h2d lcol");
  }
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  c->cur = 0;
  c->have_consts = true;
  return PDG_OK;
}

int pdg_set_critical_stretch(pdg_ctx* h, const double* sc) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || (c->N > 0 && !sc)) return PDG_E_ARG;
  if (!(c->sc > 0.0)) return c->fail(PDG_E_ARG, "damage model disabled (critical stretch <= 0 at set-up)");
  std::vector<NodeConst> nc(c->N);
  int rc = c->hip(hipMemcpyAsync(nc.data(), c->d_nc, c->N * sizeof(NodeConst), hipMemcpyDeviceToHost, c->stream), "d2h nc");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  for (int k = 0; k < c->N; k++) {
    const double v = sc[c->perm[k]];
    if (!(v > 0.0)) return c->fail(PDG_E_ARG, "critical stretch must be positive");
    nc[k].sc = (float)v;
  }
  rc = c->hip(hipMemcpyAsync(c->d_nc, nc.data(), c->N * sizeof(NodeConst), hipMemcpyHostToDevice, c->stream), "h2d nc");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  return rc;
}

int pdg_set_material_model(pdg_ctx* h, int model, int planar, double height) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (model != PDG_MATERIAL_ELASTIC && model != PDG_MATERIAL_ELASTIC_BOND_BASED &&
      model != PDG_MATERIAL_LINEAR_LPS)
    return c->fail(PDG_E_ARG, "unknown material model");
  if (planar < 0 || planar > 2 || (planar && model != PDG_MATERIAL_ELASTIC_BOND_BASED))
    return c->fail(PDG_E_ARG, "plane strain/stress is a variant of the bond-based model");
  if (planar && !(height > 0.0)) return c->fail(PDG_E_ARG, "plane strain/stress needs a height");
  c->material = model;
  c->planar = planar;
  c->height = height;
  return PDG_OK;
}

int pdg_set_critical_stretch_value(pdg_ctx* h, double sc) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (!(sc <= 0.0) && !(sc > 0.0)) return c->fail(PDG_E_ARG, "critical stretch is NaN");
  c->sc_value = sc > 0.0 ? sc : 0.0;
  c->sc_cur = c->sc_value;
  return PDG_OK;
}

int pdg_set_critical_stretch_function(pdg_ctx* h, int n, const double* times, const double* values) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || n < 0 || (n > 0 && (!times || !values))) return PDG_E_ARG;
  for (int k = 0; k < n; k++) {
    if (!(values[k] > 0.0)) return c->fail(PDG_E_ARG, "critical stretch values must be positive");
    if (k > 0 && !(times[k] > times[k - 1])) return c->fail(PDG_E_ARG, "times must increase");
  }
  c->sc_t.assign(times, times + n);
  c->sc_v.assign(values, values + n);
  return PDG_OK;
}

int pdg_set_density(pdg_ctx* h, const double* rho) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || (c->N > 0 && !rho)) return PDG_E_ARG;
  std::vector<double> r(c->N);
  for (int k = 0; k < c->N; k++) {
    const double d = rho[c->perm[k]];
    if (!(d > 0.0)) return c->fail(PDG_E_ARG, "density must be positive");
    r[k] = 1.0 / d;
  }
  int rc = c->hip(hipMemcpyAsync(c->d_inv_rho, r.data(), 8 * (size_t)c->N, hipMemcpyHostToDevice, c->stream), "h2d rho");
  if (rc) return rc;
  return c->hip(hipStreamSynchronize(c->stream), "sync");
}

int pdg_initialize(pdg_ctx* h, double* m) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || (c->N > 0 && !m)) return PDG_E_ARG;
  std::copy(c->m.begin(), c->m.end(), m);
  return PDG_OK;
}

int pdg_compute_damage(pdg_ctx* h, const double* y, const double* bdN, double* bdNP1,
                       double* damage, long long* nearCritical) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || (c->N > 0 && (!y || !bdNP1 || !damage))) return PDG_E_ARG;
  if (!(c->sc > 0.0)) return c->fail(PDG_E_ARG, "no damage model (critical stretch <= 0)");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const int N = c->N;
  int rc = c->upload_y(y);
  if (!rc) rc = c->upload_mask(bdN, false);
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_near, 0, 8, s), "zero near");
  if (!rc) rc = c->hip(c->pass1(c->d_rec[c->cur], c->d_nbroken, true, nearCritical != nullptr), "pass1");
  if (!rc) rc = c->download_mask(bdNP1);
  if (rc) return rc;
  std::vector<int> nb(N);
  rc = c->hip(hipMemcpyAsync(nb.data(), c->d_nbroken, 4 * (size_t)N, hipMemcpyDeviceToHost, s), "d2h nb");
  unsigned long long near = 0;
  if (!rc && nearCritical)
    rc = c->hip(hipMemcpyAsync(&near, c->d_near, 8, hipMemcpyDeviceToHost, s), "d2h near");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  for (int k = 0; k < N; k++) {
    const int deg = c->row[k + 1] - c->row[k];
    double t = (double)nb[k];
    damage[c->perm[k]] = deg > 0 ? t / deg : 0.0;
  }
  if (nearCritical) *nearCritical = (long long)near;
  return PDG_OK;
}

int pdg_compute_force(pdg_ctx* h, double dt, const double* y, const double* bdNP1,
                      double* dilatation, double* force) {
  (void)dt;
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || (c->N > 0 && (!y || !force))) return PDG_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->material == PDG_MATERIAL_LINEAR_LPS || c->thermal()) {
    // the linearized model and thermal strains through the FP64
    // quasi-static kernels at u = y - x
    int r = c->upload_mask(bdNP1, true);
    if (r) return r;
    std::vector<double> u(3 * (size_t)c->N);
    for (size_t q = 0; q < u.size(); q++) u[q] = y[q] - c->x[q];
    return pdg_qs_force_dilatation(h, u.data(), force, dilatation);
  }
  hipStream_t s = c->stream;
  const int N = c->N;
  // y in and f, theta out go through one pinned buffer (PCIe at full rate);
  // the stream orders the copies, and the call ends synchronised
  const size_t vbytes = 24 * (size_t)N, tbytes = 4 * (size_t)N;
  int rc = c->upload_y(y);
  char* st = N > 0 ? c->stage(vbytes + tbytes) : nullptr;  // the buffer upload_y used
  if (!rc) rc = c->upload_mask(bdNP1, true);
  if (!rc) rc = c->hip(c->pass1(c->d_rec[c->cur], nullptr, false, false), "pass1");
  if (!rc) rc = c->hip(c->tiled() ? launch_pass2_tiled_force(c->bond_args(), c->tile_args(), c->d_rec[c->cur], c->d_fout, c->tile_lanes(), c->pipe(), s)
                                  : launch_pass2_force(c->bond_args(), c->d_rec[c->cur], c->d_fout, c->lanes, s),
                       "pass2");
  if (rc) return rc;
  if (!st) {
    rc = c->download_vec3(c->d_fout, force);
    if (rc) return rc;
    if (dilatation) {
      std::vector<float> th(N);
      rc = c->hip(hipMemcpyAsync(th.data(), c->d_theta, tbytes, hipMemcpyDeviceToHost, s), "d2h theta");
      if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
      if (rc) return rc;
      for (int k = 0; k < N; k++) dilatation[c->perm[k]] = (double)th[k];
    }
    return PDG_OK;
  }
  rc = c->hip(hipMemcpyAsync(st, c->d_fout, vbytes, hipMemcpyDeviceToHost, s), "d2h force");
  if (!rc && dilatation)
    rc = c->hip(hipMemcpyAsync(st + vbytes, c->d_theta, tbytes, hipMemcpyDeviceToHost, s), "d2h theta");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  const double* fs = reinterpret_cast<const double*>(st);
  const float* ts = reinterpret_cast<const float*>(st + vbytes);
  const int* perm = c->perm.data();
  parallel_for(N, [&](long long k0, long long k1) {
    for (int k = (int)k0; k < (int)k1; k++) {
      const size_t p = (size_t)perm[k];
      for (int d = 0; d < 3; d++) force[3 * p + d] = fs[3 * (size_t)k + d];
      if (dilatation) dilatation[p] = (double)ts[k];
    }
  });
  return PDG_OK;
}

int pdg_state_upload(pdg_ctx* h, const double* u, const double* v, const double* a,
                     const double* bd) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts) return PDG_E_ARG;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const int N = c->N;
  int rc = 0;
  if (u) {
    rc = c->hip(hipMemcpyAsync(c->d_y, u, 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d u");
    if (!rc && N > 0)
      hipLaunchKernelGGL(k_set_u_disp, dim3((3 * N + 255) / 256), dim3(256), 0, s, N, c->d_y,
                         c->d_perm, c->d_rec[c->cur], c->q);
  } else if (N > 0) {
    std::vector<double> z(3 * (size_t)N, 0.0);
    rc = c->hip(hipMemcpyAsync(c->d_y, z.data(), 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d u");
    if (!rc)
      hipLaunchKernelGGL(k_set_u_disp, dim3((3 * N + 255) / 256), dim3(256), 0, s, N, c->d_y,
                         c->d_perm, c->d_rec[c->cur], c->q);
    if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  }
  if (!rc) rc = c->upload_vec3(v, c->d_v);
  if (!rc) rc = c->upload_vec3(a, c->d_a);
  if (!rc) rc = c->upload_mask(bd, true);
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  return rc;
}

int pdg_set_bcs(pdg_ctx* h, const pdg_bc* bcs, int nb) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || nb < 0 || (nb > 0 && !bcs)) return PDG_E_ARG;
  std::vector<int> slot(c->N, -1);
  std::vector<BcSlot> slots;
  for (int b = 0; b < nb; b++) {
    if (bcs[b].coord < 0 || bcs[b].coord > 2 || bcs[b].numNodes < 0)
      return c->fail(PDG_E_ARG, "invalid boundary condition");
    for (int q = 0; q < bcs[b].numNodes; q++) {
      const int p = bcs[b].nodes[q];
      if (p < 0 || p >= c->N) return c->fail(PDG_E_ARG, "boundary-condition node out of range");
      const int k = c->iperm[p];
      if (slot[k] < 0) {
        slot[k] = (int)slots.size();
        slots.push_back(BcSlot{{0, 0, 0}, {0, 0, 0}, 0, 0});
      }
      BcSlot& s = slots[slot[k]];
      s.c0[bcs[b].coord] = bcs[b].c0;
      s.c1[bcs[b].coord] = bcs[b].c1;
      s.mask |= 1 << bcs[b].coord;
    }
  }
  if (slots.empty()) slots.push_back(BcSlot{{0, 0, 0}, {0, 0, 0}, 0, 0});
  if (c->d_bcs) (void)hipFree(c->d_bcs);
  c->d_bcs = nullptr;
  int rc = c->hip(hipMalloc((void**)&c->d_bcs, slots.size() * sizeof(BcSlot)), "alloc bcs");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_bcs, slots.data(), slots.size() * sizeof(BcSlot), hipMemcpyHostToDevice, c->stream), "h2d bcs");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_bcslot, slot.data(), 4 * (size_t)c->N, hipMemcpyHostToDevice, c->stream), "h2d bcslot");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  return rc;
}

int pdg_verlet_run(pdg_ctx* h, int nsteps, double dt, double t0, int initialEval) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || nsteps < 0 || !(dt > 0.0)) return PDG_E_ARG;
  if (c->material == PDG_MATERIAL_LINEAR_LPS)
    return c->fail(PDG_E_UNSUPPORTED, "explicit runs implement the elastic and bond-based models");
  if (c->thermal()) return c->fail(PDG_E_UNSUPPORTED, "explicit runs with thermal strains");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const bool damage = c->sc > 0.0;
  BondArgs ba = c->bond_args();
  VerletArgs va{};
  va.inv_rho = c->d_inv_rho; va.v = c->d_v; va.a = c->d_a; va.f = c->d_f; va.bc_slot = c->d_bcslot;
  va.bcs = c->d_bcs; va.dt = dt; va.uq = c->q; va.inv_uq = 1.0 / c->q;
  va.err = c->d_err;
  int rc = c->hip(hipMemsetAsync(c->d_err, 0, 4, s), "zero err");
  // near-critical bonds are counted over every damage evaluation of the run
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_near, 0, 8, s), "zero near");
  if (rc) return rc;
  c->t_pass1 = c->t_pass2 = 0.0;
  rc = c->hip(hipEventRecord(c->ev[0], s), "event");
  if (rc) return rc;
  auto step_time = [&](int k) { return t0 + (k * dt); };
  float ms = 0.0f;
  // Per-pass timing: three events per force evaluation, recorded on the
  // context's stream with no host synchronisation inside the run.
  const size_t nev = c->per_pass_timing ? 3 * (size_t)(nsteps + 1) : 0;
  while (c->tev.size() < nev) {
    hipEvent_t e;
    rc = c->hip(hipEventCreate(&e), "event create");
    if (rc) return rc;
    c->tev.push_back(e);
  }
  int nevals = 0;
  const bool contact = contact_on(c);
  va.fadd = contact ? contact_force_ptr(c) : nullptr;
  // pass1 + pass2 of one force evaluation, finishing step `k` (k=0: initial)
  auto eval = [&](int k, bool end_kick) -> int {
    int r = 0;
    // contact force at the evaluation's positions and half-step velocities
    if (contact) r = contact_step(c, k, dt, c->d_rec[c->cur ^ 1], c->d_rec[c->cur]);
    if (!r && nev) r = c->hip(hipEventRecord(c->tev[3 * nevals], s), "event");
    // a time-dependent critical stretch at the time this evaluation ends
    c->sc_cur = c->sc_at(step_time(k));
    if (!r) r = c->hip(c->pass1(c->d_rec[c->cur], c->d_nbroken, damage, damage), "pass1");
    if (!r && nev) r = c->hip(hipEventRecord(c->tev[3 * nevals + 1], s), "event");
    va.end_kick = end_kick ? 1 : 0;
    va.start_step = k < nsteps ? 1 : 0;
    va.t_cur = step_time(k + 1);
    va.t_prev = step_time(k);
    va.inv_dtbc = 1.0 / (va.t_cur - va.t_prev);
    if (!r) r = c->hip(c->tiled() ? launch_pass2_tiled_verlet(ba, c->tile_args(), c->d_rec[c->cur], c->d_rec[c->cur ^ 1], va, c->tile_lanes(), c->pipe(), s)
                                  : launch_pass2_verlet(ba, c->d_rec[c->cur], c->d_rec[c->cur ^ 1], va, c->lanes, s),
                       "pass2");
    if (!r && nev) r = c->hip(hipEventRecord(c->tev[3 * nevals + 2], s), "event");
    nevals++;
    c->cur ^= 1;
    return r;
  };
  if (initialEval) {
    rc = eval(0, false);
  } else if (nsteps > 0) {
    va.t_cur = step_time(1);
    va.t_prev = step_time(0);
    va.inv_dtbc = 1.0 / (va.t_cur - va.t_prev);
    rc = c->hip(launch_verlet_prologue(c->N, c->d_rec[c->cur], c->d_rec[c->cur ^ 1], va, s), "prologue");
    c->cur ^= 1;
  }
  for (int k = 1; k <= nsteps && !rc; k++) rc = eval(k, true);
  c->sc_cur = c->sc_value;
  if (rc) return rc;
  rc = c->hip(hipEventRecord(c->ev[3], s), "event");
  if (!rc) rc = c->hip(hipEventSynchronize(c->ev[3]), "sync");
  if (!rc) { (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[3]); c->t_total = ms; }
  c->n_pass = nev ? nevals : 0;
  for (int q = 0; q < (nev ? nevals : 0) && !rc; q++) {
    (void)hipEventElapsedTime(&ms, c->tev[3 * q], c->tev[3 * q + 1]);
    c->t_pass1 += ms;
    (void)hipEventElapsedTime(&ms, c->tev[3 * q + 1], c->tev[3 * q + 2]);
    c->t_pass2 += ms;
  }
  int err = 0;
  if (!rc) rc = c->hip(hipMemcpy(&err, c->d_err, 4, hipMemcpyDeviceToHost), "d2h err");
  if (rc) return rc;
  if (err) return c->fail(PDG_E_NONFINITE, "NaN returned by force evaluation");
  return PDG_OK;
}

int pdg_verlet_run_persistent(pdg_ctx* h, int nsteps, double dt, double t0, int initialEval,
                              int* gridSize) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || nsteps < 0 || !(dt > 0.0)) return PDG_E_ARG;
  if (!c->tiled()) return c->fail(PDG_E_UNSUPPORTED, "persistent run needs the tiled kernels");
  if (c->material == PDG_MATERIAL_LINEAR_LPS)
    return c->fail(PDG_E_UNSUPPORTED, "explicit runs implement the elastic and bond-based models");
  if (!c->sc_t.empty())
    return c->fail(PDG_E_UNSUPPORTED, "persistent run with a time-dependent critical stretch");
  if (contact_on(c)) return c->fail(PDG_E_UNSUPPORTED, "persistent run with contact");
  if (c->thermal()) return c->fail(PDG_E_UNSUPPORTED, "explicit runs with thermal strains");
  const int tl = c->tile_lanes();
  if (tl != 4 && tl != 8 && tl != 16)
    return c->fail(PDG_E_UNSUPPORTED, "persistent run uses 4, 8 or 16 lanes per node");
  if (!initialEval && nsteps == 0) return PDG_OK;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  VerletArgs va{};
  va.inv_rho = c->d_inv_rho; va.v = c->d_v; va.a = c->d_a; va.f = c->d_f; va.bc_slot = c->d_bcslot;
  va.bcs = c->d_bcs; va.dt = dt; va.uq = c->q; va.inv_uq = 1.0 / c->q;
  va.err = c->d_err;
  int rc = c->hip(hipMemsetAsync(c->d_err, 0, 4, s), "zero err");
  if (!rc) rc = c->hip(hipMemsetAsync(c->d_near, 0, 8, s), "zero near");
  if (rc) return rc;
  // a continuing run starts step 1 from the stored acceleration
  if (!initialEval) {
    va.t_cur = t0 + (1 * dt);
    va.t_prev = t0 + (0 * dt);
    va.inv_dtbc = 1.0 / (va.t_cur - va.t_prev);
    rc = c->hip(launch_verlet_prologue(c->N, c->d_rec[c->cur], c->d_rec[c->cur ^ 1], va, s), "prologue");
    if (rc) return rc;
    c->cur ^= 1;
  }
  NodeRec* r0 = c->d_rec[c->cur];
  NodeRec* r1 = c->d_rec[c->cur ^ 1];
  float ms = 0.0f;
  rc = c->hip(hipEventRecord(c->ev[0], s), "event");
  if (!rc) rc = c->hip(launch_persistent(c->bond_args(), c->tile_args(), r0, r1, c->d_theta,
                                         c->d_nbroken, va, nsteps, t0, initialEval, c->sc > 0.0,
                                         tl, gridSize, s), "persistent launch");
  if (!rc) rc = c->hip(hipEventRecord(c->ev[3], s), "event");
  if (!rc) rc = c->hip(hipEventSynchronize(c->ev[3]), "sync");
  if (rc) return rc;
  (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[3]);
  c->t_total = ms;
  const int evals = nsteps + (initialEval ? 1 : 0);
  if (evals & 1) c->cur ^= 1;
  int err = 0;
  rc = c->hip(hipMemcpy(&err, c->d_err, 4, hipMemcpyDeviceToHost), "d2h err");
  if (rc) return rc;
  if (err) return c->fail(PDG_E_NONFINITE, "NaN returned by force evaluation");
  return PDG_OK;
}

int pdg_state_download(pdg_ctx* h, double* u, double* v, double* a, double* bd, double* damage,
                       double* dilatation, double* force) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts) return PDG_E_ARG;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const int N = c->N;
  int rc = 0;
  if (u) {
    std::vector<NodeRec> rec(N);
    rc = c->hip(hipMemcpyAsync(rec.data(), c->d_rec[c->cur], N * sizeof(NodeRec), hipMemcpyDeviceToHost, s), "d2h rec");
    if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    for (int k = 0; k < N; k++)
      for (int d = 0; d < 3; d++) u[3 * (size_t)c->perm[k] + d] = rec[k].u[d] * c->q;
  }
  if (v && (rc = c->download_vec3(c->d_v, v))) return rc;
  if (a && (rc = c->download_vec3(c->d_a, a))) return rc;
  if (force && (rc = c->download_vec3(c->d_f, force))) return rc;
  if (bd && (rc = c->download_mask(bd))) return rc;
  if (damage || dilatation) {
    std::vector<int> nb(N);
    std::vector<float> th(N);
    rc = c->hip(hipMemcpyAsync(nb.data(), c->d_nbroken, 4 * (size_t)N, hipMemcpyDeviceToHost, s), "d2h nb");
    if (!rc) rc = c->hip(hipMemcpyAsync(th.data(), c->d_theta, 4 * (size_t)N, hipMemcpyDeviceToHost, s), "d2h th");
    if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    for (int k = 0; k < N; k++) {
      const int deg = c->row[k + 1] - c->row[k];
      if (damage) damage[c->perm[k]] = (c->sc > 0.0 && deg > 0) ? (double)nb[k] / deg : 0.0;
      if (dilatation) dilatation[c->perm[k]] = (double)th[k];
    }
  }
  return PDG_OK;
}

int pdg_last_timing(const pdg_ctx* h, double* total, double* p1, double* p2) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c) return PDG_E_ARG;
  if (total) *total = c->t_total;
  if (p1) *p1 = c->t_pass1;
  if (p2) *p2 = c->t_pass2;
  return PDG_OK;
}

int pdg_set_per_pass_timing(pdg_ctx* h, int on) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  c->per_pass_timing = on != 0;
  return PDG_OK;
}

int pdg_near_critical_count(const pdg_ctx* h, long long* count) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !count) return PDG_E_ARG;
  unsigned long long v = 0;
  if (hipMemcpy(&v, c->d_near, 8, hipMemcpyDeviceToHost) != hipSuccess) return PDG_E_HIP;
  *count = (long long)v;
  return PDG_OK;
}

}  // extern "C"
