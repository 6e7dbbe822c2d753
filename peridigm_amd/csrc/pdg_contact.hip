// Short-range contact: the contact search of the current configuration and
// the short-range force contact model.
//
// Reference: src/core/Peridigm_ContactManager.cpp:700-750 (contactSearch: a
// neighbourhood search of the current positions within the contact search
// radius, the point's bonded neighbours dropped, broken or not),
// src/contact/Peridigm_ShortRangeForceContactModel.cpp:98-258 (computeForce)
// and src/core/Peridigm.cpp:1468-1472, 1515-1524, 1559-1565 (the search at
// every step whose number is a multiple of the search frequency, before the
// first half kick; the contact force added to the internal force before the
// acceleration).
//
// Search: positions hashed into a spatial grid of cells no smaller than the
// search radius (2^k buckets, k the smallest with 2^k >= 2N), points sorted by
// bucket. A point visits each distinct bucket of its 27 surrounding cells
// once; the distance predicate alone decides membership, as in the
// reference. The counting pass tests no bonds per candidate: a point's
// contact count is its points within the radius less its bonded neighbours
// within it, both under one predicate, so the difference is exact. Only the
// points with contacts are walked again to write their rows, finding bonded
// neighbours by binary search in the point's bond row (kept sorted in a copy
// of the bond list).
//
// Force: the search list is symmetric (j lists i exactly when i lists j: the
// squared distance and the bond lists are symmetric), and the reference adds
// the same pair term to a point once from its own list and once from its
// neighbour's, so the explicit run gathers 2 x (the point's side of every
// pair) with no atomics. The drop-in call takes any list and scatters both
// sides, as the reference does, with FP64 atomic adds.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pdg_ctx.hpp"

namespace pdg {

struct ContactState {
  double R = 0.0;       // search radius
  double rc = 0.0;      // contact radius
  double ks = 0.0;      // spring constant
  double mu = 0.0;      // friction coefficient
  double c = 0.0;       // 9 ks / (pi horizon^4)
  int freq = 1;         // search frequency (steps)
  long long step = 0;   // steps since the last initial evaluation
  bool built = false;   // a search list exists
  unsigned nbits = 1;   // log2 of the bucket count
  double inv_cs = 0.0;  // 1 / cell size
  // device, internal order unless stated
  double* y = nullptr;            // 3n positions of the search
  double* fc = nullptr;           // 3n contact force density of the last evaluation
  int* bcol = nullptr;            // bond list, each row ascending (rows: Ctx::d_row)
  int *key = nullptr, *idx = nullptr, *skey = nullptr, *sidx = nullptr;
  int *bstart = nullptr, *bend = nullptr;
  long long *cnt = nullptr, *off = nullptr;  // n+1 each (cnt[n] = 0)
  int* list = nullptr;
  long long cap = 0, size = 0;
  int* act = nullptr;   // n: points with contacts
  int* nact_d = nullptr;
  int nact = 0;
  double* ys = nullptr;  // 3n positions in bucket order (the counting pass)
  int* over = nullptr;   // n: points with contacts, in any order (the filling pass)
  int* nover_d = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  // drop-in call buffers (reference order)
  double *vref = nullptr, *dv = nullptr, *dfc = nullptr;
  long long* dhead = nullptr;
  int* dlist = nullptr;
  long long dcap = 0;
};

namespace {

struct PairParams {
  double rc, rc2, c, horizon, mu;
};

// Cell coordinate of one position component; huge or non-finite values are
// clamped (clamping keeps cells of points within one cell size of each other
// adjacent).
__device__ __forceinline__ int cell_of(double p, double inv_cs) {
  double c = floor(p * inv_cs);
  c = fmin(fmax(c, -1073741824.0), 1073741824.0);
  return (int)c;
}

__device__ __forceinline__ unsigned bucket_of(int a, int b, int c, unsigned mask) {
  return ((unsigned)a * 73856093u ^ (unsigned)b * 19349663u ^ (unsigned)c * 83492791u) & mask;
}

__global__ void k_contact_y(int n, const NodeRec* __restrict__ rec, const double* __restrict__ x,
                            double q, double* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int k = t / 3, l = t - 3 * k;
  y[t] = x[t] + rec[k].u[l] * q;
}

__global__ void k_contact_y_perm(int n, const double* __restrict__ yref, const int* __restrict__ perm,
                                 double* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int k = t / 3, l = t - 3 * k;
  y[t] = yref[3 * (size_t)perm[k] + l];
}

__global__ void k_contact_keys(int n, const double* __restrict__ y, double inv_cs, unsigned mask,
                               int* __restrict__ key, int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = cell_of(y[3 * i], inv_cs), b = cell_of(y[3 * i + 1], inv_cs),
            c = cell_of(y[3 * i + 2], inv_cs);
  key[i] = (int)bucket_of(a, b, c, mask);
  idx[i] = i;
}

__global__ void k_fill_int(long long n, int* __restrict__ p, int v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ void k_fill_f64(long long n, double* __restrict__ p) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// Bucket-ordered copy of the positions: a bucket's candidates are then read
// contiguously.
__global__ void k_contact_ys(int n, const int* __restrict__ sidx, const double* __restrict__ y,
                             double* __restrict__ ys) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int q = t / 3, l = t - 3 * q;
  ys[t] = y[3 * (size_t)sidx[q] + l];
}

__global__ void k_bucket_bounds(int n, const int* __restrict__ skey, int* __restrict__ bstart,
                                int* __restrict__ bend) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int b = skey[q];
  if (q == 0 || skey[q - 1] != b) bstart[b] = q;
  if (q == n - 1 || skey[q + 1] != b) bend[b] = q + 1;
}

// The search predicate, |p - P|^2 <= R^2 with the differences taken as
// p - P: every kernel that decides membership evaluates it this way, so the
// counting and filling passes agree bit for bit.
__device__ __forceinline__ bool within(const double* __restrict__ p, double P0, double P1,
                                       double P2, double R2) {
  const double d0 = p[0] - P0, d1 = p[1] - P1, d2 = p[2] - P2;
  return d0 * d0 + d1 * d1 + d2 * d2 <= R2;
}

// Buckets of the 27 cells around cell (a, b, c), each distinct one once: a
// bucket that two of the cells share (a hash collision) is replaced by ~0u
// the second time. No point is then met twice; points of other cells in a
// visited bucket fail the predicate, being farther than one cell.
__device__ __forceinline__ void cell_buckets(int a, int b, int c, unsigned mask,
                                             unsigned (&bk)[27]) {
#pragma unroll
  for (int k = 0; k < 27; k++)
    bk[k] = bucket_of(a + k % 3 - 1, b + (k / 3) % 3 - 1, c + k / 9 - 1, mask);
#pragma unroll
  for (int k = 1; k < 27; k++) {
    bool dup = false;
#pragma unroll
    for (int p = 0; p < k; p++) dup |= bk[p] == bk[k];
    if (dup) bk[k] = ~0u;
  }
}

// Contact count of every point: the points within the search radius in the
// buckets of its 27 cells, less its bonded neighbours within the radius.
// Every bonded neighbour within the radius is one of the former, so the
// count is that of the unbonded ones, with no bond test per candidate.
// Threads take the points in bucket order (t: the sorted position), so the
// lanes of a wave mostly share a cell; candidates are read from the
// bucket-ordered positions ys, the bonded neighbours from y. Points with
// contacts are appended to `over`.
__global__ __launch_bounds__(256) void k_contact_count(
    int n, const double* __restrict__ ys, const int* __restrict__ sidx,
    const double* __restrict__ y, double R2, double inv_cs, unsigned mask,
    const int* __restrict__ bstart, const int* __restrict__ bend, const int* __restrict__ brow,
    const int* __restrict__ bcol, long long* __restrict__ cnt, int* __restrict__ nover,
    int* __restrict__ over) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = sidx[t];
  const double Y0 = ys[3 * (size_t)t], Y1 = ys[3 * (size_t)t + 1], Y2 = ys[3 * (size_t)t + 2];
  unsigned bk[27];
  cell_buckets(cell_of(Y0, inv_cs), cell_of(Y1, inv_cs), cell_of(Y2, inv_cs), mask, bk);
  // four candidates (bonds) per iteration, so four position reads are in
  // flight at once; the remainder one at a time
  int count = 0;
  for (int k = 0; k < 27; k++) {
    if (bk[k] == ~0u) continue;
    const int s = bstart[bk[k]], e = bend[bk[k]];
    int q = s;
    for (; q + 4 <= e; q += 4) {
#pragma unroll
      for (int u = 0; u < 4; u++)
        count += (q + u != t && within(ys + 3 * (size_t)(q + u), Y0, Y1, Y2, R2)) ? 1 : 0;
    }
    for (; q < e; q++) count += (q != t && within(ys + 3 * (size_t)q, Y0, Y1, Y2, R2)) ? 1 : 0;
  }
  const int r0 = brow[i], r1 = brow[i + 1];
  int b = r0;
  for (; b + 4 <= r1; b += 4) {
    int jj[4];
#pragma unroll
    for (int u = 0; u < 4; u++) jj[u] = bcol[b + u];
#pragma unroll
    for (int u = 0; u < 4; u++) count -= within(y + 3 * (size_t)jj[u], Y0, Y1, Y2, R2) ? 1 : 0;
  }
  for (; b < r1; b++) count -= within(y + 3 * (size_t)bcol[b], Y0, Y1, Y2, R2) ? 1 : 0;
  cnt[i] = count;
  if (count > 0) over[atomicAdd(nover, 1)] = i;
}

constexpr int kFillGroup = 8;  // candidates per bond test of the filling pass

// Contact rows of the points of `pts` (those with contacts; their number in
// *npts), written at off[i] in visit order: the buckets of the 27 cells, each
// point's within the radius kept unless bonded. Candidates go in groups of
// kFillGroup: a group with a hit runs one branch-free search of the point's
// sorted bond row for all its candidates at once (the same row, so one trip
// count: the last entry <= j, then an equality test), so the wave waits for
// one chain of dependent row reads per group instead of one per hit.
__global__ __launch_bounds__(256) void k_contact_rows(
    const int* __restrict__ npts, const int* __restrict__ pts, const double* __restrict__ y,
    const double* __restrict__ ys, double R2, double inv_cs, unsigned mask,
    const int* __restrict__ bstart,
    const int* __restrict__ bend, const int* __restrict__ sidx, const int* __restrict__ brow,
    const int* __restrict__ bcol, const long long* __restrict__ off, int* __restrict__ list) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= *npts) return;
  const int i = pts[t];
  const double Y0 = y[3 * (size_t)i], Y1 = y[3 * (size_t)i + 1], Y2 = y[3 * (size_t)i + 2];
  unsigned bk[27];
  cell_buckets(cell_of(Y0, inv_cs), cell_of(Y1, inv_cs), cell_of(Y2, inv_cs), mask, bk);
  const int r0 = brow[i], r1 = brow[i + 1];
  int* const out = list + off[i];
  long long count = 0;
  for (int k = 0; k < 27; k++) {
    if (bk[k] == ~0u) continue;
    const int s = bstart[bk[k]], e = bend[bk[k]];
    for (int q0 = s; q0 < e; q0 += kFillGroup) {
      int jj[kFillGroup];
      bool hit[kFillGroup];
      bool any = false;
#pragma unroll
      for (int u = 0; u < kFillGroup; u++) {
        const int q = min(q0 + u, e - 1);  // past the bucket: a valid index, no hit
        jj[u] = sidx[q];
        hit[u] = q0 + u < e && jj[u] != i && within(ys + 3 * (size_t)q, Y0, Y1, Y2, R2);
        any |= hit[u];
      }
      if (!any) continue;
      int pos[kFillGroup];
#pragma unroll
      for (int u = 0; u < kFillGroup; u++) pos[u] = r0;
      for (int m = r1 - r0; m > 1;) {
        const int half = m >> 1;
#pragma unroll
        for (int u = 0; u < kFillGroup; u++)
          pos[u] = bcol[pos[u] + half] <= jj[u] ? pos[u] + half : pos[u];
        m -= half;
      }
#pragma unroll
      for (int u = 0; u < kFillGroup; u++)
        if (hit[u] && !(r1 > r0 && bcol[pos[u]] == jj[u])) out[count++] = jj[u];
    }
  }
}

// One pair of the short-range force model seen from `cur` (position X,
// velocity V, volume vc) with neighbour `nb` (y, v, vn); returns false beyond
// the contact radius. fc: the force density on cur; fn (BOTH only): on nb.
// Expressions in the reference's order.
template <bool BOTH>
__device__ __forceinline__ bool contact_pair(const PairParams& P, const double X[3],
                                             const double V[3], double vc, const double y[3],
                                             const double v[3], double vn, double fc[3],
                                             double fn[3]) {
  const double d0 = y[0] - X[0], d1 = y[1] - X[1], d2 = y[2] - X[2];
  const double dsq = d0 * d0 + d1 * d1 + d2 * d2;
  if (!(dsq < P.rc2)) return false;
  const double d = sqrt(dsq);
  const double temp = P.c * (P.rc - d) / P.horizon;
  const double cf0 = -(temp * vn * d0 / d), cf1 = -(temp * vn * d1 / d), cf2 = -(temp * vn * d2 / d);
  const double nf0 = temp * vc * d0 / d, nf1 = temp * vc * d1 / d, nf2 = temp * vc * d2 / d;
  fc[0] = cf0; fc[1] = cf1; fc[2] = cf2;
  if (BOTH) { fn[0] = nf0; fn[1] = nf1; fn[2] = nf2; }
  if (P.mu != 0.0) {
    const double n0 = d0 / d, n1 = d1 / d, n2 = d2 / d;
    const double cn = V[0] * n0 + V[1] * n1 + V[2] * n2;
    const double nbn = v[0] * n0 + v[1] * n1 + v[2] * n2;
    const double cp0 = V[0] - cn * n0, cp1 = V[1] - cn * n1, cp2 = V[2] - cn * n2;
    const double np0 = v[0] - nbn * n0, np1 = v[1] - nbn * n1, np2 = v[2] - nbn * n2;
    const double m0 = 0.5 * (cp0 + np0), m1 = 0.5 * (cp1 + np1), m2 = 0.5 * (cp2 + np2);
    const double cr0 = cp0 - m0, cr1 = cp1 - m1, cr2 = cp2 - m2;
    const double ncr = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
    const double ncf = sqrt(cf0 * cf0 + cf1 * cf1 + cf2 * cf2);
    if (ncr != 0.0) {
      fc[0] = cf0 + -P.mu * ncf * cr0 / ncr;
      fc[1] = cf1 + -P.mu * ncf * cr1 / ncr;
      fc[2] = cf2 + -P.mu * ncf * cr2 / ncr;
    }
    if (BOTH) {
      const double nr0 = np0 - m0, nr1 = np1 - m1, nr2 = np2 - m2;
      const double nnr = sqrt(nr0 * nr0 + nr1 * nr1 + nr2 * nr2);
      const double nnf = sqrt(nf0 * nf0 + nf1 * nf1 + nf2 * nf2);
      if (nnr != 0.0) {
        fn[0] = nf0 + -P.mu * nnf * nr0 / nnr;
        fn[1] = nf1 + -P.mu * nnf * nr1 / nnr;
        fn[2] = nf2 + -P.mu * nnf * nr2 / nnr;
      }
    }
  }
  return true;
}

// Points with a non-empty contact list, appended in any order (each point's
// force is computed on its own, so the order does not reach the results).
__global__ void k_contact_active(int n, const long long* __restrict__ cnt, int* __restrict__ nact,
                                 int* __restrict__ act) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || cnt[i] == 0) return;
  act[atomicAdd(nact, 1)] = i;
}

constexpr int kGatherLanes = 16;  // lanes per point of the contact gather

// Explicit-run contact force (internal order) of the points with contacts
// (the others keep the zero written at the search): each point gathers twice
// its own side of every pair of its (symmetric) list, kGatherLanes lanes per
// point, lane l taking entries l, l + kGatherLanes, ..., then a lane-tree sum.
// Positions as the reference forms them, y = (x + u) + dt v, from the record
// of the step's start (u) and the half-step velocity v (dt = 0 at the
// initial evaluation).
__global__ __launch_bounds__(256) void k_contact_gather(
    int nact, const int* __restrict__ act, const NodeRec* __restrict__ rec,
    const double* __restrict__ x, double q, double dt, const double* __restrict__ v,
    const double* __restrict__ vol, const long long* __restrict__ off,
    const int* __restrict__ list, PairParams P, double* __restrict__ fout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = t / kGatherLanes, lane = t % kGatherLanes;
  const bool valid = a < nact;
  const int i = valid ? act[a] : 0;
  double f[3] = {0.0, 0.0, 0.0};
  if (valid) {
    const long long e0 = off[i], e1 = off[i + 1];
    double X[3], V[3];
    for (int l = 0; l < 3; l++) {
      V[l] = v[3 * i + l];
      X[l] = (x[3 * i + l] + rec[i].u[l] * q) + dt * V[l];
    }
    const double vc = vol[i];
    for (long long e = e0 + lane; e < e1; e += kGatherLanes) {
      const int j = list[e];
      double Y[3], W[3], fc[3], fn[3];
      for (int l = 0; l < 3; l++) {
        W[l] = v[3 * j + l];
        Y[l] = (x[3 * j + l] + rec[j].u[l] * q) + dt * W[l];
      }
      if (contact_pair<false>(P, X, V, vc, Y, W, vol[j], fc, fn))
        for (int l = 0; l < 3; l++) f[l] += fc[l];
    }
  }
  // every lane takes part in the shuffles (the group of a point is aligned)
  for (int m = kGatherLanes / 2; m >= 1; m >>= 1)
    for (int l = 0; l < 3; l++) f[l] += __shfl_xor(f[l], m, kGatherLanes);
  if (valid && lane == 0)
    for (int l = 0; l < 3; l++) fout[3 * i + l] = 2.0 * f[l];
}

// Drop-in contact force (reference order, any list): both sides of every
// listed pair, the neighbour's through FP64 atomic adds.
__global__ __launch_bounds__(256) void k_contact_scatter(
    int n, const double* __restrict__ y, const double* __restrict__ v,
    const double* __restrict__ vol, const long long* __restrict__ head,
    const int* __restrict__ list, PairParams P, double* __restrict__ fout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long h = head[i];
  const int nn = list[h];
  if (nn <= 0) return;
  double X[3], V[3];
  for (int l = 0; l < 3; l++) { X[l] = y[3 * i + l]; V[l] = v[3 * i + l]; }
  const double vc = vol[i];
  double f[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nn; k++) {
    const int j = list[h + 1 + k];
    double Y[3], W[3], fc[3], fn[3];
    for (int l = 0; l < 3; l++) { Y[l] = y[3 * j + l]; W[l] = v[3 * j + l]; }
    if (contact_pair<true>(P, X, V, vc, Y, W, vol[j], fc, fn)) {
      for (int l = 0; l < 3; l++) {
        f[l] += fc[l];
        unsafeAtomicAdd(fout + 3 * (size_t)j + l, fn[l]);
      }
    }
  }
  for (int l = 0; l < 3; l++) unsafeAtomicAdd(fout + 3 * (size_t)i + l, f[l]);
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

PairParams pair_params(const Ctx* c) {
  const ContactState* S = c->contact;
  return PairParams{S->rc, S->rc * S->rc, S->c, c->horizon, S->mu};
}

template <class T>
int grow(Ctx* c, T*& p, long long& cap, long long need, const char* what) {
  if (need <= cap && p) return PDG_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const long long n = std::max<long long>(need + need / 4, 1024);
  int rc = c->hip(hipMalloc((void**)&p, (size_t)n * sizeof(T)), what);
  if (!rc) cap = n;
  return rc;
}

int tmp_reserve(Ctx* c, size_t bytes) {
  ContactState* S = c->contact;
  if (bytes <= S->tmp_bytes && S->tmp) return PDG_OK;
  if (S->tmp) (void)hipFree(S->tmp);
  S->tmp = nullptr;
  S->tmp_bytes = 0;
  int rc = c->hip(hipMalloc(&S->tmp, std::max<size_t>(bytes, 256)), "alloc contact temp");
  if (!rc) S->tmp_bytes = std::max<size_t>(bytes, 256);
  return rc;
}

// The contact list of the positions in S->y (internal order) into S->off /
// S->list; synchronises the stream once (the list size). For an explicit run
// (`run`) also the list of points with contacts and a zero contact force.
int search(Ctx* c, bool run) {
  ContactState* S = c->contact;
  const int N = c->N;
  hipStream_t s = c->stream;
  const unsigned nb = blocks_for(N);
  const unsigned mask = (1u << S->nbits) - 1u;
  hipLaunchKernelGGL(k_contact_keys, dim3(nb), dim3(256), 0, s, N, S->y, S->inv_cs, mask, S->key,
                     S->idx);
  size_t bytes = 0;
  int rc = c->hip(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, S->key, S->skey, S->idx,
                                                     S->sidx, N, 0, (int)S->nbits, s),
                  "contact sort size");
  if (!rc) rc = tmp_reserve(c, bytes);
  if (!rc) rc = c->hip(hipcub::DeviceRadixSort::SortPairs(S->tmp, bytes, S->key, S->skey, S->idx,
                                                          S->sidx, N, 0, (int)S->nbits, s),
                       "contact sort");
  if (rc) return rc;
  const long long T = 1ll << S->nbits;
  hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(T)), dim3(256), 0, s, T, S->bstart, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(T)), dim3(256), 0, s, T, S->bend, 0);
  hipLaunchKernelGGL(k_bucket_bounds, dim3(nb), dim3(256), 0, s, N, S->skey, S->bstart, S->bend);
  hipLaunchKernelGGL(k_contact_ys, dim3(blocks_for(3ll * N)), dim3(256), 0, s, N, S->sidx, S->y,
                     S->ys);
  const double R2 = S->R * S->R;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, 1ll, S->nover_d, 0);
  hipLaunchKernelGGL(k_contact_count, dim3(nb), dim3(256), 0, s, N, S->ys, S->sidx, S->y, R2,
                     S->inv_cs, mask, S->bstart, S->bend, c->d_row, S->bcol, S->cnt, S->nover_d,
                     S->over);
  rc = c->hip(hipGetLastError(), "contact count");
  bytes = 0;
  if (!rc) rc = c->hip(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, S->cnt, S->off, N + 1, s),
                       "contact scan size");
  if (!rc) rc = tmp_reserve(c, bytes);
  if (!rc) rc = c->hip(hipcub::DeviceScan::ExclusiveSum(S->tmp, bytes, S->cnt, S->off, N + 1, s),
                       "contact scan");
  // points with contacts; the force of every other point is zero until the
  // next search
  if (!rc && run) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, 1ll, S->nact_d, 0);
    hipLaunchKernelGGL(k_contact_active, dim3(nb), dim3(256), 0, s, N, S->cnt, S->nact_d, S->act);
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(3ll * N)), dim3(256), 0, s, 3ll * N, S->fc);
  }
  long long total = 0;
  int nact = 0;
  if (!rc) rc = c->hip(hipMemcpyAsync(&total, S->off + N, 8, hipMemcpyDeviceToHost, s), "d2h size");
  if (!rc && run) rc = c->hip(hipMemcpyAsync(&nact, S->nact_d, 4, hipMemcpyDeviceToHost, s), "d2h active");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (run) S->nact = nact;
  if (!rc) rc = grow(c, S->list, S->cap, total, "alloc contact list");
  if (rc) return rc;
  // the rows of the points with contacts (their number stays on the device:
  // threads past it return at once)
  if (total > 0)
    hipLaunchKernelGGL(k_contact_rows, dim3(nb), dim3(256), 0, s, S->nover_d, S->over, S->y,
                       S->ys, R2, S->inv_cs, mask, S->bstart, S->bend, S->sidx, c->d_row, S->bcol, S->off,
                       S->list);
  rc = c->hip(hipGetLastError(), "contact fill");
  S->size = total;
  S->built = run;
  return rc;
}

}  // namespace

void Ctx::free_contact() {
  if (!contact) return;
  ContactState* S = contact;
  void* ptrs[] = {S->y, S->fc, S->bcol, S->key, S->idx, S->skey, S->sidx, S->bstart,
                  S->bend, S->cnt, S->off, S->list, S->tmp, S->vref, S->dv, S->dfc,
                  S->dhead, S->dlist, S->act, S->nact_d, S->ys, S->over, S->nover_d};
  for (void* p : ptrs) if (p) (void)hipFree(p);
  delete S;
  contact = nullptr;
}

bool contact_on(const Ctx* c) { return c->contact != nullptr && c->N > 0; }

const double* contact_force_ptr(const Ctx* c) { return c->contact ? c->contact->fc : nullptr; }

// One explicit evaluation's contact force at y = x + u + dt v (u from `rec`),
// after a search in the configuration x + u of `rec_search` when that is not
// null.
int contact_eval(Ctx* c, const NodeRec* rec_search, const NodeRec* rec, double dt) {
  ContactState* S = c->contact;
  const int N = c->N;
  hipStream_t s = c->stream;
  int rc = PDG_OK;
  if (rec_search) {
    hipLaunchKernelGGL(k_contact_y, dim3(blocks_for(3ll * N)), dim3(256), 0, s, N, rec_search,
                       c->d_x, c->q, S->y);
    rc = search(c, true);
    if (rc) return rc;
  }
  if (S->nact == 0) return PDG_OK;  // all zero since the search
  hipLaunchKernelGGL(k_contact_gather, dim3(blocks_for((long long)S->nact * kGatherLanes)),
                     dim3(256), 0, s, S->nact, S->act, rec, c->d_x, c->q, dt, c->d_v, c->d_vol,
                     S->off, S->list, pair_params(c), S->fc);
  return c->hip(hipGetLastError(), "contact force");
}

// Step bookkeeping of an explicit run: evaluation k (0: initial) of a run.
// rec_prev holds the displacement at the start of step k (the previous
// evaluation's), rec the initial one at k = 0.
int contact_step(Ctx* c, int k, double dt, const NodeRec* rec_prev, const NodeRec* rec) {
  ContactState* S = c->contact;
  if (k == 0) {
    S->step = 0;
    return contact_eval(c, rec, rec, 0.0);
  }
  S->step++;
  const bool search_now = !S->built || S->step % S->freq == 0;
  return contact_eval(c, search_now ? rec_prev : nullptr, rec_prev, dt);
}

}  // namespace pdg

using namespace pdg;

extern "C" {

int pdg_set_contact(pdg_ctx* h, double searchRadius, int searchFrequency, double contactRadius,
                    double springConstant, double frictionCoefficient) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts) return PDG_E_ARG;
  if (!(contactRadius > 0.0)) {  // off
    c->free_contact();
    return PDG_OK;
  }
  if (!(searchRadius > 0.0) || searchFrequency < 1 || !std::isfinite(springConstant) ||
      !std::isfinite(frictionCoefficient) || !std::isfinite(searchRadius) ||
      !std::isfinite(contactRadius))
    return c->fail(PDG_E_ARG, "invalid contact parameters");
  if ((long long)c->col.size() != c->B || (int)c->row.size() != c->N + 1)
    return c->fail(PDG_E_ARG, "contact needs the bond lists of the constants");
  (void)hipSetDevice(c->device);
  c->free_contact();
  c->contact = new ContactState();
  ContactState* S = c->contact;
  S->R = searchRadius;
  S->rc = contactRadius;
  S->ks = springConstant;
  S->mu = frictionCoefficient;
  S->freq = searchFrequency;
  const double pi = 3.14159265358979323846;
  const double hz = c->horizon;
  S->c = 9.0 * springConstant / (pi * hz * hz * hz * hz);
  S->inv_cs = 1.0 / (searchRadius * 1.0000001);
  const int N = c->N;
  while ((1ll << S->nbits) < 2ll * std::max(N, 1)) S->nbits++;
  const long long T = 1ll << S->nbits;
  auto A = [&](auto*& p, long long n, const char* what) {
    return c->hip(hipMalloc((void**)&p, (size_t)std::max<long long>(n, 1) * sizeof(*p)), what);
  };
  const long long B = c->B;
  int rc = A(S->y, 3ll * N, "alloc contact y");
  if (!rc) rc = A(S->fc, 3ll * N, "alloc contact f");
  if (!rc) rc = A(S->bcol, B, "alloc contact bonds");
  if (!rc) rc = A(S->key, N, "alloc contact keys");
  if (!rc) rc = A(S->idx, N, "alloc contact keys");
  if (!rc) rc = A(S->skey, N, "alloc contact keys");
  if (!rc) rc = A(S->sidx, N, "alloc contact keys");
  if (!rc) rc = A(S->bstart, T, "alloc contact buckets");
  if (!rc) rc = A(S->bend, T, "alloc contact buckets");
  if (!rc) rc = A(S->cnt, N + 1ll, "alloc contact counts");
  if (!rc) rc = A(S->off, N + 1ll, "alloc contact offsets");
  if (!rc) rc = A(S->vref, N, "alloc contact volumes");
  if (!rc) rc = A(S->act, N, "alloc contact active list");
  if (!rc) rc = A(S->nact_d, 1, "alloc contact active count");
  if (!rc) rc = A(S->ys, 3ll * N, "alloc contact sorted positions");
  if (!rc) rc = A(S->over, N, "alloc contact point list");
  if (!rc) rc = A(S->nover_d, 1, "alloc contact point count");
  if (rc) { c->free_contact(); return rc; }
  // bond rows (internal ids) sorted ascending
  std::vector<int> bc(c->col.begin(), c->col.end());
  for (int k = 0; k < N; k++) std::sort(bc.begin() + c->row[k], bc.begin() + c->row[k + 1]);
  hipStream_t s = c->stream;
  const long long zero = 0;
  if (B > 0) rc = c->hip(hipMemcpyAsync(S->bcol, bc.data(), 4 * (size_t)B, hipMemcpyHostToDevice, s), "h2d bonds");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->cnt + N, &zero, 8, hipMemcpyHostToDevice, s), "h2d zero");
  if (!rc && N > 0) rc = c->hip(hipMemcpyAsync(S->vref, c->vol.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d vol");
  if (!rc) hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(3ll * N)), dim3(256), 0, s, 3ll * N, S->fc);
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) c->free_contact();
  return rc;
}

int pdg_contact_search(pdg_ctx* h, const double* y, int* contactList, long long* listSize) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !listSize || (c->N > 0 && !y)) return PDG_E_ARG;
  if (!c->contact) return c->fail(PDG_E_ARG, "no contact model (pdg_set_contact)");
  (void)hipSetDevice(c->device);
  ContactState* S = c->contact;
  const int N = c->N;
  hipStream_t s = c->stream;
  int rc = c->hip(hipMemcpyAsync(c->d_y, y, 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d y");
  if (rc) return rc;
  hipLaunchKernelGGL(k_contact_y_perm, dim3(blocks_for(3ll * N)), dim3(256), 0, s, N, c->d_y,
                     c->d_perm, S->y);
  // the explicit run rebuilds its list from its own state afterwards
  rc = search(c, false);
  if (rc) return rc;
  const long long size = S->size + N;
  const long long cap = *listSize;
  *listSize = size;
  if (!contactList) return PDG_OK;
  if (cap < size) return c->fail(PDG_E_ARG, "contact list capacity too small");
  std::vector<long long> off(N + 1);
  std::vector<int> lst((size_t)S->size);
  rc = c->hip(hipMemcpyAsync(off.data(), S->off, 8 * (size_t)(N + 1), hipMemcpyDeviceToHost, s), "d2h off");
  if (!rc && S->size > 0)
    rc = c->hip(hipMemcpyAsync(lst.data(), S->list, 4 * (size_t)S->size, hipMemcpyDeviceToHost, s), "d2h list");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  // reference order, neighbours ascending
  long long p = 0;
  std::vector<int> r;
  for (int g = 0; g < N; g++) {
    const int k = c->iperm[g];
    r.clear();
    for (long long e = off[k]; e < off[k + 1]; e++) r.push_back(c->perm[lst[e]]);
    std::sort(r.begin(), r.end());
    contactList[p++] = (int)r.size();
    for (int j : r) contactList[p++] = j;
  }
  return PDG_OK;
}

int pdg_contact_compute_force(pdg_ctx* h, double dt, const double* y, const double* v,
                              const int* contactList, long long listSize, double* contactForce) {
  (void)dt;
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || (c->N > 0 && (!y || !v || !contactList || !contactForce))) return PDG_E_ARG;
  if (!c->contact) return c->fail(PDG_E_ARG, "no contact model (pdg_set_contact)");
  const int N = c->N;
  if (N == 0) return PDG_OK;
  // list heads, checked against the stated size and the node range
  std::vector<long long> head(N);
  long long p = 0;
  for (int i = 0; i < N; i++) {
    if (p >= listSize) return c->fail(PDG_E_ARG, "contact list shorter than its counts");
    head[i] = p;
    const int nn = contactList[p];
    if (nn < 0 || p + 1 + nn > listSize) return c->fail(PDG_E_ARG, "contact list shorter than its counts");
    for (int k = 0; k < nn; k++) {
      const int j = contactList[p + 1 + k];
      if (j < 0 || j >= N) return c->fail(PDG_E_ARG, "contact neighbour out of range");
    }
    p += 1 + nn;
  }
  (void)hipSetDevice(c->device);
  ContactState* S = c->contact;
  hipStream_t s = c->stream;
  int rc = PDG_OK;
  if (!S->dv) rc = c->hip(hipMalloc((void**)&S->dv, 24 * (size_t)N), "alloc contact v");
  if (!rc && !S->dfc) rc = c->hip(hipMalloc((void**)&S->dfc, 24 * (size_t)N), "alloc contact f");
  if (!rc && !S->dhead) rc = c->hip(hipMalloc((void**)&S->dhead, 8 * (size_t)N), "alloc contact heads");
  if (!rc) rc = grow(c, S->dlist, S->dcap, p, "alloc contact list");
  if (!rc) rc = c->hip(hipMemcpyAsync(c->d_y, y, 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d y");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->dv, v, 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d v");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->dhead, head.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d heads");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->dlist, contactList, 4 * (size_t)p, hipMemcpyHostToDevice, s), "h2d list");
  if (rc) return rc;
  hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(3ll * N)), dim3(256), 0, s, 3ll * N, S->dfc);
  hipLaunchKernelGGL(k_contact_scatter, dim3(blocks_for(N)), dim3(256), 0, s, N, c->d_y, S->dv,
                     S->vref, S->dhead, S->dlist, pair_params(c), S->dfc);
  rc = c->hip(hipGetLastError(), "contact force");
  if (!rc) rc = c->hip(hipMemcpyAsync(contactForce, S->dfc, 24 * (size_t)N, hipMemcpyDeviceToHost, s), "d2h f");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  for (long long q = 0; q < 3ll * N; q++)
    if (!std::isfinite(contactForce[q]))
      return c->fail(PDG_E_NONFINITE, "NaN returned by contact force evaluation");
  return PDG_OK;
}

int pdg_contact_force_density(pdg_ctx* h, double* contactForce) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || (c->N > 0 && !contactForce)) return PDG_E_ARG;
  if (!c->contact) return c->fail(PDG_E_ARG, "no contact model (pdg_set_contact)");
  (void)hipSetDevice(c->device);
  return c->download_vec3(c->contact->fc, contactForce);
}

}  // extern "C"
