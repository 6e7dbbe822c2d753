// Neighbour-list construction on the device. Integer results from FP64
// predicates evaluated with the reference's operation order and no FMA
// contraction (built with -ffp-contract=off), so the lists are bit-identical
// to the reference builders.
//   QuickGrid lattice:  src/io/mesh_input/quick_grid/QuickGrid.cxx:160-165,
//                       :243-253, :394-408, :510-647
//   radius search:      src/io/Peridigm_ZoltanSearchTree.cpp:236-246 (<= R^2),
//                       rows sorted by id (src/io/pdneigh/NeighborhoodList.cxx:705)
//   finite-plane filter: src/io/pdneigh/BondFilter.cxx:59-153

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "pdg_internal.hpp"

namespace {

struct QGrid {
  int n[3];
  int hw[3];
  const double* c[3];
  double r;
  int spherical;
};

__device__ __forceinline__ void qg_window(int i, int h, int n, int& s, int& e) {
  s = i > h ? i - h : 0;
  e = i < (n - h) ? i + h + 1 : n;
}

// Visits the candidate window of lattice cell (i,j,k) in kk -> jj -> ii order.
template <bool FILL>
__global__ __launch_bounds__(256) void k_quickgrid(QGrid g, long long* __restrict__ count,
                                                   const long long* __restrict__ head,
                                                   int* __restrict__ nlist) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)g.n[0] * g.n[1] * g.n[2];
  if (gid >= N) return;
  const int i = (int)(gid % g.n[0]);
  const int j = (int)((gid / g.n[0]) % g.n[1]);
  const int k = (int)(gid / ((long long)g.n[0] * g.n[1]));
  const double p0 = g.c[0][i], p1 = g.c[1][j], p2 = g.c[2][k];
  int xs, xe, ys, ye, zs, ze;
  qg_window(i, g.hw[0], g.n[0], xs, xe);
  qg_window(j, g.hw[1], g.n[1], ys, ye);
  qg_window(k, g.hw[2], g.n[2], zs, ze);
  long long out = FILL ? head[gid] + 1 : 0;
  int cnt = 0;
  for (int kk = zs; kk < ze; kk++)
    for (int jj = ys; jj < ye; jj++)
      for (int ii = xs; ii < xe; ii++) {
        if (ii == i && jj == j && kk == k) continue;
        bool in = true;
        if (g.spherical) {
          const double dx = p0 - g.c[0][ii];
          const double dy = p1 - g.c[1][jj];
          const double dz = p2 - g.c[2][kk];
          in = dx * dx + dy * dy + dz * dz - g.r * g.r < 0.0;
        }
        if (in) {
          if (FILL) nlist[out++] = (int)(ii + (long long)jj * g.n[0] + (long long)kk * g.n[0] * g.n[1]);
          cnt++;
        }
      }
  if (FILL) nlist[head[gid]] = cnt;
  else count[gid] = cnt;
}

__global__ void k_head(long long n, const long long* __restrict__ excl, long long* __restrict__ head) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = excl[i] + i;
}

// ---- radius search --------------------------------------------------------
struct RGrid {
  double lo[3];
  double cs;
  int nc[3];
  double R2;
};

__device__ __forceinline__ int cell_coord(double x, double lo, double cs, int nc) {
  int c = (int)floor((x - lo) / cs);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void k_cell_ids(int n, const double* __restrict__ x, RGrid g, int* __restrict__ cell,
                           int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = cell_coord(x[3 * i], g.lo[0], g.cs, g.nc[0]);
  const int b = cell_coord(x[3 * i + 1], g.lo[1], g.cs, g.nc[1]);
  const int c = cell_coord(x[3 * i + 2], g.lo[2], g.cs, g.nc[2]);
  cell[i] = (c * g.nc[1] + b) * g.nc[0] + a;
  idx[i] = i;
}

__global__ void k_cell_start(int n, const int* __restrict__ scell, int* __restrict__ cstart) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int c = scell[q];
  if (q == 0 || scell[q - 1] != c) cstart[c] = q;
}

__global__ void k_cell_end(int n, const int* __restrict__ scell, int* __restrict__ cend) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int c = scell[q];
  if (q == n - 1 || scell[q + 1] != c) cend[c] = q + 1;
}

// Candidates from the 27 surrounding cells; membership by the predicate only.
template <bool FILL>
__global__ __launch_bounds__(256) void k_radius(int n, const double* __restrict__ x, RGrid g,
                                                const int* __restrict__ cstart,
                                                const int* __restrict__ cend,
                                                const int* __restrict__ sidx,
                                                long long* __restrict__ count,
                                                const long long* __restrict__ off,
                                                unsigned long long* __restrict__ keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double X0 = x[3 * i], X1 = x[3 * i + 1], X2 = x[3 * i + 2];
  const int a = cell_coord(X0, g.lo[0], g.cs, g.nc[0]);
  const int b = cell_coord(X1, g.lo[1], g.cs, g.nc[1]);
  const int c = cell_coord(X2, g.lo[2], g.cs, g.nc[2]);
  long long o = FILL ? off[i] : 0;
  int cnt = 0;
  for (int dz = -1; dz <= 1; dz++) {
    const int cz = c + dz;
    if (cz < 0 || cz >= g.nc[2]) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int cy = b + dy;
      if (cy < 0 || cy >= g.nc[1]) continue;
      for (int dx = -1; dx <= 1; dx++) {
        const int cx = a + dx;
        if (cx < 0 || cx >= g.nc[0]) continue;
        const int cc = (cz * g.nc[1] + cy) * g.nc[0] + cx;
        const int s = cstart[cc], e = cend[cc];
        for (int q = s; q < e; q++) {
          const int j = sidx[q];
          if (j == i) continue;
          const double ddx = (x[3 * j] - X0);
          const double ddy = (x[3 * j + 1] - X1);
          const double ddz = (x[3 * j + 2] - X2);
          if (ddx * ddx + ddy * ddy + ddz * ddz <= g.R2) {
            if (FILL) keys[o++] = ((unsigned long long)i << 32) | (unsigned)j;
            cnt++;
          }
        }
      }
    }
  }
  if (!FILL) count[i] = cnt;
}

// FinitePlane::bondIntersectInfinitePlane + bondIntersect (tolerance 1e-15),
// FinitePlaneFilter::filterBonds with p0 = the point, p1 = its neighbour.
__device__ bool plane_cuts(const double* pl, const double* p0, const double* p1) {
  const double* nrm = pl;
  const double* r0 = pl + 3;
  const double* ub = pl + 6;
  const double b = pl[9], a = pl[10];
  double ua[3];
  ua[0] = -ub[2] * nrm[1] + ub[1] * nrm[2];
  ua[1] = ub[2] * nrm[0] - ub[0] * nrm[2];
  ua[2] = -ub[1] * nrm[0] + ub[0] * nrm[1];
  const double numerator =
      (r0[0] - p0[0]) * nrm[0] + (r0[1] - p0[1]) * nrm[1] + (r0[2] - p0[2]) * nrm[2];
  const double denominator =
      (p1[0] - p0[0]) * nrm[0] + (p1[1] - p0[1]) * nrm[1] + (p1[2] - p0[2]) * nrm[2];
  double t;
  if (fabs(denominator) < 1.0e-14) t = DBL_MAX;
  else t = numerator / denominator;
  if (!(t >= 0.0 && t <= 1.0)) return false;
  double xi[3];
  xi[0] = p0[0] + t * (p1[0] - p0[0]);
  xi[1] = p0[1] + t * (p1[1] - p0[1]);
  xi[2] = p0[2] + t * (p1[2] - p0[2]);
  const double zero = 1.0e-15;
  const double one = 1.0 + zero;
  const double dr0 = xi[0] - r0[0], dr1 = xi[1] - r0[1], dr2 = xi[2] - r0[2];
  const double aa = dr0 * ua[0] + dr1 * ua[1] + dr2 * ua[2];
  const double bb = dr0 * ub[0] + dr1 * ub[1] + dr2 * ub[2];
  return -zero < aa && aa / a < one && -zero < bb && bb / b < one;
}

__global__ void k_plane_flags(long long nb, const unsigned long long* __restrict__ keys,
                              const double* __restrict__ x, int nplanes,
                              const double* __restrict__ planes, int* __restrict__ keep) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  const int i = (int)(keys[q] >> 32), j = (int)(keys[q] & 0xffffffffu);
  bool cut = false;
  for (int p = 0; p < nplanes && !cut; p++) cut = plane_cuts(planes + 11 * p, x + 3 * i, x + 3 * j);
  keep[q] = cut ? 0 : 1;
}

__global__ void k_row_count(long long nb, const unsigned long long* __restrict__ keys,
                            const int* __restrict__ keep, long long* __restrict__ count) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb || !keep[q]) return;
  atomicAdd((unsigned long long*)&count[keys[q] >> 32], 1ull);
}

__global__ void k_emit(long long nb, const unsigned long long* __restrict__ keys,
                       const int* __restrict__ keep, const long long* __restrict__ kpos,
                       const long long* __restrict__ head, int* __restrict__ nlist) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb || !keep[q]) return;
  // keys are sorted by (row, col); kpos = exclusive scan of keep, so the
  // position inside the row is kpos[q] - (kept bonds before this row)
  const int i = (int)(keys[q] >> 32);
  const long long before_row = head[i] - i;  // kept bonds of rows < i
  nlist[head[i] + 1 + (kpos[q] - before_row)] = (int)(keys[q] & 0xffffffffu);
}

__global__ void k_widen(long long nb, const int* __restrict__ a, long long* __restrict__ b) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nb) b[q] = a[q];
}

__global__ void k_emit_counts(int n, const long long* __restrict__ count,
                              const long long* __restrict__ head, int* __restrict__ nlist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) nlist[head[i]] = (int)count[i];
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  hipError_t alloc(size_t n) { return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)); }
  ~DevBuf() { if (p) (void)hipFree(p); }
};

#define PDG_CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return PDG_E_HIP; } while (0)

int exclusive_scan(const long long* in, long long* out, long long n, hipStream_t s) {
  size_t bytes = 0;
  PDG_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
  DevBuf<char> tmp;
  PDG_CK(tmp.alloc(bytes));
  PDG_CK(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, s));
  return PDG_OK;
}

}  // namespace

extern "C" {

int pdg_quickgrid(int nx, int ny, int nz, const double origin[3], const double length[3],
                  double horizon, int spherical, double* x, double* volume, int* nlist,
                  long long* listSize) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !origin || !length || !listSize) return PDG_E_ARG;
  const long long N = (long long)nx * ny * nz;
  if (N >= (1ll << 31)) return PDG_E_UNSUPPORTED;
  const int n[3] = {nx, ny, nz};
  std::vector<double> c[3];
  int hw[3];
  for (int d = 0; d < 3; d++) {
    // Spec1D::getCellSize, getDiscretization (centres by accumulation),
    // getCellNeighborhoodSize (window in cells, rounded, clamped).
    const double cs = length[d] / n[d];
    c[d].resize(n[d]);
    double p = origin[d] + cs / 2.0;
    for (int i = 0; i < n[d]; p += cs, i++) c[d][i] = p;
    const double dCx = horizon / (cs * 1.0);
    size_t nC = static_cast<size_t>(dCx);
    if (std::abs(dCx - nC) >= .5) nC += 1;
    if (nC > (size_t)n[d]) nC = n[d];
    hw[d] = (int)nC;
  }
  hipStream_t s = nullptr;
  DevBuf<double> dc0, dc1, dc2;
  DevBuf<long long> cnt, excl, head;
  PDG_CK(dc0.alloc(nx)); PDG_CK(dc1.alloc(ny)); PDG_CK(dc2.alloc(nz));
  PDG_CK(hipMemcpy(dc0.p, c[0].data(), 8 * nx, hipMemcpyHostToDevice));
  PDG_CK(hipMemcpy(dc1.p, c[1].data(), 8 * ny, hipMemcpyHostToDevice));
  PDG_CK(hipMemcpy(dc2.p, c[2].data(), 8 * nz, hipMemcpyHostToDevice));
  QGrid g;
  for (int d = 0; d < 3; d++) { g.n[d] = n[d]; g.hw[d] = hw[d]; }
  g.c[0] = dc0.p; g.c[1] = dc1.p; g.c[2] = dc2.p;
  g.r = horizon;
  g.spherical = spherical;
  PDG_CK(cnt.alloc(N)); PDG_CK(excl.alloc(N)); PDG_CK(head.alloc(N));
  const unsigned blocks = (unsigned)((N + 255) / 256);
  hipLaunchKernelGGL((k_quickgrid<false>), dim3(blocks), dim3(256), 0, s, g, cnt.p, nullptr, nullptr);
  PDG_CK(hipGetLastError());
  int rc = exclusive_scan(cnt.p, excl.p, N, s);
  if (rc) return rc;
  long long lastc = 0, laste = 0;
  PDG_CK(hipMemcpy(&lastc, cnt.p + N - 1, 8, hipMemcpyDeviceToHost));
  PDG_CK(hipMemcpy(&laste, excl.p + N - 1, 8, hipMemcpyDeviceToHost));
  const long long size = laste + lastc + N;
  const long long cap = *listSize;
  *listSize = size;
  if (x)
    for (long long gid = 0; gid < N; gid++) {
      const int i = (int)(gid % nx), j = (int)((gid / nx) % ny), k = (int)(gid / ((long long)nx * ny));
      x[3 * gid] = c[0][i]; x[3 * gid + 1] = c[1][j]; x[3 * gid + 2] = c[2][k];
    }
  if (volume) {
    const double cv = (length[0] / nx) * (length[1] / ny) * (length[2] / nz);
    for (long long gid = 0; gid < N; gid++) volume[gid] = cv;
  }
  if (!nlist) return PDG_OK;
  if (cap < size) return PDG_E_ARG;
  hipLaunchKernelGGL(k_head, dim3(blocks), dim3(256), 0, s, N, excl.p, head.p);
  DevBuf<int> dl;
  PDG_CK(dl.alloc(size));
  hipLaunchKernelGGL((k_quickgrid<true>), dim3(blocks), dim3(256), 0, s, g, nullptr, head.p, dl.p);
  PDG_CK(hipGetLastError());
  PDG_CK(hipMemcpy(nlist, dl.p, 4 * size, hipMemcpyDeviceToHost));
  return PDG_OK;
}

int pdg_radius_search(int N, const double* x, double R, int nplanes, const double* planes,
                      int* nlist, long long* listSize) {
  if (N <= 0 || !x || !(R > 0.0) || !listSize || nplanes < 0 || (nplanes > 0 && !planes))
    return PDG_E_ARG;
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = 0; i < N; i++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], x[3 * (size_t)i + d]);
      hi[d] = std::max(hi[d], x[3 * (size_t)i + d]);
    }
  RGrid g;
  g.cs = R * 1.0000001;
  long long ncell = 1;
  for (int d = 0; d < 3; d++) {
    g.lo[d] = lo[d];
    g.nc[d] = std::max(1, (int)std::floor((hi[d] - lo[d]) / g.cs) + 1);
    ncell *= g.nc[d];
  }
  if (ncell > (1ll << 30)) return PDG_E_UNSUPPORTED;
  g.R2 = R * R;
  hipStream_t s = nullptr;
  DevBuf<double> dx;
  DevBuf<int> cell, idx, scell, sidx, cstart, cend;
  PDG_CK(dx.alloc(3 * (size_t)N));
  PDG_CK(hipMemcpy(dx.p, x, 24 * (size_t)N, hipMemcpyHostToDevice));
  PDG_CK(cell.alloc(N)); PDG_CK(idx.alloc(N)); PDG_CK(scell.alloc(N)); PDG_CK(sidx.alloc(N));
  PDG_CK(cstart.alloc(ncell)); PDG_CK(cend.alloc(ncell));
  const unsigned nb = (unsigned)((N + 255) / 256);
  hipLaunchKernelGGL(k_cell_ids, dim3(nb), dim3(256), 0, s, N, dx.p, g, cell.p, idx.p);
  {
    size_t bytes = 0;
    PDG_CK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, cell.p, scell.p, idx.p, sidx.p, N, 0, 32, s));
    DevBuf<char> tmp;
    PDG_CK(tmp.alloc(bytes));
    PDG_CK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, cell.p, scell.p, idx.p, sidx.p, N, 0, 32, s));
  }
  PDG_CK(hipMemsetAsync(cstart.p, 0, 4 * ncell, s));
  PDG_CK(hipMemsetAsync(cend.p, 0, 4 * ncell, s));
  hipLaunchKernelGGL(k_cell_start, dim3(nb), dim3(256), 0, s, N, scell.p, cstart.p);
  hipLaunchKernelGGL(k_cell_end, dim3(nb), dim3(256), 0, s, N, scell.p, cend.p);
  DevBuf<long long> cnt, off;
  PDG_CK(cnt.alloc(N)); PDG_CK(off.alloc(N));
  hipLaunchKernelGGL((k_radius<false>), dim3(nb), dim3(256), 0, s, N, dx.p, g, cstart.p, cend.p,
                     sidx.p, cnt.p, nullptr, nullptr);
  PDG_CK(hipGetLastError());
  int rc = exclusive_scan(cnt.p, off.p, N, s);
  if (rc) return rc;
  long long lastc = 0, lasto = 0;
  PDG_CK(hipMemcpy(&lastc, cnt.p + N - 1, 8, hipMemcpyDeviceToHost));
  PDG_CK(hipMemcpy(&lasto, off.p + N - 1, 8, hipMemcpyDeviceToHost));
  const long long B0 = lasto + lastc;
  DevBuf<unsigned long long> keys, skeys;
  PDG_CK(keys.alloc(B0)); PDG_CK(skeys.alloc(B0));
  hipLaunchKernelGGL((k_radius<true>), dim3(nb), dim3(256), 0, s, N, dx.p, g, cstart.p, cend.p,
                     sidx.p, nullptr, off.p, keys.p);
  PDG_CK(hipGetLastError());
  {
    size_t bytes = 0;
    PDG_CK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.p, skeys.p, (int)B0, 0, 64, s));
    DevBuf<char> tmp;
    PDG_CK(tmp.alloc(bytes));
    PDG_CK(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, keys.p, skeys.p, (int)B0, 0, 64, s));
  }
  // plane filter, per-row counts of kept bonds, compaction
  DevBuf<int> keep;
  DevBuf<double> dpl;
  PDG_CK(keep.alloc(B0));
  PDG_CK(dpl.alloc(std::max(1, 11 * nplanes)));
  if (nplanes > 0) PDG_CK(hipMemcpy(dpl.p, planes, 88 * (size_t)nplanes, hipMemcpyHostToDevice));
  const unsigned bb = (unsigned)((B0 + 255) / 256);
  if (B0 > 0)
    hipLaunchKernelGGL(k_plane_flags, dim3(bb), dim3(256), 0, s, B0, skeys.p, dx.p, nplanes, dpl.p, keep.p);
  PDG_CK(hipMemsetAsync(cnt.p, 0, 8 * (size_t)N, s));
  if (B0 > 0) hipLaunchKernelGGL(k_row_count, dim3(bb), dim3(256), 0, s, B0, skeys.p, keep.p, cnt.p);
  rc = exclusive_scan(cnt.p, off.p, N, s);
  if (rc) return rc;
  PDG_CK(hipMemcpy(&lastc, cnt.p + N - 1, 8, hipMemcpyDeviceToHost));
  PDG_CK(hipMemcpy(&lasto, off.p + N - 1, 8, hipMemcpyDeviceToHost));
  const long long B = lasto + lastc;
  const long long size = B + N;
  const long long cap = *listSize;
  *listSize = size;
  if (!nlist) return PDG_OK;
  if (cap < size) return PDG_E_ARG;
  DevBuf<long long> head, keep64, kpos;
  PDG_CK(head.alloc(N)); PDG_CK(keep64.alloc(B0)); PDG_CK(kpos.alloc(B0));
  hipLaunchKernelGGL(k_head, dim3(nb), dim3(256), 0, s, (long long)N, off.p, head.p);
  // kpos: exclusive scan of keep (as 64-bit)
  if (B0 > 0) hipLaunchKernelGGL(k_widen, dim3(bb), dim3(256), 0, s, B0, keep.p, keep64.p);
  if (B0 > 0) { rc = exclusive_scan(keep64.p, kpos.p, B0, s); if (rc) return rc; }
  DevBuf<int> dl;
  PDG_CK(dl.alloc(size));
  hipLaunchKernelGGL(k_emit_counts, dim3(nb), dim3(256), 0, s, N, cnt.p, head.p, dl.p);
  if (B0 > 0) hipLaunchKernelGGL(k_emit, dim3(bb), dim3(256), 0, s, B0, skeys.p, keep.p, kpos.p, head.p, dl.p);
  PDG_CK(hipGetLastError());
  PDG_CK(hipMemcpy(nlist, dl.p, 4 * size, hipMemcpyDeviceToHost));
  return PDG_OK;
}

}  // extern "C"
