// FP64 CPU oracle for the Peridigm LPS + critical-stretch force path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, operation for operation, the
// reference loops that the GPU path replaces, so that tests can check the GPU
// results against it. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product library never does.
//
// Build: g++ -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile). No
// -march flags: contraction into FMA would change the last bits and break the
// bit-exact integer checks that depend on FP64 predicates.
//
// Every function cites the reference file:line it follows (paths relative to
// the Peridigm source root). Parity pins for each function are listed in the
// tests that exercise them (tests/test_oracle_golden.py).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <cfloat>

namespace {

// src/core/Peridigm_InfluenceFunction.hpp:69-88 (One, Parabolic Decay, Gaussian)
inline double omega_fn(int kind, double zeta, double horizon) {
  if (kind == 0) return 1.0;
  if (kind == 1) {
    if (zeta > horizon) return 0.0;
    double s = zeta / horizon;
    return s < 0.5 ? 1.0 : -4.0 * s * s + 4.0 * s;
  }
  double h2 = horizon * horizon * 0.4 * 0.4;
  double xi2 = zeta * zeta;
  return std::exp(-xi2 / h2);
}

// Forward-mode dual number used for the directional derivative J.w of the
// force (the reference obtains the tangent with Sacado DFad over the same
// templated kernel, src/materials/Peridigm_ElasticMaterial.cpp:258-383).
struct Dual {
  double v, d;
  Dual() : v(0), d(0) {}
  Dual(double a) : v(a), d(0) {}
  Dual(double a, double b) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
  return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline Dual dsqrt(Dual a) {
  double s = std::sqrt(a.v);
  return Dual(s, a.d / (2.0 * s));
}
inline double dsqrt(double a) { return std::sqrt(a); }
inline double val(double a) { return a; }
inline double val(Dual a) { return a.v; }

// src/materials/material_utilities.cxx:235-288 computeDilatation<ScalarT>
// With deltaT: the thermal strain of :281-282, e -= alphaT deltaT_p zeta.
template <typename T>
void dilatation_t(int numOwned, const double* x, const T* y, const double* m,
                  const double* vol, const double* bondDamage, T* theta,
                  const int* nlist, double horizon, int infl, const double* deltaT = nullptr,
                  double alphaT = 0.0) {
  const int* neighPtr = nlist;
  const double* bd = bondDamage;
  for (int p = 0; p < numOwned; p++) {
    int numNeigh = *neighPtr; neighPtr++;
    const double* X = x + 3 * p;
    const T* Y = y + 3 * p;
    T th = T(0.0);
    for (int n = 0; n < numNeigh; n++, neighPtr++, bd++) {
      int j = *neighPtr;
      double cellVolume = vol[j];
      const double* XP = x + 3 * j;
      const T* YP = y + 3 * j;
      double X_dx = XP[0] - X[0];
      double X_dy = XP[1] - X[1];
      double X_dz = XP[2] - X[2];
      double zetaSquared = X_dx * X_dx + X_dy * X_dy + X_dz * X_dz;
      T Y_dx = YP[0] - Y[0];
      T Y_dy = YP[1] - Y[1];
      T Y_dz = YP[2] - Y[2];
      T dY = Y_dx * Y_dx + Y_dy * Y_dy + Y_dz * Y_dz;
      double d = std::sqrt(zetaSquared);
      T e = dsqrt(dY);
      e = e - T(d);
      if (deltaT) e = e - T(alphaT * deltaT[p] * d);
      double omega = omega_fn(infl, d, horizon);
      th += 3.0 * omega * (1.0 - *bd) * d * e * cellVolume / m[p];
    }
    theta[p] = th;
  }
}

// src/materials/elastic.cxx:55-146 computeInternalForceLinearElastic<ScalarT>
// (owner-side gather plus neighbour-side scatter; f must be zeroed first as in
// src/materials/Peridigm_ElasticMaterial.cpp:149). With deltaT the thermal
// strain (:112-113); with ps the owner-side partial stress, 9 per point,
// row = force component, column = reference bond component (:132-142).
template <typename T>
void force_t(int numOwned, const double* x, const T* y, const double* m,
             const double* vol, const T* theta, const double* bondDamage, T* f,
             const int* nlist, double K, double MU, double horizon, int infl,
             const double* deltaT = nullptr, double alphaT = 0.0, T* ps = nullptr) {
  const int* neighPtr = nlist;
  const double* bd = bondDamage;
  for (int p = 0; p < numOwned; p++) {
    int numNeigh = *neighPtr; neighPtr++;
    const double* X = x + 3 * p;
    const T* Y = y + 3 * p;
    double alpha = 15.0 * MU / m[p];
    double selfCellVolume = vol[p];
    for (int n = 0; n < numNeigh; n++, neighPtr++, bd++) {
      int j = *neighPtr;
      double cellVolume = vol[j];
      const double* XP = x + 3 * j;
      const T* YP = y + 3 * j;
      double X_dx = XP[0] - X[0];
      double X_dy = XP[1] - X[1];
      double X_dz = XP[2] - X[2];
      double zeta = std::sqrt(X_dx * X_dx + X_dy * X_dy + X_dz * X_dz);
      T Y_dx = YP[0] - Y[0];
      T Y_dy = YP[1] - Y[1];
      T Y_dz = YP[2] - Y[2];
      T dY = dsqrt(Y_dx * Y_dx + Y_dy * Y_dy + Y_dz * Y_dz);
      T e = dY - T(zeta);
      if (deltaT) e = e - T(alphaT * deltaT[p] * zeta);
      double omega = omega_fn(infl, zeta, horizon);
      T c1 = omega * theta[p] * (3.0 * K / m[p] - alpha / 3.0);
      T t = (1.0 - *bd) * (c1 * zeta + (1.0 - *bd) * omega * alpha * e);
      T fx = t * Y_dx / dY;
      T fy = t * Y_dy / dY;
      T fz = t * Y_dz / dY;
      f[3 * p + 0] += fx * cellVolume;
      f[3 * p + 1] += fy * cellVolume;
      f[3 * p + 2] += fz * cellVolume;
      f[3 * j + 0] -= fx * selfCellVolume;
      f[3 * j + 1] -= fy * selfCellVolume;
      f[3 * j + 2] -= fz * selfCellVolume;
      if (ps) {
        T* q = ps + 9 * p;
        q[0] += fx * X_dx * cellVolume; q[1] += fx * X_dy * cellVolume; q[2] += fx * X_dz * cellVolume;
        q[3] += fy * X_dx * cellVolume; q[4] += fy * X_dy * cellVolume; q[5] += fy * X_dz * cellVolume;
        q[6] += fz * X_dx * cellVolume; q[7] += fz * X_dy * cellVolume; q[8] += fz * X_dz * cellVolume;
      }
    }
  }
}

// src/materials/elastic_bond_based.cxx:55-130 computeInternalForceElasticBondBased
// (owner-side gather plus neighbour-side scatter; f zeroed first as in
// src/materials/Peridigm_ElasticBondBasedMaterial.cpp:119). planar != 0: the
// plane strain / plane stress constant with thickness `height`.
void force_pmb(int numOwned, const double* x, const double* y, const double* vol,
               const double* bondDamage, double* f, const int* nlist, double K, double horizon,
               int planar, double height) {
  const double pi = 3.14159265358979323846;
  double constant = 0.0;
  if (planar)
    constant = 12.0 * K / (pi * height * horizon * horizon * horizon);
  else
    constant = 18.0 * K / (pi * horizon * horizon * horizon * horizon);
  int li = 0, b = 0;
  for (int p = 0; p < numOwned; p++) {
    const double X0 = x[p * 3], X1 = x[p * 3 + 1], X2 = x[p * 3 + 2];
    const double Y0 = y[p * 3], Y1 = y[p * 3 + 1], Y2 = y[p * 3 + 2];
    const double volume = vol[p];
    const int nn = nlist[li++];
    for (int n = 0; n < nn; n++) {
      const int j = nlist[li++];
      const double nX0 = x[j * 3], nX1 = x[j * 3 + 1], nX2 = x[j * 3 + 2];
      const double nY0 = y[j * 3], nY1 = y[j * 3 + 1], nY2 = y[j * 3 + 2];
      const double neighborVolume = vol[j];
      const double init = std::sqrt((nX0 - X0) * (nX0 - X0) + (nX1 - X1) * (nX1 - X1) +
                                    (nX2 - X2) * (nX2 - X2));
      const double cur = std::sqrt((nY0 - Y0) * (nY0 - Y0) + (nY1 - Y1) * (nY1 - Y1) +
                                   (nY2 - Y2) * (nY2 - Y2));
      const double stretch = (cur - init) / init;
      const double dmg = bondDamage[b++];
      const double t = 0.5 * (1.0 - dmg) * stretch * constant;
      const double fx = t * (nY0 - Y0) / cur, fy = t * (nY1 - Y1) / cur, fz = t * (nY2 - Y2) / cur;
      f[3 * p + 0] += fx * neighborVolume;
      f[3 * p + 1] += fy * neighborVolume;
      f[3 * p + 2] += fz * neighborVolume;
      f[3 * j + 0] -= fx * volume;
      f[3 * j + 1] -= fy * volume;
      f[3 * j + 2] -= fz * volume;
    }
  }
}

// src/damage/Peridigm_CriticalStretchDamageModel.cpp:108-184 computeDamage
// (bondDamageNP1 must hold bondDamageN on entry, as at :131).
void damage_impl(int numOwned, const double* x, const double* y,
                 double* bdNP1, double* damage, const int* nlist, double sc,
                 long long* near_count, double near_rel, const double* sc_node = nullptr) {
  int li = 0;
  long long b = 0;
  for (int i = 0; i < numOwned; ++i) {
    double X0 = x[3 * i], X1 = x[3 * i + 1], X2 = x[3 * i + 2];
    double Y0 = y[3 * i], Y1 = y[3 * i + 1], Y2 = y[3 * i + 2];
    const double sci = sc_node ? sc_node[i] : sc;
    int nn = nlist[li++];
    for (int k = 0; k < nn; ++k) {
      int j = nlist[li++];
      double init = std::sqrt((X0 - x[3 * j]) * (X0 - x[3 * j]) +
                              (X1 - x[3 * j + 1]) * (X1 - x[3 * j + 1]) +
                              (X2 - x[3 * j + 2]) * (X2 - x[3 * j + 2]));
      double cur = std::sqrt((Y0 - y[3 * j]) * (Y0 - y[3 * j]) +
                             (Y1 - y[3 * j + 1]) * (Y1 - y[3 * j + 1]) +
                             (Y2 - y[3 * j + 2]) * (Y2 - y[3 * j + 2]));
      double rel = (cur - init) / init;
      double trial = 0.0;
      if (rel > sci) trial = 1.0;
      if (near_count && bdNP1[b] == 0.0 && std::fabs(rel - sci) <= near_rel * sci)
        (*near_count)++;
      if (trial > bdNP1[b]) bdNP1[b] = trial;
      b++;
    }
  }
  li = 0;
  b = 0;
  for (int i = 0; i < numOwned; ++i) {
    int nn = nlist[li++];
    li += nn;
    double tot = 0.0;
    for (int k = 0; k < nn; ++k) tot += bdNP1[b++];
    if (nn > 0) tot /= nn; else tot = 0.0;
    damage[i] = tot;
  }
}

long long list_bonds(int N, const int* nlist) {
  long long li = 0, B = 0;
  for (int i = 0; i < N; i++) { int nn = nlist[li]; li += nn + 1; B += nn; }
  return B;
}

// Prescribed-displacement boundary conditions value(t) = c1*t + c0 per node
// set and coordinate, applied to the velocity as in
// src/core/Peridigm_BoundaryCondition.cpp:202-241 (coeff 0, deltaTCoeff 1)
// including the zero-increment rule of :136-143.
struct BCs {
  int n;
  const int* ptr;
  const int* nodes;
  const int* coord;
  const double* c0;
  const double* c1;
};

void apply_velocity_bc(const BCs& bc, double tc, double tp, const double* u,
                       double* v) {
  if (tc == tp) return;
  for (int k = 0; k < bc.n; k++) {
    for (int q = bc.ptr[k]; q < bc.ptr[k + 1]; q++) {
      int node = bc.nodes[q];
      int c = bc.coord[k];
      double cur = bc.c1[k] * tc + bc.c0[k];
      double prev = bc.c1[k] * tp + bc.c0[k];
      if (cur - prev == 0.0) prev = u[3 * node + c];
      double value = 0.0 * (cur - prev) + 1.0 * (cur - prev) * (1.0 / (tc - tp));
      v[3 * node + c] = value;
    }
  }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Neighbour lists
// ---------------------------------------------------------------------------

// src/io/mesh_input/quick_grid/QuickGrid.cxx:243-253 (cell centres by
// accumulation p += cellSize), :394-408 (candidate window), :160-165
// (SphericalNormFunction, strict '<'), :510-647 (kk->jj->ii enumeration,
// ascending global id, self skipped). Single process: local ids == gids.
// Returns the flattened list length; writes outputs only if cap suffices.
long long orc_quickgrid(int nx, int ny, int nz, double x0, double y0, double z0,
                        double Lx, double Ly, double Lz, double horizon,
                        int spherical, double* xout, double* vout, int* nlist,
                        long long cap) {
  const int n[3] = {nx, ny, nz};
  const double o[3] = {x0, y0, z0};
  const double L[3] = {Lx, Ly, Lz};
  std::vector<double> c[3];
  size_t hw[3];
  for (int d = 0; d < 3; d++) {
    double cs = L[d] / n[d];
    c[d].resize(n[d]);
    double p = o[d] + cs / 2.0;
    for (int i = 0; i < n[d]; p += cs, i++) c[d][i] = p;
    double dCx = horizon / (cs * 1.0);
    size_t nC = static_cast<size_t>(dCx);
    if (std::abs(dCx - nC) >= .5) nC += 1;
    if (nC > (size_t)n[d]) nC = n[d];
    hw[d] = nC;
  }
  auto start = [](size_t i, size_t h) { return i > h ? i - h : 0; };
  auto end = [](size_t i, size_t h, size_t nn) { return i < (nn - h) ? i + h + 1 : nn; };
  double cellVolume = (Lx / nx) * (Ly / ny) * (Lz / nz);
  long long li = 0;
  bool fill = nlist != nullptr;
  for (size_t k = 0; k < (size_t)nz; k++)
    for (size_t j = 0; j < (size_t)ny; j++)
      for (size_t i = 0; i < (size_t)nx; i++) {
        long long gid = (long long)i + (long long)j * nx + (long long)k * nx * ny;
        double pt[3] = {c[0][i], c[1][j], c[2][k]};
        if (fill && xout) { xout[3 * gid] = pt[0]; xout[3 * gid + 1] = pt[1]; xout[3 * gid + 2] = pt[2]; }
        if (fill && vout) vout[gid] = cellVolume;
        long long head = li++;
        int cnt = 0;
        for (size_t kk = start(k, hw[2]); kk < end(k, hw[2], nz); kk++)
          for (size_t jj = start(j, hw[1]); jj < end(j, hw[1], ny); jj++)
            for (size_t ii = start(i, hw[0]); ii < end(i, hw[0], nx); ii++) {
              if (ii == i && jj == j && kk == k) continue;
              double pn[3] = {c[0][ii], c[1][jj], c[2][kk]};
              bool in = true;
              if (spherical) {
                double dx = pt[0] - pn[0];
                double dy = pt[1] - pn[1];
                double dz = pt[2] - pn[2];
                in = dx * dx + dy * dy + dz * dz - horizon * horizon < 0.0;
              }
              if (in) {
                if (fill && li < cap) nlist[li] = (int)(ii + jj * nx + kk * nx * ny);
                li++;
                cnt++;
              }
            }
        if (fill && head < cap) nlist[head] = cnt;
      }
  return li;
}

// Fixed-radius search with the Zoltan search-tree criterion
// dx*dx+dy*dy+dz*dz <= R*R, dx = x_candidate - x_point
// (src/io/Peridigm_ZoltanSearchTree.cpp:236-246); rows sorted by id
// (src/io/pdneigh/NeighborhoodList.cxx:705). A uniform cell grid of cell size
// >= R supplies candidates; the predicate alone decides membership.
long long orc_radius_search(int N, const double* x, double R, int include_self,
                            int* nlist, long long cap) {
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = 0; i < N; i++)
    for (int d = 0; d < 3; d++) { lo[d] = std::min(lo[d], x[3 * i + d]); hi[d] = std::max(hi[d], x[3 * i + d]); }
  double cs = R > 0 ? R * 1.0000001 : 1.0;
  int nc[3];
  for (int d = 0; d < 3; d++) nc[d] = std::max(1, (int)std::floor((hi[d] - lo[d]) / cs) + 1);
  std::vector<long long> cell(N);
  for (int i = 0; i < N; i++) {
    int ci[3];
    for (int d = 0; d < 3; d++) ci[d] = std::min(nc[d] - 1, (int)std::floor((x[3 * i + d] - lo[d]) / cs));
    cell[i] = ((long long)ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
  }
  std::vector<int> order(N);
  for (int i = 0; i < N; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return cell[a] < cell[b] || (cell[a] == cell[b] && a < b); });
  long long ncell = (long long)nc[0] * nc[1] * nc[2];
  std::vector<int> cstart(ncell + 1, 0);
  for (int i = 0; i < N; i++) cstart[cell[i] + 1]++;
  for (long long c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  double R2 = R * R;
  long long li = 0;
  std::vector<int> row;
  for (int i = 0; i < N; i++) {
    row.clear();
    long long c = cell[i];
    int ci[3] = {(int)(c % nc[0]), (int)((c / nc[0]) % nc[1]), (int)(c / ((long long)nc[0] * nc[1]))};
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int a = ci[0] + dx, b2 = ci[1] + dy, cz = ci[2] + dz;
          if (a < 0 || b2 < 0 || cz < 0 || a >= nc[0] || b2 >= nc[1] || cz >= nc[2]) continue;
          long long cc = ((long long)cz * nc[1] + b2) * nc[0] + a;
          for (int q = cstart[cc]; q < cstart[cc + 1]; q++) {
            int j = order[q];
            if (j == i && !include_self) continue;
            double ddx = (x[3 * j] - x[3 * i]);
            double ddy = (x[3 * j + 1] - x[3 * i + 1]);
            double ddz = (x[3 * j + 2] - x[3 * i + 2]);
            if (ddx * ddx + ddy * ddy + ddz * ddz <= R2) row.push_back(j);
          }
        }
    std::sort(row.begin(), row.end());
    if (nlist && li < cap) nlist[li] = (int)row.size();
    li++;
    for (size_t q = 0; q < row.size(); q++) { if (nlist && li < cap) nlist[li] = row[q]; li++; }
  }
  return li;
}

// Brute-force O(N^2) version of the same criterion (pins the cell grid).
long long orc_radius_search_brute(int N, const double* x, double R, int include_self,
                                  int* nlist, long long cap) {
  double R2 = R * R;
  long long li = 0;
  for (int i = 0; i < N; i++) {
    long long head = li++;
    int cnt = 0;
    for (int j = 0; j < N; j++) {
      if (j == i && !include_self) continue;
      double ddx = (x[3 * j] - x[3 * i]);
      double ddy = (x[3 * j + 1] - x[3 * i + 1]);
      double ddz = (x[3 * j + 2] - x[3 * i + 2]);
      if (ddx * ddx + ddy * ddy + ddz * ddz <= R2) { if (nlist && li < cap) nlist[li] = j; li++; cnt++; }
    }
    if (nlist && head < cap) nlist[head] = cnt;
  }
  return li;
}

// FinitePlane / FinitePlaneFilter (src/io/pdneigh/BondFilter.cxx:59-106,
// :123-153; Vector3D helpers src/io/utilities/Vector3D.h). plane[] =
// normal(3), lowerLeftCorner(3), bottomUnitVector(3), lengthBottom, lengthA.
// Returns 1 if the bond p0->p1 is cut (excluded from the list).
int orc_plane_cuts(const double* plane, const double* p0, const double* p1) {
  const double* n = plane;
  const double* r0 = plane + 3;
  const double* ub = plane + 6;
  double b = plane[9], a = plane[10];
  // ua = cross(ub, n), Vector3D.h:121-126
  double ua[3];
  ua[0] = -ub[2] * n[1] + ub[1] * n[2];
  ua[1] = ub[2] * n[0] - ub[0] * n[2];
  ua[2] = -ub[1] * n[0] + ub[0] * n[1];
  double numerator = (r0[0] - p0[0]) * n[0] + (r0[1] - p0[1]) * n[1] + (r0[2] - p0[2]) * n[2];
  double denominator = (p1[0] - p0[0]) * n[0] + (p1[1] - p0[1]) * n[1] + (p1[2] - p0[2]) * n[2];
  double t;
  if (std::abs(denominator) < 1.0e-14) t = DBL_MAX;
  else t = numerator / denominator;
  int intersection = (t >= 0.0 && t <= 1.0) ? 1 : 0;
  double xi[3];
  xi[0] = p0[0] + t * (p1[0] - p0[0]);
  xi[1] = p0[1] + t * (p1[1] - p0[1]);
  xi[2] = p0[2] + t * (p1[2] - p0[2]);
  if (!intersection) return 0;
  double zero = 1.0e-15;
  double one = 1.0 + zero;
  double dr[3] = {xi[0] - r0[0], xi[1] - r0[1], xi[2] - r0[2]};
  double aa = dr[0] * ua[0] + dr[1] * ua[1] + dr[2] * ua[2];
  double bb = dr[0] * ub[0] + dr[1] * ub[1] + dr[2] * ub[2];
  return (-zero < aa && aa / a < one && -zero < bb && bb / b < one) ? 1 : 0;
}

// Removes from a flattened list every bond cut by any of nplanes planes,
// keeping order (src/io/pdneigh/NeighborhoodList.cxx:700-740).
long long orc_filter_planes(int N, const double* x, const int* nlist_in, int nplanes,
                            const double* planes, int* nlist_out, long long cap) {
  long long li = 0, lo = 0;
  for (int i = 0; i < N; i++) {
    int nn = nlist_in[li++];
    long long head = lo++;
    int cnt = 0;
    for (int k = 0; k < nn; k++) {
      int j = nlist_in[li++];
      bool cut = false;
      for (int p = 0; p < nplanes && !cut; p++) cut = orc_plane_cuts(planes + 11 * p, x + 3 * i, x + 3 * j) != 0;
      if (!cut) { if (nlist_out && lo < cap) nlist_out[lo] = j; lo++; cnt++; }
    }
    if (nlist_out && head < cap) nlist_out[head] = cnt;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Material / damage kernels
// ---------------------------------------------------------------------------

// src/materials/material_utilities.cxx:161-188 (per point), :501-520 (driver)
void orc_weighted_volume(int N, const double* x, const double* vol, const int* nlist,
                         double horizon, int infl, double* m) {
  const int* neighPtr = nlist;
  for (int p = 0; p < N; p++) {
    const double* X = x + 3 * p;
    int numNeigh = *neighPtr;
    const int* q = neighPtr + 1;
    double mm = 0.0;
    for (int n = 0; n < numNeigh; n++, q++) {
      int j = *q;
      double cellVolume = vol[j];
      const double* XP = x + 3 * j;
      double dx = XP[0] - X[0];
      double dy = XP[1] - X[1];
      double dz = XP[2] - X[2];
      double zetaSquared = dx * dx + dy * dy + dz * dz;
      double d = std::sqrt(zetaSquared);
      mm += omega_fn(infl, d, horizon) * (zetaSquared) * cellVolume;
    }
    m[p] = mm;
    neighPtr += (numNeigh + 1);
  }
}

void orc_dilatation(int N, const double* x, const double* y, const double* m,
                    const double* vol, const double* bondDamage, const int* nlist,
                    double horizon, int infl, double* theta) {
  dilatation_t<double>(N, x, y, m, vol, bondDamage, theta, nlist, horizon, infl);
}

// ElasticMaterial::computeForce (src/materials/Peridigm_ElasticMaterial.cpp:141-172):
// zero f, dilatation, then the scatter force.
void orc_compute_force(int N, const double* x, const double* y, const double* m,
                       const double* vol, const double* bondDamage, const int* nlist,
                       double K, double mu, double horizon, int infl,
                       double* theta, double* f) {
  for (int i = 0; i < 3 * N; i++) f[i] = 0.0;
  dilatation_t<double>(N, x, y, m, vol, bondDamage, theta, nlist, horizon, infl);
  force_t<double>(N, x, y, m, vol, theta, bondDamage, f, nlist, K, mu, horizon, infl);
}

// Scatter force with a caller-supplied dilatation (for isolating the force
// loop in tests).
void orc_force_given_theta(int N, const double* x, const double* y, const double* m,
                           const double* vol, const double* theta, const double* bondDamage,
                           const int* nlist, double K, double mu, double horizon, int infl,
                           double* f) {
  for (int i = 0; i < 3 * N; i++) f[i] = 0.0;
  force_t<double>(N, x, y, m, vol, theta, bondDamage, f, nlist, K, mu, horizon, infl);
}

// CriticalStretchDamageModel::computeDamage. bdN is copied to bdNP1 first
// (:131). near_count (optional) counts intact bonds with
// |s - sc| <= near_rel*sc.
void orc_compute_damage(int N, const double* x, const double* y, const double* bdN,
                        double* bdNP1, double* damage, const int* nlist, double sc,
                        long long* near_count, double near_rel) {
  long long B = list_bonds(N, nlist);
  if (bdNP1 != bdN) std::memcpy(bdNP1, bdN, sizeof(double) * B);
  if (near_count) *near_count = 0;
  damage_impl(N, x, y, bdNP1, damage, nlist, sc, near_count, near_rel);
}

// Same call with one critical stretch per node (a bond i->j uses node i's
// value): one damage-model block per specimen when several specimens share
// one list.
void orc_compute_damage_sc(int N, const double* x, const double* y, const double* bdN,
                           double* bdNP1, double* damage, const int* nlist,
                           const double* sc_node, long long* near_count, double near_rel) {
  long long B = list_bonds(N, nlist);
  if (bdNP1 != bdN) std::memcpy(bdNP1, bdN, sizeof(double) * B);
  if (near_count) *near_count = 0;
  damage_impl(N, x, y, bdNP1, damage, nlist, 0.0, near_count, near_rel, sc_node);
}

// FP64 bond stretch in list order, (|y_j-y_i| - |x_j-x_i|)/|x_j-x_i| exactly as
// the damage model forms it (used by tests to classify near-critical bonds).
void orc_bond_stretch(int N, const double* x, const double* y, const int* nlist, double* s) {
  int li = 0;
  long long b = 0;
  for (int i = 0; i < N; ++i) {
    int nn = nlist[li++];
    for (int k = 0; k < nn; ++k) {
      int j = nlist[li++];
      double init = std::sqrt((x[3 * i] - x[3 * j]) * (x[3 * i] - x[3 * j]) +
                              (x[3 * i + 1] - x[3 * j + 1]) * (x[3 * i + 1] - x[3 * j + 1]) +
                              (x[3 * i + 2] - x[3 * j + 2]) * (x[3 * i + 2] - x[3 * j + 2]));
      double cur = std::sqrt((y[3 * i] - y[3 * j]) * (y[3 * i] - y[3 * j]) +
                             (y[3 * i + 1] - y[3 * j + 1]) * (y[3 * i + 1] - y[3 * j + 1]) +
                             (y[3 * i + 2] - y[3 * j + 2]) * (y[3 * i + 2] - y[3 * j + 2]));
      s[b++] = (cur - init) / init;
    }
  }
}

// src/core/Peridigm_CriticalTimeStep.cpp:54-126 (constant horizon).
double orc_critical_time_step(int N, const double* x, const double* vol, const int* nlist,
                              double density, double K, double horizon) {
  const double pi = 3.14159265358979323846;
  double springConstant = 18.0 * K / (pi * horizon * horizon * horizon * horizon);
  double minDt = 1.0e50;
  int li = 0;
  for (int i = 0; i < N; ++i) {
    double den = 0.0;
    double X[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    int nn = nlist[li++];
    for (int k = 0; k < nn; ++k) {
      int j = nlist[li++];
      double d = std::sqrt((X[0] - x[3 * j]) * (X[0] - x[3 * j]) +
                           (X[1] - x[3 * j + 1]) * (X[1] - x[3 * j + 1]) +
                           (X[2] - x[3 * j + 2]) * (X[2] - x[3 * j + 2]));
      den += vol[j] * springConstant / d;
    }
    double dt = 1.0e50;
    if (nn > 0) dt = std::sqrt(2.0 * density / den);
    if (dt < minDt) minDt = dt;
  }
  return minDt;
}

// ---------------------------------------------------------------------------
// Explicit velocity-Verlet (src/core/Peridigm.cpp:1388-1590)
// ---------------------------------------------------------------------------
// u, v are initial conditions on entry and final state on exit; a, theta,
// force (density), damage, m receive the final values. bond_damage (B, list
// order) is the initial bond damage on entry and the final one on exit.
// sc <= 0 disables the damage model. rho is per node.
int orc_explicit_run_sc(int N, const double* x, const double* vol, const double* rho,
                        const int* nlist, double K, double mu, double horizon, int infl,
                        double sc, const double* sc_node, double dt, double t0, int nsteps,
                        int nbc, const int* bc_ptr, const int* bc_nodes, const int* bc_coord,
                        const double* bc_c0, const double* bc_c1, double* u, double* v,
                        double* a, double* bond_damage, double* damage, double* theta,
                        double* force, double* m);

int orc_explicit_run(int N, const double* x, const double* vol, const double* rho,
                     const int* nlist, double K, double mu, double horizon, int infl,
                     double sc, double dt, double t0, int nsteps, int nbc,
                     const int* bc_ptr, const int* bc_nodes, const int* bc_coord,
                     const double* bc_c0, const double* bc_c1, double* u, double* v,
                     double* a, double* bond_damage, double* damage, double* theta,
                     double* force, double* m) {
  return orc_explicit_run_sc(N, x, vol, rho, nlist, K, mu, horizon, infl, sc, nullptr, dt, t0,
                             nsteps, nbc, bc_ptr, bc_nodes, bc_coord, bc_c0, bc_c1, u, v, a,
                             bond_damage, damage, theta, force, m);
}

// Same run with a critical stretch per node (one damage-model block per
// specimen when several specimens share one list); sc > 0 enables damage.
int orc_explicit_run_sc(int N, const double* x, const double* vol, const double* rho,
                        const int* nlist, double K, double mu, double horizon, int infl,
                        double sc, const double* sc_node, double dt, double t0, int nsteps,
                        int nbc, const int* bc_ptr, const int* bc_nodes, const int* bc_coord,
                        const double* bc_c0, const double* bc_c1, double* u, double* v,
                        double* a, double* bond_damage, double* damage, double* theta,
                        double* force, double* m) {
  BCs bc{nbc, bc_ptr, bc_nodes, bc_coord, bc_c0, bc_c1};
  std::vector<double> y(3 * N);
  orc_weighted_volume(N, x, vol, nlist, horizon, infl, m);
  for (int i = 0; i < 3 * N; i++) y[i] = x[i] + u[i];
  for (int i = 0; i < N; i++) damage[i] = 0.0;
  auto eval = [&]() {
    if (sc > 0.0) damage_impl(N, x, y.data(), bond_damage, damage, nlist, sc, nullptr, 0.0, sc_node);
    orc_compute_force(N, x, y.data(), m, vol, bond_damage, nlist, K, mu, horizon, infl, theta, force);
  };
  eval();
  for (int i = 0; i < 3 * N; i++) { a[i] = force[i]; a[i] += 0.0; a[i] /= rho[i / 3]; }
  double dt2 = dt / 2.0;
  double tc = t0, tp = t0;
  for (int step = 1; step <= nsteps; step++) {
    tp = tc;
    tc = t0 + (step * dt);
    for (int i = 0; i < 3 * N; i++) v[i] = v[i] + dt2 * a[i];
    apply_velocity_bc(bc, tc, tp, u, v);
    for (int i = 0; i < 3 * N; i++) y[i] = x[i] + u[i] + dt * v[i];
    for (int i = 0; i < 3 * N; i++) u[i] = u[i] + dt * v[i];
    eval();
    for (int i = 0; i < 3 * N; i++) {
      if (!std::isfinite(force[i])) return -1;
    }
    for (int i = 0; i < 3 * N; i++) { a[i] = force[i]; a[i] += 0.0; a[i] /= rho[i / 3]; }
    for (int i = 0; i < 3 * N; i++) v[i] = v[i] + dt2 * a[i];
  }
  return 0;
}

// Bond-based force density (theta is not part of the model).
void orc_compute_force_pmb(int N, const double* x, const double* y, const double* vol,
                           const double* bondDamage, const int* nlist, double K, double horizon,
                           int planar, double height, double* f) {
  for (int i = 0; i < 3 * N; i++) f[i] = 0.0;
  force_pmb(N, x, y, vol, bondDamage, f, nlist, K, horizon, planar, height);
}

// ElasticMaterial::computeForce with thermal strains (deltaT per point, may be
// null) and the partial stress (ps, 9 per point, may be null).
void orc_compute_force_thermal(int N, const double* x, const double* y, const double* m,
                               const double* vol, const double* bondDamage, const int* nlist,
                               double K, double mu, double horizon, int infl, double alphaT,
                               const double* deltaT, double* theta, double* f, double* ps) {
  for (int i = 0; i < 3 * N; i++) f[i] = 0.0;
  if (ps)
    for (int i = 0; i < 9 * N; i++) ps[i] = 0.0;
  dilatation_t<double>(N, x, y, m, vol, bondDamage, theta, nlist, horizon, infl, deltaT, alphaT);
  force_t<double>(N, x, y, m, vol, theta, bondDamage, f, nlist, K, mu, horizon, infl, deltaT,
                  alphaT, ps);
}

// Linearized LPS without partial volumes ("Linear LPS Partial Volume",
// src/materials/Peridigm_LinearLPSPVMaterial.cpp:229-289 precompute and
// :296-394 computeForce, the influence function value of each bond taken
// from the influence function at its reference length as initialize stores
// it): theta_i = 3/m_i sum_j w (1-d) xi.(u_j - u_i) V_j, and every bond i->j
// adds (1-d)(T1 xi + T2 (xi xi^T)(u_j - u_i)) with
// T1 = (9K - 15mu) w theta_i / (3 m_i), T2 = 15 mu w / (m_i |xi|^2) to f_i
// times V_j and subtracts it from f_j times V_i, in list order.
void orc_compute_force_linear(int N, const double* x, const double* y, const double* m,
                              const double* vol, const double* bondDamage, const int* nlist,
                              double K, double mu, double horizon, int infl, double* theta,
                              double* f) {
  long long li = 0, bi = 0;
  for (int p = 0; p < N; p++) {
    theta[p] = 0.0;
    const int nn = nlist[li++];
    const double* xs = x + 3 * p;
    const double* ys = y + 3 * p;
    for (int k = 0; k < nn; k++, bi++) {
      const int j = nlist[li++];
      const double* xn = x + 3 * j;
      const double* yn = y + 3 * j;
      double zeta[3], u[3], un[3];
      for (int c = 0; c < 3; c++) {
        zeta[c] = xn[c] - xs[c];
        u[c] = ys[c] - xs[c];
        un[c] = yn[c] - xn[c];
      }
      const double nz = std::sqrt(zeta[0] * zeta[0] + zeta[1] * zeta[1] + zeta[2] * zeta[2]);
      const double w = omega_fn(infl, nz, horizon);
      const double dot = zeta[0] * (un[0] - u[0]) + zeta[1] * (un[1] - u[1]) + zeta[2] * (un[2] - u[2]);
      theta[p] += w * (1.0 - bondDamage[bi]) * dot * vol[j];
    }
    if (nn > 0) theta[p] *= 3.0 / m[p];
  }
  for (int i = 0; i < 3 * N; i++) f[i] = 0.0;
  li = 0;
  bi = 0;
  for (int p = 0; p < N; p++) {
    const int nn = nlist[li++];
    const double* xs = x + 3 * p;
    const double* ys = y + 3 * p;
    for (int k = 0; k < nn; k++, bi++) {
      const int j = nlist[li++];
      const double* xn = x + 3 * j;
      const double* yn = y + 3 * j;
      double zeta[3], u[3], un[3];
      for (int c = 0; c < 3; c++) {
        zeta[c] = xn[c] - xs[c];
        u[c] = ys[c] - xs[c];
        un[c] = yn[c] - xn[c];
      }
      const double z2 = zeta[0] * zeta[0] + zeta[1] * zeta[1] + zeta[2] * zeta[2];
      const double w = omega_fn(infl, std::sqrt(z2), horizon);
      const double t1 = (9.0 * K - 15.0 * mu) * w * theta[p] / (3.0 * m[p]);
      const double t2 = 15.0 * mu * w / (m[p] * z2);
      double mv[3];
      for (int r = 0; r < 3; r++)
        mv[r] = zeta[r] * zeta[0] * (un[0] - u[0]) + zeta[r] * zeta[1] * (un[1] - u[1]) +
                zeta[r] * zeta[2] * (un[2] - u[2]);
      const double dmg = 1.0 - bondDamage[bi];
      for (int c = 0; c < 3; c++) {
        const double fc = dmg * (t1 * zeta[c] + t2 * mv[c]);
        f[3 * p + c] += fc * vol[j];
        f[3 * j + c] -= fc * vol[p];
      }
    }
  }
}

// Explicit run as orc_explicit_run_sc with a material model (0: elastic
// LPS, 1: elastic bond based) and optionally one critical stretch per force
// evaluation (sc_eval[k] for the evaluation that ends step k, k = 0..nsteps:
// "Time Dependent Critical Stretch" evaluated at every step,
// src/core/Peridigm.cpp:1447-1463); sc > 0 enables damage.
int orc_contact_search(int N, const double* y, double R, const int* nlist, int* out,
                       long long cap, long long* size);
void orc_contact_force(int N, const double* y, const double* v, const double* vol,
                       const int* clist, double rc, double ks, double muf, double horizon,
                       double* f);

// With contact (crc > 0; src/core/Peridigm.cpp:1468-1472, 1515-1524,
// 1559-1565): the contact list is rebuilt from the configuration at the start
// of every step whose number is a multiple of cfreq (and at step 0), and
// every evaluation adds the contact force density (velocity after the first
// half kick) to the internal one; fcontact receives the last evaluation's.
int orc_explicit_run_ext(int N, const double* x, const double* vol, const double* rho,
                         const int* nlist, int material, int planar, double height, double K,
                         double mu, double horizon, int infl, double sc, const double* sc_eval,
                         double dt, double t0, int nsteps, int nbc, const int* bc_ptr,
                         const int* bc_nodes, const int* bc_coord, const double* bc_c0,
                         const double* bc_c1, double csearch, int cfreq, double crc, double cks,
                         double cmu, double* u, double* v, double* a, double* bond_damage,
                         double* damage, double* theta, double* force, double* fcontact,
                         double* m) {
  BCs bc{nbc, bc_ptr, bc_nodes, bc_coord, bc_c0, bc_c1};
  std::vector<double> y(3 * N);
  orc_weighted_volume(N, x, vol, nlist, horizon, infl, m);
  for (int i = 0; i < 3 * N; i++) y[i] = x[i] + u[i];
  for (int i = 0; i < N; i++) damage[i] = 0.0;
  const bool contact = crc > 0.0;
  std::vector<int> clist;
  auto search = [&]() {
    long long sz = 0;
    orc_contact_search(N, y.data(), csearch, nlist, nullptr, 0, &sz);
    clist.assign((size_t)sz, 0);
    orc_contact_search(N, y.data(), csearch, nlist, clist.data(), sz, &sz);
  };
  auto eval = [&](int k) {
    if (sc > 0.0)
      damage_impl(N, x, y.data(), bond_damage, damage, nlist, sc_eval ? sc_eval[k] : sc, nullptr,
                  0.0, nullptr);
    if (material == 1) {
      for (int i = 0; i < N; i++) theta[i] = 0.0;
      orc_compute_force_pmb(N, x, y.data(), vol, bond_damage, nlist, K, horizon, planar, height,
                            force);
    } else {
      orc_compute_force(N, x, y.data(), m, vol, bond_damage, nlist, K, mu, horizon, infl, theta,
                        force);
    }
    if (contact) {
      orc_contact_force(N, y.data(), v, vol, clist.data(), crc, cks, cmu, horizon, fcontact);
      for (int i = 0; i < 3 * N; i++) force[i] += fcontact[i];
    }
  };
  if (contact) search();
  eval(0);
  for (int i = 0; i < 3 * N; i++) { a[i] = force[i]; a[i] += 0.0; a[i] /= rho[i / 3]; }
  double dt2 = dt / 2.0;
  double tc = t0, tp = t0;
  for (int step = 1; step <= nsteps; step++) {
    if (contact && cfreq > 0 && step % cfreq == 0) search();
    tp = tc;
    tc = t0 + (step * dt);
    for (int i = 0; i < 3 * N; i++) v[i] = v[i] + dt2 * a[i];
    apply_velocity_bc(bc, tc, tp, u, v);
    for (int i = 0; i < 3 * N; i++) y[i] = x[i] + u[i] + dt * v[i];
    for (int i = 0; i < 3 * N; i++) u[i] = u[i] + dt * v[i];
    eval(step);
    for (int i = 0; i < 3 * N; i++) {
      if (!std::isfinite(force[i])) return -1;
    }
    for (int i = 0; i < 3 * N; i++) { a[i] = force[i]; a[i] += 0.0; a[i] /= rho[i / 3]; }
    for (int i = 0; i < 3 * N; i++) v[i] = v[i] + dt2 * a[i];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Short-range contact
// ---------------------------------------------------------------------------

// Contact search (src/core/Peridigm_ContactManager.cpp:700-750, the search
// itself the neighbourhood search of src/io/pdneigh/NeighborhoodList.cxx:652
// with the contact search radius): for every point, the points within
// searchRadius of it in the configuration y (dx^2 + dy^2 + dz^2 <= R^2), the
// point itself and its bonded neighbours (nlist, broken or not) excluded.
// Flat [n, j...] layout, neighbours in ascending index (brute force: test
// sizes). *size receives the list length; out may be null (size query).
int orc_contact_search(int N, const double* y, double R, const int* nlist, int* out,
                       long long cap, long long* size) {
  std::vector<std::vector<int>> bonded(N);
  long long p = 0;
  for (int i = 0; i < N; i++) {
    const int nn = nlist[p++];
    for (int k = 0; k < nn; k++) bonded[i].push_back(nlist[p++]);
    std::sort(bonded[i].begin(), bonded[i].end());
  }
  const double R2 = R * R;
  long long n = 0;
  std::vector<int> row;
  for (int i = 0; i < N; i++) {
    row.clear();
    for (int j = 0; j < N; j++) {
      if (j == i) continue;
      const double dx = y[3 * j] - y[3 * i], dy = y[3 * j + 1] - y[3 * i + 1],
                   dz = y[3 * j + 2] - y[3 * i + 2];
      if (dx * dx + dy * dy + dz * dz <= R2 &&
          !std::binary_search(bonded[i].begin(), bonded[i].end(), j))
        row.push_back(j);
    }
    if (out && n + 1 + (long long)row.size() <= cap) {
      out[n] = (int)row.size();
      for (size_t k = 0; k < row.size(); k++) out[n + 1 + k] = row[k];
    }
    n += 1 + (long long)row.size();
  }
  *size = n;
  return (out && n > cap) ? -1 : 0;
}

// ShortRangeForceContactModel::computeForce
// (src/contact/Peridigm_ShortRangeForceContactModel.cpp:98-258): the
// contact force density is zeroed, then every listed pair (i, j) closer than
// the contact radius adds its normal force (and, with a nonzero friction
// coefficient, its friction force) to both points, in list order, with the
// reference's expressions.
void orc_contact_force(int N, const double* y, const double* v, const double* vol,
                       const int* clist, double rc, double ks, double muf, double horizon,
                       double* f) {
  for (int i = 0; i < 3 * N; i++) f[i] = 0.0;
  const double pi = 3.14159265358979323846;
  const double rc2 = rc * rc;
  long long p = 0;
  for (int i = 0; i < N; i++) {
    const int nn = clist[p++];
    if (nn <= 0) continue;
    const double X0 = y[3 * i], X1 = y[3 * i + 1], X2 = y[3 * i + 2];
    const double V0 = v[3 * i], V1 = v[3 * i + 1], V2 = v[3 * i + 2];
    const double vi = vol[i];
    for (int k = 0; k < nn; k++) {
      const int j = clist[p++];
      const double d0 = y[3 * j] - X0, d1 = y[3 * j + 1] - X1, d2 = y[3 * j + 2] - X2;
      const double dsq = d0 * d0 + d1 * d1 + d2 * d2;
      if (!(dsq < rc2)) continue;
      const double d = std::sqrt(dsq);
      const double c = 9.0 * ks / (pi * horizon * horizon * horizon * horizon);
      const double temp = c * (rc - d) / horizon;
      const double vj = vol[j];
      if (muf != 0.0) {
        const double n0 = d0 / d, n1 = d1 / d, n2 = d2 / d;
        const double cn = V0 * n0 + V1 * n1 + V2 * n2;
        const double nbn = v[3 * j] * n0 + v[3 * j + 1] * n1 + v[3 * j + 2] * n2;
        const double cp0 = V0 - cn * n0, cp1 = V1 - cn * n1, cp2 = V2 - cn * n2;
        const double np0 = v[3 * j] - nbn * n0, np1 = v[3 * j + 1] - nbn * n1,
                     np2 = v[3 * j + 2] - nbn * n2;
        const double m0 = 0.5 * (cp0 + np0), m1 = 0.5 * (cp1 + np1), m2 = 0.5 * (cp2 + np2);
        const double cr0 = cp0 - m0, cr1 = cp1 - m1, cr2 = cp2 - m2;
        const double nr0 = np0 - m0, nr1 = np1 - m1, nr2 = np2 - m2;
        const double ncr = std::sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
        const double nnr = std::sqrt(nr0 * nr0 + nr1 * nr1 + nr2 * nr2);
        const double cf0 = -(temp * vj * d0 / d), cf1 = -(temp * vj * d1 / d),
                     cf2 = -(temp * vj * d2 / d);
        const double nf0 = (temp * vi * d0 / d), nf1 = (temp * vi * d1 / d),
                     nf2 = (temp * vi * d2 / d);
        const double ncf = std::sqrt(cf0 * cf0 + cf1 * cf1 + cf2 * cf2);
        const double nnf = std::sqrt(nf0 * nf0 + nf1 * nf1 + nf2 * nf2);
        double cfr0 = 0.0, cfr1 = 0.0, cfr2 = 0.0, nfr0 = 0.0, nfr1 = 0.0, nfr2 = 0.0;
        if (ncr != 0.0) {
          cfr0 = -muf * ncf * cr0 / ncr; cfr1 = -muf * ncf * cr1 / ncr; cfr2 = -muf * ncf * cr2 / ncr;
        }
        if (nnr != 0.0) {
          nfr0 = -muf * nnf * nr0 / nnr; nfr1 = -muf * nnf * nr1 / nnr; nfr2 = -muf * nnf * nr2 / nnr;
        }
        f[3 * i] += cf0 + cfr0; f[3 * i + 1] += cf1 + cfr1; f[3 * i + 2] += cf2 + cfr2;
        f[3 * j] += nf0 + nfr0; f[3 * j + 1] += nf1 + nfr1; f[3 * j + 2] += nf2 + nfr2;
      } else {
        f[3 * i] -= temp * vj * d0 / d;
        f[3 * i + 1] -= temp * vj * d1 / d;
        f[3 * i + 2] -= temp * vj * d2 / d;
        f[3 * j] += temp * vi * d0 / d;
        f[3 * j + 1] += temp * vi * d1 / d;
        f[3 * j + 2] += temp * vi * d2 / d;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Quasi-static residual and tangent (src/core/Peridigm.cpp:4176-4317)
// ---------------------------------------------------------------------------

// Internal force density f(u) with y = x + u and bond damage frozen.
void orc_force_at(int N, const double* x, const double* u, const double* m, const double* vol,
                  const double* bondDamage, const int* nlist, double K, double mu,
                  double horizon, int infl, double* f) {
  std::vector<double> y(3 * N), th(N);
  for (int i = 0; i < 3 * N; i++) y[i] = x[i] + u[i];
  orc_compute_force(N, x, y.data(), m, vol, bondDamage, nlist, K, mu, horizon, infl, th.data(), f);
}

// Directional derivative Jw = d f(u + eps w)/d eps at eps = 0, by forward-mode
// dual numbers through the same dilatation and scatter force loops (the
// counterpart of the Sacado tangent, applied to a vector).
void orc_force_directional_thermal(int N, const double* x, const double* u, const double* w,
                                   const double* m, const double* vol, const double* bondDamage,
                                   const int* nlist, double K, double mu, double horizon,
                                   int infl, double alphaT, const double* deltaT, double* f,
                                   double* Jw) {
  std::vector<Dual> y(3 * N), th(N), ff(3 * N);
  for (int i = 0; i < 3 * N; i++) y[i] = Dual(x[i] + u[i], w[i]);
  dilatation_t<Dual>(N, x, y.data(), m, vol, bondDamage, th.data(), nlist, horizon, infl, deltaT,
                     alphaT);
  force_t<Dual>(N, x, y.data(), m, vol, th.data(), bondDamage, ff.data(), nlist, K, mu, horizon,
                infl, deltaT, alphaT);
  for (int i = 0; i < 3 * N; i++) { if (f) f[i] = ff[i].v; Jw[i] = ff[i].d; }
}

void orc_force_directional(int N, const double* x, const double* u, const double* w,
                           const double* m, const double* vol, const double* bondDamage,
                           const int* nlist, double K, double mu, double horizon, int infl,
                           double* f, double* Jw) {
  orc_force_directional_thermal(N, x, u, w, m, vol, bondDamage, nlist, K, mu, horizon, infl,
                                0.0, nullptr, f, Jw);
}

// ---------------------------------------------------------------------------
// Message-passing network (FP64 reference for gradient checks)
// ---------------------------------------------------------------------------
// Layer l: h' = act(h W_self + (sum_j a_ij h_j) W_nbr + b), act = tanh for
// l < L-1 and identity for the last layer. Weights row-major [din][dout].
// feat: N x F0 input features. a_ij (edge weights, B) in CSR order
// (row_ptr N+1, col B). out: N x F_L. All layers store their activations in
// acts (sum_l N*F_l, including the input) for the backward pass.
void orc_gnn_forward(int N, const int* row_ptr, const int* col, const double* ew,
                     int L, const int* dims, const double* params, const double* feat,
                     double* acts) {
  int off_in = 0;
  std::memcpy(acts, feat, sizeof(double) * N * dims[0]);
  const double* P = params;
  for (int l = 0; l < L; l++) {
    int din = dims[l], dout = dims[l + 1];
    const double* Ws = P; P += din * dout;
    const double* Wn = P; P += din * dout;
    const double* bb = P; P += dout;
    const double* h = acts + off_in;
    double* ho = acts + off_in + (long long)N * din;
    std::vector<double> agg(din);
    for (int i = 0; i < N; i++) {
      std::fill(agg.begin(), agg.end(), 0.0);
      for (int e = row_ptr[i]; e < row_ptr[i + 1]; e++)
        for (int c = 0; c < din; c++) agg[c] += ew[e] * h[(long long)col[e] * din + c];
      for (int o = 0; o < dout; o++) {
        double s = bb[o];
        for (int c = 0; c < din; c++) s += h[(long long)i * din + c] * Ws[c * dout + o] + agg[c] * Wn[c * dout + o];
        ho[(long long)i * dout + o] = (l < L - 1) ? std::tanh(s) : s;
      }
    }
    off_in += N * din;
  }
}

// Backward of orc_gnn_forward for loss = sum(gout * out): writes dparams
// (same layout as params). Requires the acts written by orc_gnn_forward.
void orc_gnn_backward(int N, const int* row_ptr, const int* col, const double* ew,
                      int L, const int* dims, const double* params, const double* acts,
                      const double* gout, double* dparams) {
  std::vector<long long> act_off(L + 1), par_off(L + 1);
  act_off[0] = 0; par_off[0] = 0;
  for (int l = 0; l < L; l++) {
    act_off[l + 1] = act_off[l] + (long long)N * dims[l];
    par_off[l + 1] = par_off[l] + 2LL * dims[l] * dims[l + 1] + dims[l + 1];
  }
  std::vector<double> g(gout, gout + (long long)N * dims[L]);
  for (int l = L - 1; l >= 0; l--) {
    int din = dims[l], dout = dims[l + 1];
    const double* Ws = params + par_off[l];
    const double* Wn = Ws + din * dout;
    double* dWs = dparams + par_off[l];
    double* dWn = dWs + din * dout;
    double* db = dWn + din * dout;
    const double* h = acts + act_off[l];
    const double* ho = acts + act_off[l + 1];
    std::vector<double> gz(g);
    if (l < L - 1)
      for (long long q = 0; q < (long long)N * dout; q++) gz[q] = g[q] * (1.0 - ho[q] * ho[q]);
    for (int q = 0; q < din * dout; q++) { dWs[q] = 0.0; dWn[q] = 0.0; }
    for (int o = 0; o < dout; o++) db[o] = 0.0;
    std::vector<double> gh((long long)N * din, 0.0), gagg((long long)N * din, 0.0);
    std::vector<double> agg(din);
    for (int i = 0; i < N; i++) {
      std::fill(agg.begin(), agg.end(), 0.0);
      for (int e = row_ptr[i]; e < row_ptr[i + 1]; e++)
        for (int c = 0; c < din; c++) agg[c] += ew[e] * h[(long long)col[e] * din + c];
      for (int o = 0; o < dout; o++) {
        double go = gz[(long long)i * dout + o];
        db[o] += go;
        for (int c = 0; c < din; c++) {
          dWs[c * dout + o] += h[(long long)i * din + c] * go;
          dWn[c * dout + o] += agg[c] * go;
          gh[(long long)i * din + c] += Ws[c * dout + o] * go;
          gagg[(long long)i * din + c] += Wn[c * dout + o] * go;
        }
      }
    }
    // scatter gagg back through the aggregation: agg_i = sum_e ew_e h_col(e)
    for (int i = 0; i < N; i++)
      for (int e = row_ptr[i]; e < row_ptr[i + 1]; e++)
        for (int c = 0; c < din; c++) gh[(long long)col[e] * din + c] += ew[e] * gagg[(long long)i * din + c];
    g.swap(gh);
  }
}

}  // extern "C"
