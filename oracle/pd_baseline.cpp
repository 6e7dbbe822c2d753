// Multi-threaded FP64 CPU baseline of the explicit LPS + critical-stretch step.
//
// TEST INFRASTRUCTURE ONLY: the CPU figure bench.py reports beside the GPU
// (its cpu_baseline leg). It is kept apart from the pinned oracle functions in
// pd_oracle.cpp and is itself checked against them (tests/test_baseline.py).
//
// It runs the same loops as the reference, in the same order inside each
// point's neighbourhood, on all the host threads it is given (OpenMP):
//   - damage: src/damage/Peridigm_CriticalStretchDamageModel.cpp:131-183,
//     parallel over owned points (each point's bonds are its own);
//   - dilatation: src/materials/material_utilities.cxx:235-288, parallel over
//     owned points;
//   - force: src/materials/elastic.cxx:94-145, the owner-side gather and the
//     neighbour-side scatter, with each thread scattering into its own force
//     buffer over a contiguous range of owned points, then the buffers summed
//     (the shared-memory counterpart of an MPI rank's overlap vector and the
//     Export(Add) of src/core/Peridigm.cpp:1541-1546);
//   - velocity Verlet: src/core/Peridigm.cpp:1447-1590, parallel node loops.
// Only the force buffers' summation order differs from the serial oracle, so
// results agree to FP64 rounding.

#include <omp.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Mesh {
  int N;
  const double* x;
  const double* vol;
  const int* nlist;
  std::vector<long long> head;  // position of n_i in nlist
  std::vector<long long> boff;  // first bond of point i (list order)
};

Mesh make_mesh(int N, const double* x, const double* vol, const int* nlist) {
  Mesh M{N, x, vol, nlist, std::vector<long long>(N + 1), std::vector<long long>(N + 1)};
  long long li = 0, b = 0;
  for (int i = 0; i < N; i++) {
    M.head[i] = li;
    M.boff[i] = b;
    const int nn = nlist[li];
    li += nn + 1;
    b += nn;
  }
  M.head[N] = li;
  M.boff[N] = b;
  return M;
}

// src/damage/Peridigm_CriticalStretchDamageModel.cpp:131-183 (bdNP1 holds
// bdN on entry).
void damage(const Mesh& M, const double* y, double* bd, double* dmg, double sc, int nt) {
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int i = 0; i < M.N; ++i) {
    const double* x = M.x;
    const double X0 = x[3 * i], X1 = x[3 * i + 1], X2 = x[3 * i + 2];
    const double Y0 = y[3 * i], Y1 = y[3 * i + 1], Y2 = y[3 * i + 2];
    const int nn = M.nlist[M.head[i]];
    const int* nb = M.nlist + M.head[i] + 1;
    double* d = bd + M.boff[i];
    double tot = 0.0;
    for (int k = 0; k < nn; ++k) {
      const int j = nb[k];
      const double init = std::sqrt((X0 - x[3 * j]) * (X0 - x[3 * j]) +
                                    (X1 - x[3 * j + 1]) * (X1 - x[3 * j + 1]) +
                                    (X2 - x[3 * j + 2]) * (X2 - x[3 * j + 2]));
      const double cur = std::sqrt((Y0 - y[3 * j]) * (Y0 - y[3 * j]) +
                                   (Y1 - y[3 * j + 1]) * (Y1 - y[3 * j + 1]) +
                                   (Y2 - y[3 * j + 2]) * (Y2 - y[3 * j + 2]));
      const double rel = (cur - init) / init;
      double trial = 0.0;
      if (rel > sc) trial = 1.0;
      if (trial > d[k]) d[k] = trial;
    }
    for (int k = 0; k < nn; ++k) tot += d[k];
    dmg[i] = nn > 0 ? tot / nn : 0.0;
  }
}

// src/materials/material_utilities.cxx:235-288 ("One" influence function).
void dilatation(const Mesh& M, const double* y, const double* m, const double* bd, double* theta,
                int nt) {
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int p = 0; p < M.N; p++) {
    const double* x = M.x;
    const double* X = x + 3 * p;
    const double* Y = y + 3 * p;
    const int nn = M.nlist[M.head[p]];
    const int* nb = M.nlist + M.head[p] + 1;
    const double* d = bd + M.boff[p];
    double th = 0.0;
    for (int n = 0; n < nn; n++) {
      const int j = nb[n];
      const double* XP = x + 3 * j;
      const double* YP = y + 3 * j;
      const double X_dx = XP[0] - X[0], X_dy = XP[1] - X[1], X_dz = XP[2] - X[2];
      const double zetaSquared = X_dx * X_dx + X_dy * X_dy + X_dz * X_dz;
      const double Y_dx = YP[0] - Y[0], Y_dy = YP[1] - Y[1], Y_dz = YP[2] - Y[2];
      const double dY = Y_dx * Y_dx + Y_dy * Y_dy + Y_dz * Y_dz;
      const double zeta = std::sqrt(zetaSquared);
      double e = std::sqrt(dY);
      e = e - zeta;
      th += 3.0 * 1.0 * (1.0 - d[n]) * zeta * e * M.vol[j] / m[p];
    }
    theta[p] = th;
  }
}

// src/materials/elastic.cxx:94-145: thread t owns points [p0, p1) and
// scatters into its own buffer fb[t]; the buffers are then summed into f.
void force(const Mesh& M, const double* y, const double* m, const double* theta, const double* bd,
           double K, double mu, double* f, std::vector<std::vector<double>>& fb, int nt) {
  const int N = M.N;
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    std::vector<double>& g = fb[t];
    std::fill(g.begin(), g.end(), 0.0);
    const int p0 = (int)((long long)N * t / T), p1 = (int)((long long)N * (t + 1) / T);
    const double* x = M.x;
    for (int p = p0; p < p1; p++) {
      const double* X = x + 3 * p;
      const double* Y = y + 3 * p;
      const double alpha = 15.0 * mu / m[p];
      const double selfCellVolume = M.vol[p];
      const int nn = M.nlist[M.head[p]];
      const int* nb = M.nlist + M.head[p] + 1;
      const double* d = bd + M.boff[p];
      for (int n = 0; n < nn; n++) {
        const int j = nb[n];
        const double cellVolume = M.vol[j];
        const double* XP = x + 3 * j;
        const double* YP = y + 3 * j;
        const double X_dx = XP[0] - X[0], X_dy = XP[1] - X[1], X_dz = XP[2] - X[2];
        const double zeta = std::sqrt(X_dx * X_dx + X_dy * X_dy + X_dz * X_dz);
        const double Y_dx = YP[0] - Y[0], Y_dy = YP[1] - Y[1], Y_dz = YP[2] - Y[2];
        const double dY = std::sqrt(Y_dx * Y_dx + Y_dy * Y_dy + Y_dz * Y_dz);
        const double e = dY - zeta;
        const double c1 = 1.0 * theta[p] * (3.0 * K / m[p] - alpha / 3.0);
        const double t1 = (1.0 - d[n]) * (c1 * zeta + (1.0 - d[n]) * 1.0 * alpha * e);
        const double fx = t1 * Y_dx / dY, fy = t1 * Y_dy / dY, fz = t1 * Y_dz / dY;
        g[3 * p + 0] += fx * cellVolume;
        g[3 * p + 1] += fy * cellVolume;
        g[3 * p + 2] += fz * cellVolume;
        g[3 * j + 0] -= fx * selfCellVolume;
        g[3 * j + 1] -= fy * selfCellVolume;
        g[3 * j + 2] -= fz * selfCellVolume;
      }
    }
#pragma omp barrier
#pragma omp for schedule(static)
    for (long long q = 0; q < 3LL * N; q++) {
      double s = 0.0;
      for (int k = 0; k < T; k++) s += fb[k][q];
      f[q] = s;
    }
  }
}

}  // namespace

extern "C" {

// Explicit run of nsteps velocity-Verlet steps on nthreads threads ("One"
// influence function). Arguments as the oracle's orc_explicit_run: u, v are
// the initial conditions on entry and the final state on exit, bond_damage
// the initial damage on entry; prescribed-displacement sets (bc_*) set the
// velocity of their coordinate to (c1 tc + c0 - (c1 tp + c0)) / (tc - tp)
// after the half kick (src/core/Peridigm_BoundaryCondition.cpp:202-241). m
// (weighted volume) is an input here. Returns 0, or -1 on a non-finite force.
int bl_explicit_run(int nthreads, int N, const double* x, const double* vol, const double* rho,
                    const int* nlist, const double* m, double K, double mu, double sc, double dt,
                    double t0, int nsteps, int nbc, const int* bc_ptr, const int* bc_nodes,
                    const int* bc_coord, const double* bc_c0, const double* bc_c1, double* u,
                    double* v, double* a, double* bond_damage, double* dmg, double* theta,
                    double* f) {
  const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
  Mesh M = make_mesh(N, x, vol, nlist);
  std::vector<double> y(3 * (size_t)N);
  std::vector<std::vector<double>> fb(nt, std::vector<double>(3 * (size_t)N));
  const long long n3 = 3LL * N;
  int bad = 0;
  auto eval = [&]() {
    if (sc > 0.0) damage(M, y.data(), bond_damage, dmg, sc, nt);
    dilatation(M, y.data(), m, bond_damage, theta, nt);
    force(M, y.data(), m, theta, bond_damage, K, mu, f, fb, nt);
  };
#pragma omp parallel for schedule(static) num_threads(nt)
  for (long long i = 0; i < n3; i++) y[i] = x[i] + u[i];
  eval();
#pragma omp parallel for schedule(static) num_threads(nt)
  for (long long i = 0; i < n3; i++) a[i] = f[i] / rho[i / 3];
  const double dt2 = dt / 2.0;
  double tc = t0, tp = t0;
  for (int step = 1; step <= nsteps; step++) {
    tp = tc;
    tc = t0 + (step * dt);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (long long i = 0; i < n3; i++) v[i] = v[i] + dt2 * a[i];
    if (tc != tp) {
      for (int k = 0; k < nbc; k++) {
        const double cur = bc_c1[k] * tc + bc_c0[k];
        const double prev0 = bc_c1[k] * tp + bc_c0[k];
        for (int q = bc_ptr[k]; q < bc_ptr[k + 1]; q++) {
          const int node = bc_nodes[q], c = bc_coord[k];
          const double prev = (cur - prev0 == 0.0) ? u[3 * node + c] : prev0;
          v[3 * node + c] = 0.0 * (cur - prev) + 1.0 * (cur - prev) * (1.0 / (tc - tp));
        }
      }
    }
#pragma omp parallel for schedule(static) num_threads(nt)
    for (long long i = 0; i < n3; i++) {
      y[i] = x[i] + u[i] + dt * v[i];
      u[i] = u[i] + dt * v[i];
    }
    eval();
#pragma omp parallel for schedule(static) num_threads(nt) reduction(|| : bad)
    for (long long i = 0; i < n3; i++) {
      if (!std::isfinite(f[i])) bad = 1;
      a[i] = f[i] / rho[i / 3];
      v[i] = v[i] + dt2 * a[i];
    }
    if (bad) return -1;
  }
  return 0;
}

int bl_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
