// The reference's ElasticMaterial unit tests (testTwoPts, testEightPts,
// testThreePts: src/materials/unit_test/utPeridigm_ElasticMaterial.cpp:78-659)
// restated against pdgpu::ElasticMaterial through the C++ interface mirror,
// plus a CriticalStretchDamageModel case. Inputs and expected values are the
// reference's; tolerances: weighted volume as in the reference (FP64 on the
// device), dilatation and force density 2e-6 relative (FP32 bond arithmetic).
// Needs a GPU. Exit code = number of failed checks.
#include <cmath>
#include <cstdio>
#include <vector>

#include "pdgpu_peridigm.hpp"

static int failures = 0;

static void expect_rel(const char* what, double got, double want, double tol) {
  const double den = std::max(std::fabs(got), std::fabs(want));
  const bool ok = den == 0.0 || std::fabs(got - want) / den <= tol;
  if (!ok) {
    std::printf("FAIL %s: got %.17g want %.17g (tol %g)\n", what, got, want, tol);
    failures++;
  }
}

static void expect_small(const char* what, double got, double bound) {
  if (!(got <= bound)) {
    std::printf("FAIL %s: %.17g > %g\n", what, got, bound);
    failures++;
  }
}

static pdgpu::ElasticParams params() {
  pdgpu::ElasticParams p;
  p.density = 7800.0;
  p.bulkModulus = 130.0e9;
  p.shearModulus = 78.0e9;
  p.horizon = 10.0;
  return p;
}

static void run(pdgpu::ElasticMaterial& mat, int n, const std::vector<int>& nl,
                const double* x, const double* y, const double* vol, pdgpu::DataManager& dm) {
  double* X = dm.getData("Model_Coordinates", pdgpu::STEP_NONE);
  double* Y = dm.getData("Coordinates", pdgpu::STEP_NP1);
  double* V = dm.getData("Volume", pdgpu::STEP_NONE);
  for (int i = 0; i < 3 * n; i++) { X[i] = x[i]; Y[i] = y[i]; }
  for (int i = 0; i < n; i++) V[i] = vol[i];
  std::vector<int> owned(n);
  for (int i = 0; i < n; i++) owned[i] = i;
  mat.initialize(1.0, n, owned.data(), nl.data(), dm);
  mat.computeForce(1.0, n, owned.data(), nl.data(), dm);
}

static void testTwoPts() {
  pdgpu::ElasticMaterial mat(params());
  const double x[] = {0, 0, 0, 1, 0, 0}, y[] = {0, 0, 0, 2, 0, 0}, v[] = {1, 1};
  std::vector<int> nl = {1, 1, 1, 0};
  pdgpu::DataManager dm(2, 2);
  run(mat, 2, nl, x, y, v, dm);
  const double* m = dm.getData("Weighted_Volume", pdgpu::STEP_NONE);
  const double* th = dm.getData("Dilatation", pdgpu::STEP_NP1);
  const double* f = dm.getData("Force_Density", pdgpu::STEP_NP1);
  for (int i = 0; i < 2; i++) {
    expect_rel("two: weighted volume", m[i], 1.0, 1e-15);
    expect_rel("two: dilatation", th[i], 3.0, 2e-6);
  }
  expect_rel("two: f0x", f[0], 2.34e12, 1e-2);
  expect_rel("two: f1x", f[3], -2.34e12, 1e-2);
  for (int c : {1, 2, 4, 5}) expect_small("two: transverse force", std::fabs(f[c]), 1e-14);
}

static void testEightPts() {
  pdgpu::ElasticMaterial mat(params());
  std::vector<double> x = {0, 0, 0, 0, 0, 1, 0, 1, 0, 0, 1, 1, 1, 0, 0, 1, 0, 1, 1, 1, 0, 1, 1, 1};
  std::vector<double> y = x;
  for (int i = 0; i < 8; i++) if (x[3 * i + 2] == 1.0) y[3 * i + 2] = 0.98;
  std::vector<double> v(8, 1.0);
  std::vector<int> nl;
  for (int i = 0; i < 8; i++) {
    nl.push_back(7);
    for (int j = 0; j < 8; j++) if (j != i) nl.push_back(j);
  }
  pdgpu::DataManager dm(8, 56);
  run(mat, 8, nl, x.data(), y.data(), v.data(), dm);
  const double* m = dm.getData("Weighted_Volume", pdgpu::STEP_NONE);
  const double* th = dm.getData("Dilatation", pdgpu::STEP_NP1);
  const double* f = dm.getData("Force_Density", pdgpu::STEP_NP1);
  // hand solution: pair force (|Y| - |X|) * 9.75e10 along Y, factor 2
  double f0[3] = {0, 0, 0};
  auto add = [&](double t, double a, double b, double c) {
    f0[0] += 2.0 * a * t; f0[1] += 2.0 * b * t; f0[2] += 2.0 * c * t;
  };
  add((0.98 - 1.0) * 9.75e10, 0, 0, 1);
  double vm = std::sqrt(1.0 + 0.98 * 0.98);
  double t = (vm - std::sqrt(2.0)) * 9.75e10;
  add(t, 0, 1 / vm, 0.98 / vm);
  add(t, 1 / vm, 0, 0.98 / vm);
  vm = std::sqrt(2.0 + 0.98 * 0.98);
  add((vm - std::sqrt(3.0)) * 9.75e10, 1 / vm, 1 / vm, 0.98 / vm);
  for (int i = 0; i < 8; i++) {
    expect_rel("eight: weighted volume", m[i], 12.0, 1e-12);
    expect_rel("eight: dilatation", th[i], -0.01991593994643333, 2e-6);
    const double sx = x[3 * i] == 0 ? 1 : -1, sy = x[3 * i + 1] == 0 ? 1 : -1, sz = x[3 * i + 2] == 0 ? 1 : -1;
    expect_rel("eight: fx", sx * f[3 * i], f0[0], 2e-6);
    expect_rel("eight: fy", sy * f[3 * i + 1], f0[1], 2e-6);
    expect_rel("eight: fz", sz * f[3 * i + 2], f0[2], 2e-6);
  }
}

static void testThreePts() {
  pdgpu::ElasticMaterial mat(params());
  const double x[] = {1.1, 2.6, -0.1, -2.0, 0.9, -0.3, 0.0, 0.01, 1.8};
  const double y[] = {1.2, 2.4, -0.1, -1.9, 0.7, -0.8, 0.1, 0.21, 1.6};
  const double v[] = {0.9, 1.1, 0.8};
  std::vector<int> nl = {2, 1, 2, 2, 0, 2, 2, 0, 1};
  pdgpu::DataManager dm(3, 6);
  run(mat, 3, nl, x, y, v, dm);
  const double* m = dm.getData("Weighted_Volume", pdgpu::STEP_NONE);
  const double* th = dm.getData("Dilatation", pdgpu::STEP_NP1);
  const double* f = dm.getData("Force_Density", pdgpu::STEP_NP1);
  const double mref[] = {23.016479999999931, 18.64767999999997, 20.497599999999963};
  const double tref[] = {-0.114127034572639, 0.08257537372985, -0.12166177890553};
  for (int i = 0; i < 3; i++) {
    expect_rel("three: weighted volume", m[i], mref[i], 1e-12);
    expect_rel("three: dilatation", th[i], tref[i], 2e-6);
  }
  // per-bond reference values t_i <x_j - x_i> of the reference test
  const double t01[] = {-2753550531.144094, -1510011581.595148, -621769474.7744727};
  const double t10[] = {3398655528.6806289, 1863778838.308732, 767438345.18594845};
  const double t02[] = {7736514797.0571524, 15402697641.41379, -11956431959.088315};
  const double t20[] = {-8687228655.850972, -17295482505.73966, 13425717013.587864};
  const double t12[] = {5110873051.7924152, -1252163897.689138, 6133047662.15088};
  const double t21[] = {-4649613866.523322, 1139155397.2982139, -5579536639.827986};
  for (int c = 0; c < 3; c++) {
    const double r0 = t01[c] * v[1] - t10[c] * v[1] + t02[c] * v[2] - t20[c] * v[2];
    const double r1 = t10[c] * v[0] - t01[c] * v[0] + t12[c] * v[2] - t21[c] * v[2];
    const double r2 = t20[c] * v[0] - t02[c] * v[0] + t21[c] * v[1] - t12[c] * v[1];
    expect_rel("three: f0", f[c], r0, 2e-6);
    expect_rel("three: f1", f[3 + c], r1, 2e-6);
    expect_rel("three: f2", f[6 + c], r2, 2e-6);
  }
}

static void testCriticalStretch() {
  // two bonds of a three-point chain, one stretched past s_c = 0.02
  const double x[] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
  const double y[] = {0, 0, 0, 1.03, 0, 0, 0, 1.01, 0};
  const double v[] = {1, 1, 1};
  std::vector<int> nl = {2, 1, 2, 1, 0, 1, 0};
  pdgpu::DataManager dm(3, 4);
  double* X = dm.getData("Model_Coordinates", pdgpu::STEP_NONE);
  double* Y = dm.getData("Coordinates", pdgpu::STEP_NP1);
  double* V = dm.getData("Volume", pdgpu::STEP_NONE);
  for (int i = 0; i < 9; i++) { X[i] = x[i]; Y[i] = y[i]; }
  for (int i = 0; i < 3; i++) V[i] = v[i];
  pdgpu::CriticalStretchDamageModel dmg(0.02, 1.5);
  int owned[] = {0, 1, 2};
  dmg.initialize(1.0, 3, owned, nl.data(), dm);
  dm.updateState();
  dmg.computeDamage(1.0, 3, owned, nl.data(), dm);
  const double* bd = dm.getData("Bond_Damage", pdgpu::STEP_NP1);
  const double* d = dm.getData("Damage", pdgpu::STEP_NP1);
  const double bref[] = {1, 0, 1, 0}, dref[] = {0.5, 1.0, 0.0};
  for (int b = 0; b < 4; b++) expect_rel("damage: bond", bd[b], bref[b], 0.0);
  for (int i = 0; i < 3; i++) expect_rel("damage: node", d[i], dref[i], 0.0);
}

// A material and a damage model bound to the same neighbourhood list share
// one device context, each with its own constants: the damage model breaks
// the stretched bond and the material's force on the chain is unchanged by
// the damage model's binding.
static void testSharedContext() {
  const double x[] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
  const double y[] = {0, 0, 0, 1.03, 0, 0, 0, 1.01, 0};
  std::vector<int> nl = {2, 1, 2, 1, 0, 1, 0};
  int owned[] = {0, 1, 2};
  auto setup = [&](pdgpu::DataManager& dm) {
    double* X = dm.getData("Model_Coordinates", pdgpu::STEP_NONE);
    double* Y = dm.getData("Coordinates", pdgpu::STEP_NP1);
    double* V = dm.getData("Volume", pdgpu::STEP_NONE);
    for (int i = 0; i < 9; i++) { X[i] = x[i]; Y[i] = y[i]; }
    for (int i = 0; i < 3; i++) V[i] = 1.0;
  };
  pdgpu::ElasticParams p;
  p.density = 7800.0; p.bulkModulus = 130.0e9; p.shearModulus = 78.0e9; p.horizon = 1.5;
  // alone
  pdgpu::DataManager a(3, 4);
  setup(a);
  std::vector<double> f_alone(9);
  {
    pdgpu::ElasticMaterial mat(p);
    mat.initialize(1.0, 3, owned, nl.data(), a);
    mat.computeForce(1.0, 3, owned, nl.data(), a);
    const double* f = a.getData("Force_Density", pdgpu::STEP_NP1);
    for (int i = 0; i < 9; i++) f_alone[i] = f[i];
  }
  // with a damage model on the same list
  pdgpu::DataManager b(3, 4);
  setup(b);
  pdgpu::ElasticMaterial mat(p);
  pdgpu::CriticalStretchDamageModel dmg(0.02, 1.5);
  mat.initialize(1.0, 3, owned, nl.data(), b);
  dmg.initialize(1.0, 3, owned, nl.data(), b);
  b.updateState();
  dmg.computeDamage(1.0, 3, owned, nl.data(), b);
  expect_rel("shared: one context", mat.context() == dmg.context() ? 1.0 : 0.0, 1.0, 0.0);
  const double* bd = b.getData("Bond_Damage", pdgpu::STEP_NP1);
  const double bref[] = {1, 0, 1, 0};
  for (int k = 0; k < 4; k++) expect_rel("shared: bond damage", bd[k], bref[k], 0.0);
  // the force with the undamaged field equals the material's alone
  double* bdn = b.getData("Bond_Damage", pdgpu::STEP_NP1);
  for (int k = 0; k < 4; k++) bdn[k] = 0.0;
  mat.computeForce(1.0, 3, owned, nl.data(), b);
  const double* f = b.getData("Force_Density", pdgpu::STEP_NP1);
  for (int i = 0; i < 9; i++) expect_rel("shared: force", f[i], f_alone[i], 1e-12);
}

int main() {
  try {
    testTwoPts();
    testEightPts();
    testThreePts();
    testCriticalStretch();
    testSharedContext();
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 100;
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
  return failures;
}
