// CPU checks of the LDS bond scheduler (peridigm_amd/csrc/pdg_sched.hpp):
// every row's new order is a permutation of its bonds, the conflict cost
// never grows, and on a lattice tile it falls well below the unscheduled
// order (scrambled slots, as along a space-filling curve, cost about 2.6 per
// 16-lane group iteration; scheduled about 1.1 at 4 lanes per node). Prints
// the figures and PASSED.
#include <array>
#include <cstdio>
#include <random>
#include <set>
#include <vector>

#include "../../peridigm_amd/csrc/pdg_sched.hpp"

static int check(int G, const std::vector<int>& ts, const std::vector<int>& row,
                 const std::vector<int>& slot, double* ratio) {
  std::vector<int> order(slot.size());
  const long long nt = (long long)ts.size() - 1;
  pdg::SchedStats st = pdg::schedule_bonds(G, 512, 128, ts, row, slot, order, 0, nt);
  for (size_t i = 0; i + 1 < row.size(); i++) {
    std::multiset<int> a, b;
    for (int k = row[i]; k < row[i + 1]; k++) {
      a.insert(k);
      b.insert(order[k]);
      if (order[k] < row[i] || order[k] >= row[i + 1]) {
        printf("G=%d: bond moved out of its row\n", G);
        return 1;
      }
    }
    if (a != b) {
      printf("G=%d: row %zu is not a permutation\n", G, i);
      return 1;
    }
  }
  // recompute the cost of the new order independently
  long long cost = 0, groups = 0;
  const int per = G <= 16 ? 16 / G : 1;
  for (long long t = 0; t < nt; t++)
    for (int g0 = ts[t]; g0 < ts[t + 1]; g0 += per) {
      int nit = 0;
      const int gn = std::min(ts[t + 1], g0 + per) - g0;
      for (int a = 0; a < gn; a++) nit = std::max(nit, (row[g0 + a + 1] - row[g0 + a] + G - 1) / G);
      for (int it = 0; it < nit; it++) {
        std::vector<int> s;
        for (int a = 0; a < gn; a++)
          for (int l = 0; l < G; l++) {
            const int k = row[g0 + a] + it * G + l;
            if (k < row[g0 + a + 1]) s.push_back(slot[order[k]]);
          }
        if (!s.empty()) { groups++; cost += pdg::lds_group_cost(s); }
      }
    }
  if (G > 16) {  // rows left in neighbour order
    for (size_t k = 0; k < order.size(); k++)
      if (order[k] != (int)k) { printf("G=%d: rows reordered\n", G); return 1; }
    *ratio = 1.0;
    return 0;
  }
  if (cost != st.cost_after || groups != st.groups || st.cost_after > st.cost_before) {
    printf("G=%d: inconsistent figures %lld/%lld %lld/%lld before %lld\n", G, cost, st.cost_after,
           groups, st.groups, st.cost_before);
    return 1;
  }
  *ratio = (double)st.cost_after / (double)st.groups;
  printf("G=%d: %lld group iterations, cost per group %.3f in neighbour order, %.3f scheduled\n", G,
         st.groups, (double)st.cost_before / st.groups, *ratio);
  return 0;
}

int main() {
  // one tile of a cubic lattice: 128 own nodes in an 8 x 4 x 4 block, halo
  // = own nodes plus all lattice points within 3.015 spacings, slots in
  // sorted (x-major) order; rows list neighbours in slot order
  const double R = 3.015;
  std::vector<int> ts = {0, 128};
  std::vector<int> row = {0}, slot;
  std::vector<std::array<int, 3>> halo;
  for (int x = -3; x < 8 + 3; x++)
    for (int y = -3; y < 4 + 3; y++)
      for (int z = -3; z < 4 + 3; z++) halo.push_back({x, y, z});
  auto inside = [&](int x, int y, int z) { return x >= 0 && x < 8 && y >= 0 && y < 4 && z >= 0 && z < 4; };
  // slots in a scrambled order (as along a space-filling curve, neighbours
  // of a node are not consecutive slots); rows list neighbours by slot
  std::mt19937 rng(7);
  for (size_t h = halo.size() - 1; h > 0; h--) std::swap(halo[h], halo[rng() % (h + 1)]);
  std::vector<int> own;
  for (size_t h = 0; h < halo.size(); h++)
    if (inside(halo[h][0], halo[h][1], halo[h][2])) own.push_back((int)h);
  for (int i : own) {
    for (size_t h = 0; h < halo.size(); h++) {
      const int dx = halo[h][0] - halo[i][0], dy = halo[h][1] - halo[i][1], dz = halo[h][2] - halo[i][2];
      const int d2 = dx * dx + dy * dy + dz * dz;
      if (d2 > 0 && d2 <= R * R) slot.push_back((int)h);
    }
    row.push_back((int)slot.size());
  }
  double r4 = 0, r8 = 0, r16 = 0, r32 = 0;
  if (check(4, ts, row, slot, &r4) || check(8, ts, row, slot, &r8) || check(16, ts, row, slot, &r16) ||
      check(32, ts, row, slot, &r32))
    return 1;
  // random slots: the order must still be a valid permutation
  std::vector<int> rrow = {0}, rslot;
  std::vector<int> rts = {0};
  for (int t = 0; t < 6; t++) {
    for (int i = 0; i < 128; i++) {
      const int deg = 40 + rng() % 89;
      for (int k = 0; k < deg; k++) rslot.push_back(rng() % 700);
      rrow.push_back((int)rslot.size());
    }
    rts.push_back(rts.back() + 128);
  }
  double rr = 0;
  if (check(4, rts, rrow, rslot, &rr)) return 1;
  if (r4 > 1.2 || r8 > 1.4 || r16 > 1.8 || rr > 1.3) {
    printf("lattice schedule not close to conflict-free\n");
    return 1;
  }
  printf("PASSED\n");
  return 0;
}
