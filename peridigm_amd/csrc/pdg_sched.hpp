// Bond order inside each row of a tiled CSR, chosen for conflict-free LDS
// reads.
//
// The tiled bond loops give a node G lanes; in iteration `it` lane l of the
// node reads the halo slot of its bond it * G + l with 16-byte LDS reads.
// gfx950 serves a 16-byte read 16 lanes at a time, conflict-free when the 16
// lanes address distinct slot residues modulo 16 (lanes reading the same
// slot are served together). A row's order is free (bond fields are mapped
// through the context's bond index), so this picks, iteration by iteration,
// which bonds of the 16 / G nodes sharing a 16-lane group go together:
// a bond whose slot another lane of the group already reads, else one whose
// residue is least used so far in the iteration (ties: the residue the node
// has most bonds left in, which keeps later iterations flexible).
//
// Host-only, header-only so it can be tested on the CPU.
#pragma once

#include <algorithm>
#include <vector>

namespace pdg {

struct SchedStats {
  long long groups = 0;       // (16-lane group, iteration) pairs with an active lane
  long long cost_before = 0;  // sum over them of the largest number of distinct
  long long cost_after = 0;   //   slots sharing one residue (1 = conflict-free)
};

// Cost of one (group, iteration): the largest number of distinct slots with
// one residue modulo 16.
inline int lds_group_cost(std::vector<int>& slots) {
  if (slots.empty()) return 0;
  std::sort(slots.begin(), slots.end());
  slots.erase(std::unique(slots.begin(), slots.end()), slots.end());
  int cnt[16] = {0};
  int mx = 0;
  for (int s : slots) mx = std::max(mx, ++cnt[s & 15]);
  return mx;
}

// order[b] for b in [row[i], row[i+1]) receives the old bond index placed at
// position b. Nodes of tile t are [tile_start[t], tile_start[t+1]), processed
// in rounds of `wg / G` nodes. slot[b] is the halo slot of bond b. Rows are
// left in place for G > 16 and for rows longer than maxdeg.
inline SchedStats schedule_bonds(int G, int wg, int maxdeg, const std::vector<int>& tile_start,
                                 const std::vector<int>& row, const std::vector<int>& slot,
                                 std::vector<int>& order, long long t0, long long t1) {
  SchedStats st;
  const int per_group = G <= 16 ? 16 / G : 0;
  const int ng = wg / G;
  std::vector<int> buck[16];
  std::vector<int> used;
  std::vector<int> before, after;
  for (long long t = t0; t < t1; t++) {
    const int n0 = tile_start[t], n1 = tile_start[t + 1];
    for (int i = n0; i < n1; i++)
      for (int b = row[i]; b < row[i + 1]; b++) order[b] = b;
    if (per_group == 0) continue;
    for (int base = n0; base < n1; base += ng) {
      const int rn1 = std::min(n1, base + ng);
      for (int g0 = base; g0 < rn1; g0 += per_group) {
        const int gn = std::min(rn1, g0 + per_group) - g0;
        bool ok = true;
        int nit = 0;
        for (int a = 0; a < gn; a++) {
          const int deg = row[g0 + a + 1] - row[g0 + a];
          if (deg > maxdeg) ok = false;
          nit = std::max(nit, (deg + G - 1) / G);
        }
        // cost of the original order
        for (int it = 0; it < nit; it++) {
          before.clear();
          for (int a = 0; a < gn; a++) {
            const int rs = row[g0 + a], deg = row[g0 + a + 1] - rs;
            for (int l = 0; l < G && it * G + l < deg; l++) before.push_back(slot[rs + it * G + l]);
          }
          if (!before.empty()) {
            st.groups++;
            st.cost_before += lds_group_cost(before);
          }
        }
        if (!ok) {
          st.cost_after = st.cost_before;
          continue;
        }
        // per node: remaining bonds by residue
        std::vector<std::vector<int>> rem(gn * 16);
        std::vector<int> left(gn);
        for (int a = 0; a < gn; a++) {
          const int rs = row[g0 + a], re = row[g0 + a + 1];
          left[a] = re - rs;
          for (int b = rs; b < re; b++) rem[a * 16 + (slot[b] & 15)].push_back(b);
          for (int r = 0; r < 16; r++) std::reverse(rem[a * 16 + r].begin(), rem[a * 16 + r].end());
        }
        std::vector<int> nodes(gn);
        for (int it = 0; it < nit; it++) {
          int cnt[16] = {0};
          used.clear();
          for (int a = 0; a < gn; a++) nodes[a] = a;
          std::stable_sort(nodes.begin(), nodes.end(),
                           [&](int x, int y) { return left[x] > left[y]; });
          after.clear();
          for (int a : nodes) {
            const int rs = row[g0 + a];
            const int need = std::min(G, left[a]);
            for (int p = 0; p < need; p++) {
              int pick = -1, pr = -1;
              // a slot another lane already reads: served together
              for (int s : used) {
                std::vector<int>& v = rem[a * 16 + (s & 15)];
                for (size_t q = 0; q < v.size(); q++)
                  if (slot[v[q]] == s) {
                    pick = v[q];
                    pr = s & 15;
                    v.erase(v.begin() + q);
                    break;
                  }
                if (pick >= 0) break;
              }
              if (pick < 0) {
                int best = -1;
                for (int r = 0; r < 16; r++) {
                  if (rem[a * 16 + r].empty()) continue;
                  if (best < 0 || cnt[r] < cnt[best] ||
                      (cnt[r] == cnt[best] && rem[a * 16 + r].size() > rem[a * 16 + best].size()))
                    best = r;
                }
                pr = best;
                pick = rem[a * 16 + best].back();
                rem[a * 16 + best].pop_back();
                used.push_back(slot[pick]);
                cnt[best]++;
              }
              (void)pr;
              order[rs + it * G + p] = pick;
              after.push_back(slot[pick]);
            }
            left[a] -= need;
          }
          if (!after.empty()) st.cost_after += lds_group_cost(after);
        }
      }
    }
  }
  return st;
}

}  // namespace pdg
