// flowplan.cpp -- task lists of the dataflow panel factorization (host).
//
// For the supernodes of one tree level, the tasks over 128 x 128 tiles of
// each panel (diagonal blocks D(k), TRSM row tiles T(k, R), trailing-update
// tiles U(k, R, C); the TRSM and update that lead to the next diagonal block
// split into four 32-row quarters), their dependencies, and a topological
// order that favours the chain of diagonal blocks.  The device executes the
// list in order with per-tile operation counters; flow_check replays that
// protocol on the host.
#include <algorithm>
#include <queue>
#include <utility>

#include "common.hpp"
#include "symbolic.hpp"

namespace hykkt {

namespace {
constexpr int NB = 128;
inline int cdiv(int a, int b) { return (a + b - 1) / b; }
}  // namespace

// Kahn's algorithm over the tasks of the level's supernodes, taking the ready
// task of smallest priority: per step k of a supernode, the diagonal block,
// the short TRSM and update that lead to the next diagonal block, the tiles
// that feed the next step's chain, the other TRSM tiles and the rest of the
// next block column, then the bulk of the trailing update.
void flow_tasks(const CholPlan& pl, int b0, int b1, std::vector<int64_t>& cnt_off, int64_t& ncnt,
                std::vector<int>& flow, double& flops, double& bytes)
{
  struct Task {
    int type, shrt, qr, s, k, R, C;
    double pri;
  };
  std::vector<Task> tk;
  std::vector<std::vector<int>> succ;
  std::vector<int> indeg;
  for (int q = b0; q < b1; q++) {
    const int s = pl.level_sn[q];
    const int nc = pl.nc(s), f = nc + pl.nr(s);
    const int ncb = cdiv(nc, NB), nrb = cdiv(f, NB);
    cnt_off[s] = ncnt;
    ncnt += static_cast<int64_t>(nrb) * ncb;
    // ids: D[k], T[k][R] (1 or 4 quarters), U[(k * nrb + R) * ncb + C]
    const int base = static_cast<int>(tk.size());
    auto add = [&](int type, int shrt, int qr, int k, int R, int C, double pri) {
      tk.push_back({type, shrt, qr, s, k, R, C, pri});
      succ.emplace_back();
      indeg.push_back(0);
      return static_cast<int>(tk.size()) - 1;
    };
    auto dep = [&](int from, int to) {
      succ[from].push_back(to);
      indeg[to]++;
    };
    std::vector<int> D(ncb);
    std::vector<std::vector<int>> T(static_cast<size_t>(ncb) * nrb);
    std::vector<std::vector<int>> U(static_cast<size_t>(ncb) * nrb * ncb);
    auto Uid = [&](int k, int R, int C) -> std::vector<int>& {
      return U[(static_cast<size_t>(k) * nrb + R) * ncb + C];
    };
    for (int k = 0; k < ncb; k++) {
      const int k0 = k * NB, b = std::min(NB, nc - k0), cs = k0 + b;
      D[k] = add(0, 0, 0, k, k, k, k - 0.85);
      if (k > 0)
        for (int id : Uid(k - 1, k, k)) dep(id, D[k]);
      if (cs >= f) continue;
      for (int R = cs / NB; R < nrb; R++) {
        std::vector<int>& t = T[static_cast<size_t>(k) * nrb + R];
        const bool chain = (R == k + 1);
        const double pri = chain ? k - 0.8 : (R == k + 2 ? k - 0.7 : k - 0.5 + 1e-6 * R);
        if (chain && k + 1 < ncb) {
          for (int qr = 0; qr < 4; qr++) t.push_back(add(1, 1, qr, k, R, k, pri));
        } else {
          t.push_back(add(1, 0, 0, k, R, k, pri));
        }
        for (int id : t) {
          dep(D[k], id);
          if (k > 0 && R > k)
            for (int u : Uid(k - 1, R, k)) dep(u, id);
        }
      }
    }
    for (int k = 0; k + 1 < ncb; k++) {
      for (int C = k + 1; C < ncb; C++)
        for (int R = C; R < nrb; R++) {
          std::vector<int>& u = Uid(k, R, C);
          double pri;
          if (C == k + 1) pri = (R == C) ? k - 0.75 : (R == k + 2 ? k - 0.65 : k - 0.4 + 1e-6 * R);
          else pri = (C == k + 2 && R == C) ? k - 0.6 : k + 0.5 + 1e-9 * (1e4 * C + R);
          if (C == k + 1 && R == C) {
            for (int qr = 0; qr < 4; qr++) u.push_back(add(2, 1, qr, k, R, C, pri));
          } else {
            u.push_back(add(2, 0, 0, k, R, C, pri));
          }
          for (int id : u) {
            for (int t : T[static_cast<size_t>(k) * nrb + R]) dep(t, id);
            for (int t : T[static_cast<size_t>(k) * nrb + C]) dep(t, id);
            if (k > 0)
              for (int p : Uid(k - 1, R, C)) dep(p, id);
          }
        }
    }
    (void)base;
    // algorithmic work, as the per-step launches count it
    for (int k = 0; k < ncb; k++) {
      const int k0 = k * NB, b = std::min(NB, nc - k0);
      const double db = b, r = f - (k0 + b), cc = nc - (k0 + b);
      const double tri = db * (db + 1) / 2, pairs = cc * r - cc * (cc - 1) / 2;
      flops += db * db * db / 3 + db * db / 2 + db / 6 + r * db * db + 2.0 * db * pairs;
      bytes += 16.0 * tri + 8.0 * (2 * r * db + tri) + 8.0 * 2 * pairs;
    }
  }
  // priority-ordered topological sort
  using Key = std::pair<double, int>;
  std::priority_queue<Key, std::vector<Key>, std::greater<Key>> ready;
  for (int i = 0; i < static_cast<int>(tk.size()); i++)
    if (indeg[i] == 0) ready.push({tk[i].pri, i});
  int placed = 0;
  while (!ready.empty()) {
    const int i = ready.top().second;
    ready.pop();
    const Task& t = tk[i];
    flow.insert(flow.end(), {t.type | (t.shrt << 2) | (t.qr << 4), t.s, t.k, t.R | (t.C << 16)});
    placed++;
    for (int j : succ[i])
      if (--indeg[j] == 0) ready.push({tk[j].pri, j});
  }
  HYKKT_REQUIRE(placed == static_cast<int>(tk.size()), HYKKT_E_STATE,
                "dataflow task graph has a cycle");
}


int flow_check(const CholPlan& pl, int b0, int b1, const std::vector<int64_t>& cnt_off,
               int64_t cnt_lo, int64_t cnt_n, const int* tasks, int ntasks)
{
  std::vector<int64_t> cnt(static_cast<size_t>(std::max<int64_t>(cnt_n, 1)), 0);
  std::vector<int> lvl(pl.ns, 0);
  for (int q = b0; q < b1; q++) lvl[pl.level_sn[q]] = 1;
  for (int t = 0; t < ntasks; t++) {
    const int w0 = tasks[4 * t], s = tasks[4 * t + 1], k = tasks[4 * t + 2], rc = tasks[4 * t + 3];
    const int type = w0 & 3, shrt = (w0 >> 2) & 1;
    const int R = rc & 0xffff, C = rc >> 16;
    if (s < 0 || s >= pl.ns || !lvl[s]) return t + 1;
    const int nc = pl.nc(s), f = nc + pl.nr(s), ncb = cdiv(nc, NB), nrb = cdiv(f, NB);
    if (k < 0 || k >= ncb || R < 0 || R >= nrb || C < 0 || C >= ncb) return t + 1;
    auto tile = [&](int r, int c) -> int64_t& { return cnt[cnt_off[s] - cnt_lo + r * ncb + c]; };
    bool ok;
    if (type == 0) ok = tile(k, k) >= 4 * k && R == k && C == k;
    else if (type == 1) ok = tile(k, k) >= 4 * k + 4 && (R == k || tile(R, k) >= 4 * k) && C == k && R >= k;
    else ok = C > k && R >= C && tile(R, k) >= 4 * k + 4 && tile(C, k) >= 4 * k + 4 && tile(R, C) >= 4 * k;
    if (!ok) return t + 1;
    int64_t& ct = (type == 0) ? tile(k, k) : (type == 1 ? tile(R, k) : tile(R, C));
    ct += shrt ? 1 : 4;
  }
  // every tile of every panel ends complete: updates from steps 0..C-1 and
  // its D or T (the partial last diagonal block also its own T)
  for (int q = b0; q < b1; q++) {
    const int s = pl.level_sn[q];
    const int nc = pl.nc(s), f = nc + pl.nr(s), ncb = cdiv(nc, NB), nrb = cdiv(f, NB);
    for (int C = 0; C < ncb; C++)
      for (int R = C; R < nrb; R++) {
        const int cs = C * NB + std::min(NB, nc - C * NB);
        int64_t want = 4 * (C + 1);
        if (R == C && cs < f && cs / NB == C) want += 4;   // T(C, C) of a partial block
        if (cnt[cnt_off[s] - cnt_lo + R * ncb + C] != want) return -(R * 65536 + C + 1);
      }
  }
  return 0;
}

}  // namespace hykkt
