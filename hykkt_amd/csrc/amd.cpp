// amd.cpp -- approximate minimum degree ordering (host, once per pattern).
//
// Written from the published method: P. Amestoy, T. Davis, I. Duff, "An
// approximate minimum degree ordering algorithm", SIAM J. Matrix Anal. Appl.
// 17(4):886-905, 1996.  The elimination graph is kept as a quotient graph:
// each uneliminated (super)variable i has a list of adjacent variables A_i and
// of adjacent elements E_i; each element e (an eliminated pivot) has its
// variable list L_e.  Eliminating pivot p forms L_p = A_p U (U_{e in E_p} L_e),
// absorbs the elements of E_p, and bounds each i in L_p's external degree by
//   min(n-k, d_i + |L_p\i|, |A_i\i| + |L_p\i| + sum_{e in E_i\p} |L_e\L_p|),
// with |L_e\L_p| computed for all touched elements in one pass.  Elements
// with |L_e\L_p| = 0 are absorbed into p (aggressive absorption), and
// variables of L_p with identical (A_i, E_i) are merged into supervariables.
// The CUDA reference calls the closed cusolverSpXcsrsymamdHost here
// (PermClass.cpp:71-99); this is an independent ordering, so the permutation
// itself is not expected to match the reference's, only to be a valid
// fill-reducing permutation.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "symbolic.hpp"

namespace hykkt {

namespace {

enum : char { VAR = 0, ELEM = 1, DEAD = 2 };

struct Buckets {
  std::vector<int> head, next, prev;
  int mindeg = 0;
  explicit Buckets(int n) : head(n + 1, -1), next(n, -1), prev(n, -1) {}
  void insert(int i, int d)
  {
    next[i] = head[d];
    prev[i] = -1;
    if (head[d] >= 0) prev[head[d]] = i;
    head[d] = i;
    if (d < mindeg) mindeg = d;
  }
  void remove(int i, int d)
  {
    if (prev[i] >= 0) next[prev[i]] = next[i];
    else head[d] = next[i];
    if (next[i] >= 0) prev[next[i]] = prev[i];
    next[i] = prev[i] = -1;
  }
};

}  // namespace

std::vector<int> amd_order(const Pattern& a)
{
  const int n = a.n;
  std::vector<std::vector<int>> adjV(n), adjE(n), Le(n);
  for (int i = 0; i < n; i++) {
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      int j = a.ci[p];
      if (j == i) continue;
      adjV[i].push_back(j);
      adjV[j].push_back(i);
    }
  }
  for (int i = 0; i < n; i++) {
    std::sort(adjV[i].begin(), adjV[i].end());
    adjV[i].erase(std::unique(adjV[i].begin(), adjV[i].end()), adjV[i].end());
  }
  std::vector<int> nv(n, 1), deg(n), esize(n, 0), w(n, 0), wtag(n, 0), mark(n, 0);
  std::vector<int> mem_next(n, -1), mem_tail(n);
  std::vector<char> status(n, VAR);
  for (int i = 0; i < n; i++) mem_tail[i] = i;
  Buckets bk(n);
  bk.mindeg = n;
  for (int i = 0; i < n; i++) {
    deg[i] = static_cast<int>(adjV[i].size());
    bk.insert(i, deg[i]);
  }
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> Lp, newE, newV;
  std::vector<std::pair<uint64_t, int>> hashes;
  int wround = 0, mround = 0, k = 0;

  while (k < n) {
    while (bk.mindeg <= n && bk.head[bk.mindeg] < 0) bk.mindeg++;
    HYKKT_REQUIRE(bk.mindeg <= n, HYKKT_E_STATE, "amd: degree lists exhausted early");
    const int p = bk.head[bk.mindeg];
    bk.remove(p, deg[p]);

    // ---- form the new element L_p ----
    ++mround;
    mark[p] = mround;
    Lp.clear();
    for (int v : adjV[p])
      if (status[v] == VAR && nv[v] > 0 && mark[v] != mround) { mark[v] = mround; Lp.push_back(v); }
    for (int e : adjE[p]) {
      if (status[e] != ELEM) continue;
      for (int v : Le[e])
        if (status[v] == VAR && nv[v] > 0 && mark[v] != mround) { mark[v] = mround; Lp.push_back(v); }
      status[e] = DEAD;                       // absorbed into p
      std::vector<int>().swap(Le[e]);
    }
    status[p] = ELEM;
    std::vector<int>().swap(adjV[p]);
    std::vector<int>().swap(adjE[p]);
    for (int v = p; v >= 0; v = mem_next[v]) order.push_back(v);
    k += nv[p];
    int lpsize = 0;
    for (int v : Lp) lpsize += nv[v];
    esize[p] = lpsize;
    for (int v : Lp) bk.remove(v, deg[v]);

    // ---- |L_e \ L_p| for every element adjacent to L_p ----
    ++wround;
    for (int i : Lp)
      for (int e : adjE[i]) {
        if (status[e] != ELEM) continue;
        if (wtag[e] != wround) { wtag[e] = wround; w[e] = esize[e]; }
        w[e] -= nv[i];
      }

    // ---- prune lists and bound degrees ----
    for (int i : Lp) {
      newE.clear();
      long long dext = 0;
      for (int e : adjE[i]) {
        if (status[e] != ELEM) continue;
        if (w[e] == 0) {                      // L_e within L_p: aggressive absorption
          status[e] = DEAD;
          std::vector<int>().swap(Le[e]);
          continue;
        }
        newE.push_back(e);
        dext += w[e];
      }
      newE.push_back(p);
      adjE[i].swap(newE);
      newV.clear();
      long long av = 0;
      for (int v : adjV[i]) {
        if (status[v] != VAR || nv[v] == 0 || mark[v] == mround) continue;
        newV.push_back(v);
        av += nv[v];
      }
      adjV[i].swap(newV);
      long long others = lpsize - nv[i];
      long long d = std::min<long long>(static_cast<long long>(n - k) - nv[i],
                                        std::min<long long>(deg[i] + others, av + others + dext));
      deg[i] = static_cast<int>(std::max<long long>(d, 0));
    }

    // ---- supervariable detection among L_p ----
    hashes.clear();
    for (int i : Lp) {
      uint64_t h = 0;
      for (int e : adjE[i]) h += static_cast<uint64_t>(e) * 0x9E3779B97F4A7C15ull + 1;
      for (int v : adjV[i]) h += static_cast<uint64_t>(v) * 0xC2B2AE3D27D4EB4Full + 7;
      h ^= (static_cast<uint64_t>(adjE[i].size()) << 40) ^ adjV[i].size();
      hashes.push_back({h, i});
    }
    std::sort(hashes.begin(), hashes.end());
    for (size_t s = 0; s < hashes.size();) {
      size_t t = s;
      while (t < hashes.size() && hashes[t].first == hashes[s].first) t++;
      for (size_t x = s; x < t; x++) {
        int i = hashes[x].second;
        if (nv[i] == 0) continue;
        std::vector<int> si_e(adjE[i]), si_v(adjV[i]);
        std::sort(si_e.begin(), si_e.end());
        std::sort(si_v.begin(), si_v.end());
        for (size_t y = x + 1; y < t; y++) {
          int j = hashes[y].second;
          if (nv[j] == 0 || adjE[j].size() != si_e.size() || adjV[j].size() != si_v.size())
            continue;
          std::vector<int> sj_e(adjE[j]), sj_v(adjV[j]);
          std::sort(sj_e.begin(), sj_e.end());
          std::sort(sj_v.begin(), sj_v.end());
          if (sj_e != si_e || sj_v != si_v) continue;
          // j is indistinguishable from i: merge j into i
          deg[i] -= nv[j];
          nv[i] += nv[j];
          nv[j] = 0;
          status[j] = DEAD;
          mem_next[mem_tail[i]] = j;
          mem_tail[i] = mem_tail[j];
          std::vector<int>().swap(adjV[j]);
          std::vector<int>().swap(adjE[j]);
        }
      }
      s = t;
    }

    // ---- reinsert the surviving principal variables, compact L_p ----
    int cnt = 0;
    for (int i : Lp) {
      if (nv[i] == 0) continue;
      if (deg[i] < 0) deg[i] = 0;
      bk.insert(i, deg[i]);
      Lp[cnt++] = i;
    }
    Lp.resize(cnt);
    Le[p] = Lp;
  }
  HYKKT_REQUIRE(static_cast<int>(order.size()) == n, HYKKT_E_STATE, "amd: incomplete ordering");
  return order;
}

}  // namespace hykkt
