// cholplan.cpp -- symbolic analysis for the supernodal multifrontal
// Cholesky (host, once per sparsity pattern).
//
// Methods (all written from the published descriptions):
//  * elimination tree with path compression (J. Liu, "The role of elimination
//    trees in sparse factorization", SIAM J. Matrix Anal. Appl. 11(1), 1990);
//  * column counts from row subtrees: colcount(j) = number of rows i whose
//    row subtree T_r(i) contains j.  Each T_r(i) is the union of the tree
//    paths from its leaves (first-descendant test in postorder) to i; the
//    counts are accumulated as +1 at each leaf, -1 at the least common
//    ancestor of consecutive leaves, -1 above i, then summed over subtrees
//    (Gilbert, Ng & Peyton, SIAM J. Matrix Anal. Appl. 15(4), 1994);
//  * fundamental supernodes, then relaxed amalgamation of a supernode with
//    its contiguous parent when the explicit zeros stay small.
// The reference delegates this step to cusolverSpXcsrcholAnalysis
// (CholeskyClass.cpp:31-58).
#include <algorithm>
#include <numeric>

#include "symbolic.hpp"

namespace hykkt {

namespace {

int find_root(std::vector<int>& anc, int i)
{
  int r = i;
  while (anc[r] != r) r = anc[r];
  while (anc[i] != r) {
    int nx = anc[i];
    anc[i] = r;
    i = nx;
  }
  return r;
}

}  // namespace

// Update-matrix space with reuse.  U(s) (nr x nr lower triangle, packed by
// columns: nr (nr+1)/2 doubles) is written at s's level
// and read at its parent's level, then dead: allocate the update matrices of
// a level before freeing those its supernodes consume (their children's), so
// the two never share space, with best fit over a list of free ranges.
// u_off[s] is the offset of U(s) (unused for nr = 0), u_off[ns] the pool size.
static void assign_update_space(CholPlan& pl)
{
  const int ns = pl.ns;
  pl.u_off.assign(ns + 1, 0);
  std::vector<std::pair<int64_t, int64_t>> freel;   // (offset, size), sorted by offset
  int64_t top = 0;
  auto alloc = [&](int64_t sz) -> int64_t {
    int best = -1;
    for (int i = 0; i < static_cast<int>(freel.size()); i++)
      if (freel[i].second >= sz && (best < 0 || freel[i].second < freel[best].second)) best = i;
    if (best < 0) {
      // grow the pool, absorbing a free range that ends at the top
      if (!freel.empty() && freel.back().first + freel.back().second == top) {
        const int64_t off = freel.back().first;
        freel.pop_back();
        top = off + sz;
        return off;
      }
      const int64_t off = top;
      top += sz;
      return off;
    }
    const int64_t off = freel[best].first;
    freel[best].first += sz;
    freel[best].second -= sz;
    if (freel[best].second == 0) freel.erase(freel.begin() + best);
    return off;
  };
  auto release = [&](int64_t off, int64_t sz) {
    auto it = std::lower_bound(freel.begin(), freel.end(), std::make_pair(off, int64_t{0}));
    it = freel.insert(it, {off, sz});
    // coalesce with the next and the previous range
    if (it + 1 != freel.end() && it->first + it->second == (it + 1)->first) {
      it->second += (it + 1)->second;
      freel.erase(it + 1);
    }
    if (it != freel.begin() && (it - 1)->first + (it - 1)->second == it->first) {
      (it - 1)->second += it->second;
      freel.erase(it);
    }
  };
  const int nlev = static_cast<int>(pl.level_ptr.size()) - 1;
  for (int l = 0; l < nlev; l++) {
    for (int q = pl.level_ptr[l]; q < pl.level_ptr[l + 1]; q++) {
      const int s = pl.level_sn[q];
      const int64_t nr = pl.nr(s);
      if (nr > 0) pl.u_off[s] = alloc(nr * (nr + 1) / 2);
    }
    for (int q = pl.level_ptr[l]; q < pl.level_ptr[l + 1]; q++) {
      const int s = pl.level_sn[q];
      for (int t = pl.child_ptr[s]; t < pl.child_ptr[s + 1]; t++) {
        const int c = pl.child_list[t];
        const int64_t nr = pl.nr(c);
        if (nr > 0) release(pl.u_off[c], nr * (nr + 1) / 2);
      }
    }
  }
  pl.u_off[ns] = top;
}

CholPlan chol_analyze(const Pattern& a, const CholOptions& opt)
{
  const int n = a.n;
  HYKKT_REQUIRE(a.m == n, HYKKT_E_ARG, "chol_analyze: matrix must be square");
  CholPlan pl;
  pl.n = n;

  // ---- 1. elimination tree of the lower triangle ----
  std::vector<int> parent(n, -1), anc(n, -1);
  for (int i = 0; i < n; i++) {
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      int j = a.ci[p];
      while (j != -1 && j < i) {
        int nx = anc[j];
        anc[j] = i;
        if (nx == -1) { parent[j] = i; break; }
        j = nx;
      }
    }
  }

  // ---- 2. postorder (children visited in increasing index) ----
  std::vector<int> head(n, -1), nxt(n, -1);
  for (int j = n - 1; j >= 0; j--) {
    if (parent[j] >= 0) { nxt[j] = head[parent[j]]; head[parent[j]] = j; }
  }
  pl.post.resize(n);
  pl.ipost.resize(n);
  {
    std::vector<int> stack;
    int k = 0;
    for (int r = 0; r < n; r++) {
      if (parent[r] != -1) continue;
      stack.push_back(r);
      while (!stack.empty()) {
        int t = stack.back();
        if (head[t] >= 0) {
          int c = head[t];
          head[t] = nxt[c];
          stack.push_back(c);
        } else {
          stack.pop_back();
          pl.post[k] = t;
          pl.ipost[t] = k++;
        }
      }
    }
    HYKKT_REQUIRE(k == n, HYKKT_E_STATE, "chol_analyze: postorder incomplete");
  }
  std::vector<int> par(n, -1);
  for (int j = 0; j < n; j++)
    if (parent[j] >= 0) par[pl.ipost[j]] = pl.ipost[parent[j]];

  // ---- 3. lower triangle in internal order, by column (CSC) ----
  std::vector<int> cp(n + 1, 0);
  for (int i = 0; i < n; i++)
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      if (a.ci[p] > i) continue;
      int r = pl.ipost[i], c = pl.ipost[a.ci[p]];
      cp[std::min(r, c) + 1]++;
    }
  for (int j = 0; j < n; j++) cp[j + 1] += cp[j];
  const int nlow = cp[n];
  std::vector<int> ri(nlow), src(nlow), fill(cp.begin(), cp.end() - 1);
  for (int i = 0; i < n; i++)
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      if (a.ci[p] > i) continue;
      int r = pl.ipost[i], c = pl.ipost[a.ci[p]];
      int lo = std::min(r, c), hi = std::max(r, c);
      int q = fill[lo]++;
      ri[q] = hi;
      src[q] = p;
    }
  for (int j = 0; j < n; j++) {
    std::vector<std::pair<int, int>> t;
    for (int q = cp[j]; q < cp[j + 1]; q++) t.push_back({ri[q], src[q]});
    std::sort(t.begin(), t.end());
    for (int q = cp[j]; q < cp[j + 1]; q++) { ri[q] = t[q - cp[j]].first; src[q] = t[q - cp[j]].second; }
  }

  // ---- 4. column counts ----
  std::vector<int> first(n, -1), delta(n, 0), maxfirst(n, -1), prevleaf(n, -1), nleaf(n, 0);
  for (int j = 0; j < n; j++)
    for (int d = j; d != -1 && first[d] == -1; d = par[d]) first[d] = j;
  std::vector<int> uf(n);
  std::iota(uf.begin(), uf.end(), 0);
  for (int j = 0; j < n; j++) {
    for (int q = cp[j]; q < cp[j + 1]; q++) {
      int i = ri[q];
      if (i <= j) continue;
      if (first[j] > maxfirst[i]) {            // j is a leaf of T_r(i)
        maxfirst[i] = first[j];
        delta[j] += 1;
        if (prevleaf[i] != -1) delta[find_root(uf, prevleaf[i])] -= 1;
        prevleaf[i] = j;
        nleaf[i]++;
      }
    }
    if (par[j] != -1) uf[j] = par[j];
  }
  for (int i = 0; i < n; i++) {
    if (nleaf[i] == 0) delta[i] += 1;
    if (par[i] != -1) delta[par[i]] -= 1;
  }
  std::vector<int64_t> cc(n);
  for (int j = 0; j < n; j++) cc[j] = delta[j];
  for (int j = 0; j < n; j++)
    if (par[j] != -1) cc[par[j]] += cc[j];
  double flops = 0.0;
  int64_t nnzl = 0;
  std::vector<int> nchild(n, 0);
  for (int j = 0; j < n; j++) {
    HYKKT_REQUIRE(cc[j] >= 1, HYKKT_E_STATE, "chol_analyze: bad column count");
    flops += static_cast<double>(cc[j]) * static_cast<double>(cc[j]);
    nnzl += cc[j];
    if (par[j] != -1) nchild[par[j]]++;
  }
  pl.flops = flops;
  pl.nnz_l_true = nnzl;

  // ---- 5. fundamental supernodes ----
  std::vector<int> sfirst;
  for (int j = 0; j < n; j++) {
    bool join = j > 0 && par[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1;
    if (!join) sfirst.push_back(j);
  }
  sfirst.push_back(n);
  int ns0 = static_cast<int>(sfirst.size()) - 1;

  // ---- 6. relaxed amalgamation with the contiguous parent ----
  std::vector<int> s_of(n);
  for (int s = 0; s < ns0; s++)
    for (int j = sfirst[s]; j < sfirst[s + 1]; j++) s_of[j] = s;
  std::vector<int> sfa(sfirst.begin(), sfirst.end() - 1), sla(ns0);   // first col, one-past-last
  std::vector<int64_t> strue(ns0, 0);
  std::vector<int> snr(ns0);
  std::vector<char> alive(ns0, 1);
  for (int s = 0; s < ns0; s++) {
    sla[s] = sfirst[s + 1];
    for (int j = sfa[s]; j < sla[s]; j++) strue[s] += cc[j];
    snr[s] = static_cast<int>(cc[sla[s] - 1] - 1);
  }
  auto stored = [](int64_t nc, int64_t nr) { return nc * (nc + 1) / 2 + nc * nr; };
  for (int s = 0; s < ns0; s++) {
    int lastcol = sla[s] - 1;
    if (par[lastcol] == -1) continue;
    int p = s_of[par[lastcol]];                 // current owner (s_of follows merges)
    if (sla[s] != sfa[p]) continue;            // only the contiguous (last) child
    int64_t ncm = sla[p] - sfa[s];
    int64_t st = stored(ncm, snr[p]);
    int64_t tr = strue[s] + strue[p];
    double frac = static_cast<double>(st - tr) / static_cast<double>(st);
    bool merge = ncm <= opt.relax_small ||
                 (ncm <= opt.relax_mid_cols && frac <= opt.relax_frac_mid) ||
                 frac <= opt.relax_frac_big;
    if (!merge) continue;
    sfa[p] = sfa[s];
    strue[p] = tr;
    alive[s] = 0;
    for (int j = sfa[s]; j < sla[s]; j++) s_of[j] = p;
  }
  pl.sn_first.clear();
  for (int s = 0; s < ns0; s++)
    if (alive[s]) pl.sn_first.push_back(sfa[s]);
  pl.sn_first.push_back(n);
  const int ns = static_cast<int>(pl.sn_first.size()) - 1;
  pl.ns = ns;
  pl.sn_of.resize(n);
  for (int s = 0; s < ns; s++)
    for (int j = pl.sn_first[s]; j < pl.sn_first[s + 1]; j++) pl.sn_of[j] = s;
  pl.sn_parent.assign(ns, -1);
  for (int s = 0; s < ns; s++) {
    int last = pl.sn_first[s + 1] - 1;
    if (par[last] != -1) pl.sn_parent[s] = pl.sn_of[par[last]];
  }

  // ---- 7. children, row structures ----
  pl.child_ptr.assign(ns + 1, 0);
  for (int s = 0; s < ns; s++)
    if (pl.sn_parent[s] >= 0) pl.child_ptr[pl.sn_parent[s] + 1]++;
  for (int s = 0; s < ns; s++) pl.child_ptr[s + 1] += pl.child_ptr[s];
  pl.child_list.resize(pl.child_ptr[ns]);
  {
    std::vector<int> f(pl.child_ptr.begin(), pl.child_ptr.end() - 1);
    for (int s = 0; s < ns; s++)
      if (pl.sn_parent[s] >= 0) pl.child_list[f[pl.sn_parent[s]]++] = s;
  }
  pl.rows_ptr.assign(ns + 1, 0);
  std::vector<int> mark(n, -1), buf;
  for (int s = 0; s < ns; s++) {
    int a0 = pl.sn_first[s], e0 = pl.sn_first[s + 1];
    buf.clear();
    for (int j = a0; j < e0; j++)
      for (int q = cp[j]; q < cp[j + 1]; q++) {
        int r = ri[q];
        if (r >= e0 && mark[r] != s) { mark[r] = s; buf.push_back(r); }
      }
    for (int t = pl.child_ptr[s]; t < pl.child_ptr[s + 1]; t++) {
      int c = pl.child_list[t];
      for (int q = pl.rows_ptr[c]; q < pl.rows_ptr[c + 1]; q++) {
        int r = pl.rows[q];
        if (r >= e0 && mark[r] != s) { mark[r] = s; buf.push_back(r); }
      }
    }
    std::sort(buf.begin(), buf.end());
    HYKKT_REQUIRE(static_cast<int64_t>(buf.size()) == cc[e0 - 1] - 1, HYKKT_E_STATE,
                  "chol_analyze: supernode structure inconsistent with column counts");
    pl.rows.insert(pl.rows.end(), buf.begin(), buf.end());
    pl.rows_ptr[s + 1] = static_cast<int>(pl.rows.size());
  }

  // ---- 8. levels (height), offsets, relative indices ----
  pl.level.assign(ns, 0);
  int nlev = 0;
  for (int s = 0; s < ns; s++) {
    for (int t = pl.child_ptr[s]; t < pl.child_ptr[s + 1]; t++)
      pl.level[s] = std::max(pl.level[s], pl.level[pl.child_list[t]] + 1);
    nlev = std::max(nlev, pl.level[s] + 1);
  }
  pl.level_ptr.assign(nlev + 1, 0);
  for (int s = 0; s < ns; s++) pl.level_ptr[pl.level[s] + 1]++;
  for (int l = 0; l < nlev; l++) pl.level_ptr[l + 1] += pl.level_ptr[l];
  pl.level_sn.resize(ns);
  {
    std::vector<int> f(pl.level_ptr.begin(), pl.level_ptr.end() - 1);
    for (int s = 0; s < ns; s++) pl.level_sn[f[pl.level[s]]++] = s;
  }
  pl.l_off.assign(ns + 1, 0);
  pl.nnz_l_stored = 0;
  pl.max_front = 0;
  for (int s = 0; s < ns; s++) {
    int64_t nc = pl.nc(s), nr = pl.nr(s), f = nc + nr;
    pl.l_off[s + 1] = pl.l_off[s] + f * nc;
    pl.nnz_l_stored += nc * (nc + 1) / 2 + nc * nr;
    pl.max_front = std::max<int>(pl.max_front, static_cast<int>(f));
  }
  assign_update_space(pl);
  pl.rel.resize(pl.rows.size());
  for (int s = 0; s < ns; s++) {
    int p = pl.sn_parent[s];
    if (p < 0) {
      HYKKT_REQUIRE(pl.nr(s) == 0, HYKKT_E_STATE, "chol_analyze: root with off-diagonal rows");
      continue;
    }
    int ap = pl.sn_first[p], ep = pl.sn_first[p + 1], ncp = ep - ap;
    const int* rb = pl.rows.data() + pl.rows_ptr[p];
    const int* re = pl.rows.data() + pl.rows_ptr[p + 1];
    for (int q = pl.rows_ptr[s]; q < pl.rows_ptr[s + 1]; q++) {
      int r = pl.rows[q];
      if (r < ep) {
        HYKKT_REQUIRE(r >= ap, HYKKT_E_STATE, "chol_analyze: row below parent front");
        pl.rel[q] = r - ap;
      } else {
        const int* it = std::lower_bound(rb, re, r);
        HYKKT_REQUIRE(it != re && *it == r, HYKKT_E_STATE, "chol_analyze: row not in parent");
        pl.rel[q] = ncp + static_cast<int>(it - rb);
      }
    }
  }

  // ---- 9. assembly targets of the input's lower-triangle entries ----
  pl.a_src.resize(nlow);
  pl.a_dst.resize(nlow);
  for (int c = 0; c < n; c++) {
    int s = pl.sn_of[c], a0 = pl.sn_first[s], e0 = pl.sn_first[s + 1];
    int64_t f = pl.nc(s) + pl.nr(s);
    const int* rb = pl.rows.data() + pl.rows_ptr[s];
    const int* re = pl.rows.data() + pl.rows_ptr[s + 1];
    for (int q = cp[c]; q < cp[c + 1]; q++) {
      int r = ri[q], lr;
      if (r < e0) lr = r - a0;
      else lr = (e0 - a0) + static_cast<int>(std::lower_bound(rb, re, r) - rb);
      pl.a_src[q] = src[q];
      pl.a_dst[q] = pl.l_off[s] + static_cast<int64_t>(c - a0) * f + lr;
    }
  }
  return pl;
}

}  // namespace hykkt
