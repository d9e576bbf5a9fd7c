// symbolic.cpp -- transposes, SpGEMM term lists and the reference's
// permutation maps (host, once per sparsity pattern).
#include "symbolic.hpp"

#include <algorithm>
#include <numeric>

namespace hykkt {

TransposeMap transpose_map(const Pattern& a)
{
  TransposeMap t;
  t.at.n = a.m;
  t.at.m = a.n;
  t.at.rp.assign(static_cast<size_t>(a.m) + 1, 0);
  const int nnz = a.nnz();
  for (int p = 0; p < nnz; p++) t.at.rp[a.ci[p] + 1]++;
  for (int j = 0; j < a.m; j++) t.at.rp[j + 1] += t.at.rp[j];
  t.at.ci.resize(nnz);
  t.map.resize(nnz);
  std::vector<int> next(t.at.rp.begin(), t.at.rp.end() - 1);
  for (int i = 0; i < a.n; i++) {
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      int q = next[a.ci[p]]++;
      t.at.ci[q] = i;
      t.map[q] = p;
    }
  }
  return t;
}

SpgemmPlan spgemm_plan(const Pattern& a, const std::vector<int>& a_pos, const Pattern& b,
                       const Pattern& d)
{
  HYKKT_REQUIRE(a.m == b.n && d.n == a.n && d.m == b.m, HYKKT_E_ARG, "spgemm shape mismatch");
  SpgemmPlan pl;
  const int n = a.n, ncols = b.m;
  pl.e.n = n;
  pl.e.m = ncols;
  pl.e.rp.assign(static_cast<size_t>(n) + 1, 0);
  std::vector<int> mark(ncols, -1), rowbuf;
  // pass 1: pattern (union of D's row and the structural product row)
  for (int i = 0; i < n; i++) {
    rowbuf.clear();
    for (int p = d.rp[i]; p < d.rp[i + 1]; p++)
      if (mark[d.ci[p]] != i) { mark[d.ci[p]] = i; rowbuf.push_back(d.ci[p]); }
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      int k = a.ci[p];
      for (int q = b.rp[k]; q < b.rp[k + 1]; q++)
        if (mark[b.ci[q]] != i) { mark[b.ci[q]] = i; rowbuf.push_back(b.ci[q]); }
    }
    std::sort(rowbuf.begin(), rowbuf.end());
    pl.e.ci.insert(pl.e.ci.end(), rowbuf.begin(), rowbuf.end());
    pl.e.rp[i + 1] = static_cast<int>(pl.e.ci.size());
  }
  const int nnz = pl.e.nnz();
  pl.d_src.assign(nnz, -1);
  pl.term_ptr.assign(static_cast<size_t>(nnz) + 1, 0);
  std::vector<int> pos(ncols, -1);
  // pass 2: count terms per output entry, D sources
  for (int i = 0; i < n; i++) {
    for (int p = pl.e.rp[i]; p < pl.e.rp[i + 1]; p++) pos[pl.e.ci[p]] = p;
    for (int p = d.rp[i]; p < d.rp[i + 1]; p++) pl.d_src[pos[d.ci[p]]] = p;
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      int k = a.ci[p];
      for (int q = b.rp[k]; q < b.rp[k + 1]; q++) pl.term_ptr[pos[b.ci[q]] + 1]++;
    }
  }
  for (int p = 0; p < nnz; p++) pl.term_ptr[p + 1] += pl.term_ptr[p];
  const int nterms = pl.term_ptr[nnz];
  pl.ta.resize(nterms);
  pl.tb.resize(nterms);
  std::vector<int> fill(pl.term_ptr.begin(), pl.term_ptr.end() - 1);
  // pass 3: fill terms in increasing k (A's row is sorted by k)
  for (int i = 0; i < n; i++) {
    for (int p = pl.e.rp[i]; p < pl.e.rp[i + 1]; p++) pos[pl.e.ci[p]] = p;
    for (int p = a.rp[i]; p < a.rp[i + 1]; p++) {
      int k = a.ci[p];
      for (int q = b.rp[k]; q < b.rp[k + 1]; q++) {
        int t = fill[pos[b.ci[q]]]++;
        pl.ta[t] = a_pos[p];
        pl.tb[t] = q;
      }
    }
  }
  return pl;
}

// ------------------------------------------------------------------ maps
void perm_reverse(const std::vector<int>& perm, std::vector<int>& rev)
{
  rev.assign(perm.size(), -1);
  for (size_t i = 0; i < perm.size(); i++) rev[perm[i]] = static_cast<int>(i);
}

// Sort (col, map) pairs of one row by col.  Columns within a row are
// distinct, so the result equals the reference's insertion sort.
static void sort_pairs(int* c, int* m, int len)
{
  if (len < 2) return;
  std::vector<std::pair<int, int>> t(len);
  for (int i = 0; i < len; i++) t[i] = {c[i], m[i]};
  std::sort(t.begin(), t.end());
  for (int i = 0; i < len; i++) { c[i] = t[i].first; m[i] = t[i].second; }
}

// permcheck.cu:162-195 (make_vec_map_rc)
void perm_map_rc(const Pattern& a, const std::vector<int>& perm, const std::vector<int>& rev,
                 Pattern& pa, std::vector<int>& map)
{
  const int n = a.n;
  pa.n = n;
  pa.m = a.m;
  pa.rp.assign(static_cast<size_t>(n) + 1, 0);
  pa.ci.resize(a.nnz());
  map.resize(a.nnz());
  int cnt = 0;
  for (int i = 0; i < n; i++) {
    int src = perm[i], s = a.rp[src], len = a.rp[src + 1] - s;
    pa.rp[i + 1] = pa.rp[i] + len;
    for (int j = 0; j < len; j++) {
      map[cnt + j] = s + j;
      pa.ci[cnt + j] = rev[a.ci[s + j]];
    }
    sort_pairs(&pa.ci[cnt], &map[cnt], len);
    cnt += len;
  }
}

// permcheck.cu:99-124 (make_vec_map_c)
void perm_map_c(const Pattern& a, const std::vector<int>& rev, Pattern& pa,
                std::vector<int>& map)
{
  pa.n = a.n;
  pa.m = a.m;
  pa.rp = a.rp;
  pa.ci.resize(a.nnz());
  map.resize(a.nnz());
  for (int i = 0; i < a.n; i++) {
    int s = a.rp[i], len = a.rp[i + 1] - s;
    for (int j = 0; j < len; j++) {
      map[s + j] = s + j;
      pa.ci[s + j] = rev[a.ci[s + j]];
    }
    sort_pairs(&pa.ci[s], &map[s], len);
  }
}

// permcheck.cu:134-160 (make_vec_map_r)
void perm_map_r(const Pattern& a, const std::vector<int>& perm, Pattern& pa,
                std::vector<int>& map)
{
  const int n = a.n;
  pa.n = n;
  pa.m = a.m;
  pa.rp.assign(static_cast<size_t>(n) + 1, 0);
  pa.ci.resize(a.nnz());
  map.resize(a.nnz());
  int cnt = 0;
  for (int i = 0; i < n; i++) {
    int src = perm[i], s = a.rp[src], len = a.rp[src + 1] - s;
    pa.rp[i + 1] = pa.rp[i] + len;
    for (int j = 0; j < len; j++) {
      map[cnt + j] = s + j;
      pa.ci[cnt + j] = a.ci[s + j];
    }
    cnt += len;
  }
}

}  // namespace hykkt
