// mmio.cpp -- MatrixMarket ingest (host).  Same semantics as the reference's
// input_functions.cpp: `skip` lines are skipped and the last of them is the
// size line (:26-31); entries are 1-based "row col value" (:44-50); a
// symmetric file has each off-diagonal entry mirrored (:54-119); rows are
// sorted by column (:106-111, :158-164); vectors skip 2 lines (:198-217).
#include "mmio.hpp"

#include <algorithm>
#include <cstdio>

#include "common.hpp"

namespace hykkt {

namespace {

void sort_rows(CsrHost& a)
{
  std::vector<std::pair<int, double>> t;
  for (int i = 0; i < a.n; i++) {
    int b = a.rp[i], e = a.rp[i + 1];
    t.clear();
    for (int p = b; p < e; p++) t.push_back({a.ci[p], a.v[p]});
    std::stable_sort(t.begin(), t.end(),
                     [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
                       return x.first < y.first;
                     });
    for (int p = b; p < e; p++) {
      a.ci[p] = t[p - b].first;
      a.v[p] = t[p - b].second;
    }
  }
}

}  // namespace

CsrHost read_mm_csr(const std::string& path, int skip, bool symmetric)
{
  FILE* f = std::fopen(path.c_str(), "r");
  HYKKT_REQUIRE(f, HYKKT_E_IO, "cannot open " + path);
  char line[256];
  line[0] = 0;
  for (int i = 0; i < skip; i++) {
    if (!std::fgets(line, sizeof line, f)) {
      std::fclose(f);
      throw Error(HYKKT_E_IO, "unexpected end of file in header of " + path);
    }
  }
  int n = 0, m = 0, nnz = 0;
  if (std::sscanf(line, "%d %d %d", &n, &m, &nnz) != 3) {
    std::fclose(f);
    throw Error(HYKKT_E_IO, "no size line at line " + std::to_string(skip) + " of " + path);
  }
  std::vector<int> r, c;
  std::vector<double> v;
  r.reserve(nnz);
  c.reserve(nnz);
  v.reserve(nnz);
  while (std::fgets(line, sizeof line, f)) {
    int ri, cj;
    double x;
    if (std::sscanf(line, "%d %d %lf", &ri, &cj, &x) != 3) continue;
    if (ri < 1 || ri > n || cj < 1 || cj > m) {
      std::fclose(f);
      throw Error(HYKKT_E_IO, "entry out of range in " + path);
    }
    r.push_back(ri - 1);
    c.push_back(cj - 1);
    v.push_back(x);
  }
  std::fclose(f);
  HYKKT_REQUIRE(static_cast<int>(r.size()) == nnz, HYKKT_E_IO,
                "entry count differs from the size line in " + path);
  CsrHost a;
  a.n = n;
  a.m = m;
  a.rp.assign(static_cast<size_t>(n) + 1, 0);
  for (int i = 0; i < nnz; i++) {
    a.rp[r[i] + 1]++;
    if (symmetric && r[i] != c[i]) a.rp[c[i] + 1]++;
  }
  for (int i = 0; i < n; i++) a.rp[i + 1] += a.rp[i];
  a.ci.resize(a.rp[n]);
  a.v.resize(a.rp[n]);
  std::vector<int> fill(a.rp.begin(), a.rp.end() - 1);
  for (int i = 0; i < nnz; i++) {
    int p = fill[r[i]]++;
    a.ci[p] = c[i];
    a.v[p] = v[i];
    if (symmetric && r[i] != c[i]) {
      p = fill[c[i]]++;
      a.ci[p] = r[i];
      a.v[p] = v[i];
    }
  }
  sort_rows(a);
  return a;
}

std::vector<double> read_rhs(const std::string& path)
{
  FILE* f = std::fopen(path.c_str(), "r");
  HYKKT_REQUIRE(f, HYKKT_E_IO, "cannot open " + path);
  char line[256];
  for (int i = 0; i < 2; i++) {
    if (!std::fgets(line, sizeof line, f)) {
      std::fclose(f);
      throw Error(HYKKT_E_IO, "unexpected end of file in " + path);
    }
  }
  std::vector<double> out;
  while (std::fgets(line, sizeof line, f)) {
    double x;
    if (std::sscanf(line, "%lf", &x) == 1) out.push_back(x);
  }
  std::fclose(f);
  return out;
}

}  // namespace hykkt
