// solution_dump.hpp -- optional output of the example drivers: when the
// environment variable HYKKT_DRIVER_SOLUTION_DIR names a directory, each
// solve's solution blocks and the ordering used are written there as raw
// little-endian arrays, <tag>_dx.bin, <tag>_ds.bin, <tag>_dyc.bin,
// <tag>_dyd.bin (float64) and <tag>_perm.bin (int32), so that a test can
// compare them with a host solve of the same system.  The reference's
// drivers print residuals only (their copy-back is disabled,
// HykktSolver.cpp:451-456).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/HykktSolver.hpp"

inline bool dump_solution(const HykktSolver& hs, const std::string& tag)
{
  const char* dir = std::getenv("HYKKT_DRIVER_SOLUTION_DIR");
  if (!dir || !*dir) return true;
  const hykkt_stats st = hs.stats();
  std::vector<double> dx(st.n), ds(st.d), dyc(st.m), dyd(st.d);
  std::vector<int> perm(st.n);
  hs.get_solution(dx.data(), ds.data(), dyc.data(), dyd.data());
  if (hykkt_get_permutation(hs.handle(), perm.data(), st.n) != HYKKT_OK) return false;
  auto put = [&](const char* what, const void* p, size_t bytes) {
    const std::string f = std::string(dir) + "/" + tag + "_" + what + ".bin";
    FILE* fp = std::fopen(f.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(p, 1, bytes, fp) == bytes;
    return std::fclose(fp) == 0 && ok;
  };
  return put("dx", dx.data(), dx.size() * 8) && put("ds", ds.data(), ds.size() * 8) &&
         put("dyc", dyc.data(), dyc.size() * 8) && put("dyd", dyd.data(), dyd.size() * 8) &&
         put("perm", perm.data(), perm.size() * 4);
}
