// mmio.hpp -- MatrixMarket ingest with the reference's semantics
// (src/input_functions.cpp:16-217).
#pragma once

#include <string>
#include <vector>

namespace hykkt {

struct CsrHost {
  int n = 0, m = 0;
  std::vector<int> rp, ci;
  std::vector<double> v;
  int nnz() const { return rp.empty() ? 0 : rp.back(); }
};

// read_mm_file_into_coo + sym_coo_to_csr (symmetric=true) or coo_to_csr.
// The last of the `skip` skipped lines holds "rows cols nnz".
CsrHost read_mm_csr(const std::string& path, int skip, bool symmetric);
// read_rhs: skips two lines, then one value per line.
std::vector<double> read_rhs(const std::string& path);

}  // namespace hykkt
