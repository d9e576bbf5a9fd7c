// common.hpp -- shared host-side types and error handling for the HyKKT
// MI355X library.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hykkt.h"

namespace hykkt {

// Error carried up to the C-ABI, which turns it into a return code and a
// message retrievable with hykkt_last_error() (the reference instead calls
// exit(EXIT_FAILURE), cuda_check_errors.hpp:3-14).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// Error codes: HYKKT_OK, HYKKT_E_* from the C ABI header.


#define HYKKT_REQUIRE(cond, code, msg)                     \
  do {                                                     \
    if (!(cond)) throw ::hykkt::Error((code), (msg));      \
  } while (0)

// Host CSR pattern (0-based, int32 indices, sorted columns per row).
struct Pattern {
  int n = 0;   // rows
  int m = 0;   // cols
  std::vector<int> rp;
  std::vector<int> ci;
  int nnz() const { return rp.empty() ? 0 : rp.back(); }
};

}  // namespace hykkt
