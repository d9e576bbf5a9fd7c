// ref_harness.cpp -- extern "C" entry points over the reference's own host
// sources (src/input_functions.cpp, src/MMatrix.cpp), which oracle/Makefile
// compiles where they lie into oracle/_ref/libhykkt_ref.so.
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's and the
// product's MatrixMarket ingest against the reference code itself.
#include <cstring>
#include "MMatrix.hpp"
#include "input_functions.hpp"

extern "C" {

// read_mm_file_into_coo + (sym_)coo_to_csr (input_functions.cpp:16-175).
// Pass arrays of at least n+1 / nnz_cap; returns nnz of the CSR or <0.
int ref_read_csr(const char* path, int skip, int symmetric, int* n, int* m, int cap,
                 int* rp, int* ci, double* cv)
{
  MMatrix a;
  read_mm_file_into_coo(path, a, skip);
  if (symmetric) sym_coo_to_csr(a); else coo_to_csr(a);
  *n = a.n_;
  *m = a.m_;
  if (a.nnz_ > cap) return -1;
  std::memcpy(rp, a.csr_rows, sizeof(int) * (a.n_ + 1));
  std::memcpy(ci, a.csr_cols, sizeof(int) * a.nnz_);
  std::memcpy(cv, a.csr_vals, sizeof(double) * a.nnz_);
  return a.nnz_;
}

// read_rhs (input_functions.cpp:198-217).
void ref_read_rhs(const char* path, double* out) { read_rhs(path, out); }

}
