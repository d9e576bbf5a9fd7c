// ref_harness.cpp -- extern "C" entry points over the reference's own host
// sources (src/input_functions.cpp, src/MMatrix.cpp, src/permcheck.cu), which
// oracle/Makefile compiles where they lie into oracle/_ref/libhykkt_ref.so.
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's ingest and
// permutation-map restatements against the reference code itself.
#include <cstring>
#include "MMatrix.hpp"
#include "input_functions.hpp"
#include "permcheck.hpp"

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

// permcheck.cu:99-195
void ref_reverse_perm(int n, int* perm, int* rev) { reverse_perm(n, perm, rev); }
void ref_map_rc(int n, int* rows, int* cols, int* perm, int* rev, int* prows, int* pcols,
                int* pmap)
{
  make_vec_map_rc(n, rows, cols, perm, rev, prows, pcols, pmap);
}
void ref_map_c(int n, int* rows, int* cols, int* rev, int* pcols, int* pmap)
{
  make_vec_map_c(n, rows, cols, rev, pcols, pmap);
}
void ref_map_r(int n, int* rows, int* cols, int* perm, int* prows, int* pcols, int* pmap)
{
  make_vec_map_r(n, rows, cols, perm, prows, pcols, pmap);
}
}
