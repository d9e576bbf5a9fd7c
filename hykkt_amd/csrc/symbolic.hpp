// symbolic.hpp -- once-per-pattern host analysis: transposes, SpGEMM term
// lists, fill-reducing ordering, permutation maps, supernodal Cholesky plan.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace hykkt {

// Pattern of A^T with a value map: at row j lists the rows i with A(i,j) != 0
// in increasing i; at_v[q] = a_v[map[q]].  Same semantics as the csr2csc
// transpose the reference runs every solve (cusparse_utils.cpp:49-98).
struct TransposeMap {
  Pattern at;
  std::vector<int> map;
};
TransposeMap transpose_map(const Pattern& a);

// Numeric plan for E = alpha_s*D + beta_s*(alpha_p*A*B) on E's structural
// pattern (SpgemmClass.hpp:4-5; matrix_matrix_ops.cpp:13-328).  For output
// entry p: d_src[p] is the position in D (or -1), and terms
// [term_ptr[p], term_ptr[p+1]) are (position in A, position in B) pairs in
// increasing inner index k.  The value is acc = fma(A, B, acc) over the terms.
struct SpgemmPlan {
  Pattern e;
  std::vector<int> d_src;
  std::vector<int> term_ptr;
  std::vector<int> ta;
  std::vector<int> tb;
};
// A is given as the transpose map of B^T-shaped data: a_pat (n x k) with
// a_pos[q] giving, for the q-th entry of a_pat, the position used for the
// A operand value (lets A = B^T share B's value array).
SpgemmPlan spgemm_plan(const Pattern& a_pat, const std::vector<int>& a_pos, const Pattern& b,
                       const Pattern& d);

// Approximate minimum degree ordering of a symmetric pattern (diagonal
// entries ignored).  Written from Amestoy, Davis & Duff, "An approximate
// minimum degree ordering algorithm", SIAM J. Matrix Anal. Appl. 17(4), 1996:
// quotient graph of variables and elements, element absorption (including
// aggressive absorption), supervariable detection by hashing, approximate
// external degrees.  Returns perm with perm[k] = k-th eliminated variable.
std::vector<int> amd_order(const Pattern& a);

// Reference permutation maps (permcheck.cu:99-195, PermClass.cpp:101-123).
void perm_reverse(const std::vector<int>& perm, std::vector<int>& rev);
void perm_map_rc(const Pattern& a, const std::vector<int>& perm, const std::vector<int>& rev,
                 Pattern& pa, std::vector<int>& map);
void perm_map_c(const Pattern& a, const std::vector<int>& rev, Pattern& pa,
                std::vector<int>& map);
void perm_map_r(const Pattern& a, const std::vector<int>& perm, Pattern& pa,
                std::vector<int>& map);

// ---------------------------------------------------------------------------
// Supernodal multifrontal Cholesky plan for a symmetric matrix given by its
// lower triangle.  Internal indices are a postorder of the elimination tree.
// Front of supernode s: index list [cols(s); rows(s)] (f = nc + nr entries).
// Factor panel: f x nc column-major at l_off[s] (lower part used).
// Update matrix: nr x nr lower triangle at u_off[s], packed by columns
// (column c holds rows c..nr-1 at offset c*nr - c(c-1)/2).
struct CholPlan {
  int n = 0;
  int ns = 0;
  std::vector<int> post;        // internal k -> input index
  std::vector<int> ipost;       // input index -> internal k
  std::vector<int> sn_first;    // ns+1, columns of s: [sn_first[s], sn_first[s+1])
  std::vector<int> sn_of;       // column -> supernode
  std::vector<int> sn_parent;   // -1 for roots
  std::vector<int> rows_ptr;    // ns+1
  std::vector<int> rows;        // off-diagonal-block row indices (internal), ascending
  std::vector<int> rel;         // per entry of rows: position in the parent's front
  std::vector<int> child_ptr;   // ns+1
  std::vector<int> child_list;  // children in increasing order
  std::vector<int64_t> l_off;   // ns+1 (doubles)
  std::vector<int64_t> u_off;   // ns+1 (doubles): U(s) offsets (reused space), [ns] = pool size
  std::vector<int> level;       // height in the supernodal tree (leaves 0)
  std::vector<int> level_ptr;   // nlevels+1
  std::vector<int> level_sn;    // supernodes grouped by level
  // lower-triangle entries of the input matrix, as (internal row, internal col)
  // with their input CSR position and the factor-storage offset they land at
  std::vector<int> a_src;       // input position of each assembled entry
  std::vector<int64_t> a_dst;   // offset in the factor storage
  int64_t nnz_l_true = 0;       // nnz(L) without amalgamation zeros
  int64_t nnz_l_stored = 0;     // sum of nc*(nc+1)/2 + nr*nc
  double flops = 0.0;           // sum over columns of colcount^2 (true)
  int max_front = 0;
  int nc(int s) const { return sn_first[s + 1] - sn_first[s]; }
  int nr(int s) const { return rows_ptr[s + 1] - rows_ptr[s]; }
};

struct CholOptions {
  int relax_small = 8;          // always merge while the merged supernode has <= this many cols
  double relax_frac_mid = 0.5;  // zero fraction allowed up to relax_mid_cols columns
  int relax_mid_cols = 32;
  double relax_frac_big = 0.05; // zero fraction allowed beyond that
};

// Analyse the lower triangle of the n x n symmetric matrix `a` (CSR, sorted;
// entries with col <= row are used).  a_src of the plan holds CSR positions in a.
CholPlan chol_analyze(const Pattern& a, const CholOptions& opt);

}  // namespace hykkt
