// kernels.hpp -- launchers for the gfx950 HIP kernels (device pointers only).
//
// Systems of a launch.  Every per-system launcher takes a Batch: `count`
// systems that share one sparsity pattern, whose per-system buffers (values,
// vectors, factor storage) sit `vs` doubles apart -- system k's copy of a
// buffer p is p + k*vs -- while index arrays (patterns, maps, task lists) are
// shared.  System k runs in blockIdx.y == k (or, for the triangular sweeps, in
// every count-th item), with exactly the arithmetic it gets alone, so a
// system's results do not depend on the batch it is in.  `mask`, when not
// null, holds one int per system; systems with 0 are skipped.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace hykkt {
namespace dev {

void check(hipError_t e, const char* what);
#define HIPCHK(x) ::hykkt::dev::check((x), #x)

struct Batch {
  int count = 1;
  int64_t vs = 0;             // per-system stride of the buffers (doubles)
  const int* mask = nullptr;  // per system: 0 = skip
};
inline Batch masked(const Batch& b, const int* m) { return Batch{b.count, b.vs, m}; }

// ---------------------------------------------------------------- vectors
void fill(double* p, double v, int64_t n, const Batch& b, hipStream_t st);
// p[i] = v for i < n, one array (no batch)
void fill_i(int* p, int v, int64_t n, hipStream_t st);
void copy(double* dst, const double* src, int64_t n, const Batch& b, hipStream_t st);
// dst[k] <- src[k] for each system k with keep[k] != 0 (both per-system)
void copy_where(double* dst, const double* src, int64_t n, const int* keep, const Batch& b,
                hipStream_t st);
// dst[i] = src[map[i]]  (map_idx, permcheck.cu:12-19; map shared)
void gather(double* dst, const double* src, const int* map, int64_t n, const Batch& b,
            hipStream_t st);
// dst[i] = src[map[i]] with 64-bit positions (one system)
void gather64(double* dst, const double* src, const int64_t* map, int64_t n, hipStream_t st);
// dst[i] = src[map[i]] * scale[i]
void gather_scale(double* dst, const double* src, const int* map, const double* scale, int n,
                  const Batch& b, hipStream_t st);
// The recovery's two scaled results in one launch: dx[i] = (h0[map[i]] -
// cz[map[i]]) * max_d[i] for i < n, dyc[i] = cx[i] * max_d[n + i] for i < m.
void recover_xy(double* dx, const double* h0, const double* cz, const int* map,
                const double* max_d, int n, double* dyc, const double* cx, int m, const Batch& b,
                hipStream_t st);
// y[i] *= s[i]  (vec_scale, matrix_vector_ops.cu:385-392)
void vec_scale(double* y, const double* s, int n, const Batch& b, hipStream_t st);
// out = a + b
// out = a - b (per system)
void sub(double* out, const double* a, const double* bv, int64_t n, const Batch& b,
         hipStream_t st);
// CG tracking: q = beta q + t, z += alpha q (alpha, beta from the CG scalars sc)
void cg_track(int n, const double* sc, double* q, double* z, const double* t, const Batch& b,
              hipStream_t st);
void add(double* out, const double* a, const double* bv, int64_t n, const Batch& b,
         hipStream_t st);
// out[k*nslot + j] = base[k*vs + j] for j < nslot and each system k (b.count
// systems, mask ignored): per-system scalars gathered for one readback.
void collect(double* out, const double* base, int nslot, const Batch& b, hipStream_t st);
// Per system k: the arrays src[k*nsrc + j] (j < nsrc, device pointers to
// caller memory, lengths len[j]) copied to dst[j] + k*vs (arena slots).
void copy_in(int nsrc, double* const* dst, const double* const* src_tab, const int64_t* len,
             const Batch& b, hipStream_t st);

// ---------------------------------------------------------------- sparse
// jd_vs = Ds-row-scaled Jd (skipped if jd_vs is null); ryd_s = ryd*ds + rs
// (row_scale :490-509, add_vecs :310-317)
void row_scale(int d, const int* jd_rp, const double* jd_v, const double* ds,
               const double* ryd, const double* rs, double* jd_vs, double* ryd_s,
               const Batch& b, hipStream_t st);
// Row blocks of a CSR matrix for the LDS-staged SpMV: consecutive rows with
// at most spmv_lds_entries() entries together (a longer row is a block of its
// own); blocks[k] = first row of block k, blocks[nblk] = nrows.
int spmv_lds_entries();
// y_out[i] = alpha*t + beta*y_in[i] (beta == 0: alpha*t),
// t = sum over the row of fma(v[map ? map[q] : q], x[ci[q]], t) in column
// order.  The row block's values and x entries are staged through LDS with
// coalesced loads, then each row is summed by one thread in that order.
// alpha_ps (may be null): per-system alpha (arena slot) instead of alpha.
// dot_w (may be null): also dot_part[blk] = sum over the block's rows of
// y_out[i] * dot_w[i] (per-system, for a fused w.r).
void spmv(int nrows, const int* rp, const int* ci, const int* map, const double* v, double alpha,
          const double* alpha_ps, const double* x, double beta, const double* y_in,
          double* y_out, const int* blocks, int nblk, const Batch& b, hipStream_t st,
          const double* dot_w = nullptr, double* dot_part = nullptr);
// out[p] = d_v[d_src[p]] (or 0) + alpha_p*acc, acc = sum fma(av[ta], bv[tb]) over terms of p
void spgemm_values(int nnz_out, const int* d_src, const double* d_v, const int* term_ptr,
                   const int* ta, const int* tb, const double* av, const double* bv,
                   double alpha_p, double* out, const Batch& b, hipStream_t st);
// out[ent_dst[e]] = value of entry ent_src[e] (as spgemm_values, alpha_p from
// the per-system slot alpha_ps)
void spgemm_scatter(int nent, const int* ent_src, const int64_t* ent_dst, const int* d_src,
                    const double* d_v, const int* term_ptr, const int* ta, const int* tb,
                    const double* av, const double* bv, const double* alpha_ps, double* out,
                    const Batch& b, hipStream_t st);
// out[dst[e]] = vals[src[e]] for e < n
void scatter(double* out, const int64_t* dst, const double* vals, const int* src, int n,
             const Batch& b, hipStream_t st);
// add `delta` at out[pos[i]] for i < n (add_diag hook, matrix_vector_ops.cu:344-358)
void add_at(double* out, const int64_t* pos, int n, double delta, const Batch& b,
            hipStream_t st);

// ---------------------------------------------------------------- Ruiz
// adapt_row_max (matrix_vector_ops.cu:218-266). Jt rows read as jt_v[jt_map ? jt_map[q] : q].
// zero_row (may be null; one int per system): receives the smallest row index
// whose entries are all zero (its scale is 1/sqrt(0)); the caller initialises it.
void ruiz_rowmax(int n, int totn, const int* h_rp, const double* h_v, const int* jt_rp,
                 const int* jt_map, const double* jt_v, const int* j_rp, const double* j_v,
                 double* scale, int* zero_row, const Batch& b, hipStream_t st);
// adapt_diag_scale (matrix_vector_ops.cu:146-186); jt_v_scaled may be null (Jt shares J).
// first != 0: max_d is written as 1*scale (reset, :268-283, fused).
void ruiz_scale(int n, int totn, const int* h_rp, const int* h_ci, double* h_v,
                const int* jt_rp, const int* jt_ci, double* jt_v_scaled,
                const int* j_rp, const int* j_ci, double* j_v, const double* scale,
                double* rhs1, double* rhs2, double* max_d, int first, const Batch& b,
                hipStream_t st);

// ---------------------------------------------------------------- reductions
// Block partial sums of a length-n reduction: reduce_blocks(n) blocks, a
// function of n only (the same for every batch).
int reduce_blocks(int64_t n);
// Deterministic sum of a[i]*b[i] (bv == null: sum of a[i]); result to *out.
// `part` needs reduce_blocks(n) doubles per system.
void dot(int n, const double* a, const double* bv, double* part, double* out, const Batch& b,
         hipStream_t st);
// Up to DOTS_MAX such sums in two launches, out[k] per system for sum k
// (same bits as dot() each); part needs the sum of their reduce_blocks(n[k])
// doubles per system.  off is filled in by dots().
constexpr int DOTS_MAX = 8;
struct DotSet {
  const double* a[DOTS_MAX];
  const double* b[DOTS_MAX];
  int n[DOTS_MAX];
  int off[DOTS_MAX + 1];
  int count = 0;
};
void dots(DotSet d, double* part, double* out, const Batch& b, hipStream_t st);
// Per system k (b.count systems, stride b.vs between sc slots): sc[0], sc[1],
// iters[k], status[k] as out[4k .. 4k+3].
void cg_pack(double* out, const double* sc, const int* iters, const int* status, const Batch& b,
             hipStream_t st);
// out[0] = *w, then a[0..count) and bv[0..count) where not null.
void collect_ints(int* out, const int* w, const int* a, const int* bv, int count, hipStream_t st);
// The first stage of dot(): reduce_blocks(n) partials per system in part.
void dot_partial(int n, const double* a, const double* bv, double* part, const Batch& b,
                 hipStream_t st);

// ---------------------------------------------------------------- CG
// CG scalars per system: sc[0]=gam [1]=gam1 [2]=delta [3]=alpha [4]=beta.
// Per-system ints (one array each, indexed by system): live (1 while the
// system iterates), iters, status.
struct CgFlags {
  int* live;
  int* iters;
  int* status;
};
// p = beta p + r; s = beta s + w; x += alpha p; r -= alpha s, and the block
// partials of r.r in part (reduce_blocks(m) per system)
void cg_update_dot(int m, const double* sc, double* p, double* s, double* x, double* r,
                   const double* w, double* part, const Batch& b, hipStream_t st);
// sc[slot] = sum of the np partials in part; with check (slot 1, iteration `it`): the
// stopping test of SchurComplementConjugateGradient.cpp:164-170 -- sqrt(gam1)
// < tol ends the system (iters it+1, status 0), it+1 == itmax ends it with
// status 1.  With it < 0 (the start, slot 0): r.r == 0 ends it (iters 0).
// With scalars (slot 2): then beta = gam1/gam, gam = gam1,
// alpha = gam/(delta - beta gam/alpha) (:181-183), or alpha = gam/delta for
// it < 0 (:143).
void dot_final(int np, double* part, double* sc, int slot, bool check, bool scalars, int it,
               double tol, int itmax, const CgFlags& f, const Batch& b, hipStream_t st);

// ---------------------------------------------------------------- recover / residual
// Residual pieces, squares written per element for dot():
//   n rows: r1 = rx - H dx - Jc^T dyc - Jd^T dyd   (Jc^T, Jd^T via transposes + maps)
void resid_rows_n(int n, const int* h_rp, const int* h_ci, const double* h_v,
                  const int* jct_rp, const int* jct_ci, const int* jct_map, const double* jc_v,
                  const int* jdt_rp, const int* jdt_ci, const int* jdt_map, const double* jd_v,
                  const double* rx, const double* dx, const double* dyc, const double* dyd,
                  double* sq, double* e, const Batch& b, hipStream_t st);
//   m rows: r3 = ry - Jc dx
void resid_rows_m(int m, const int* jc_rp, const int* jc_ci, const double* jc_v, const double* ry,
                  const double* dx, double* sq, double* e, const Batch& b, hipStream_t st);
//   d rows: r2 = rs - (Ds ds - dyd), r4 = ryd - (Jd dx - ds)
// e, e2, e4 (may be null): the signed residual rows, for iterative refinement.
void resid_rows_d(int d, const int* jd_rp, const int* jd_ci, const double* jd_v, int has_jd,
                  const double* ds, const double* rs, const double* ryd, const double* dx,
                  const double* dsv, const double* dyd, double* sq2, double* sq4, double* e2,
                  double* e4, const Batch& b, hipStream_t st);
// ds = Jd dx - ryd (or -ryd); dyd = ds*Ds - rs
void recover_d(int d, const int* jd_rp, const int* jd_ci, const double* jd_v, int has_jd,
               const double* dx, const double* ryd, const double* ds_diag, const double* rs,
               double* dsv, double* dyd, const Batch& b, hipStream_t st);

// ok[s] = (system s's factorization met no failing pivot); a system that
// did gets a zero solution (dx n, dsv d, dyc m, dyd d).
void pivot_split(const int* pivot, int* ok, int n, int m, int d, double* dx, double* dsv,
                 double* dyc, double* dyd, const Batch& b, hipStream_t st);

// ---------------------------------------------------------------- scratch
// The allocation-free entry point's device scratch (noalloc.hip): its size
// in bytes, and its address on the current device (taken once per device).
size_t noalloc_scratch_bytes();
void* noalloc_scratch();

// ---------------------------------------------------------------- Cholesky
// Offset of (r, c), r >= c, in an n x n lower triangle packed by columns.
__host__ __device__ inline int64_t packed_lower(int64_t r, int64_t c, int64_t n)
{
  return c * n - c * (c + 1) / 2 + r;
}

// The factor storage of system 0 (system k's at + k*vs, vs of the Batch).
struct CholDev {
  int n, ns;
  const int* sn_first;      // ns+1
  const int* rows_ptr;      // ns+1
  const int* rows;
  const int* rel;
  const int64_t* l_off;     // ns+1
  const int64_t* u_off;     // ns+1: U(s) offsets in the reused update space
  const int* child_ptr;
  const int* child_list;
  const int* sn_parent;
  double* L;
  double* U;
  int* pivot;               // one int per system: first failing pivot (min), INT_MAX if none
  double pivot_tol;
  double* dinv;             // inverted 128x128 diagonal blocks (column-major)
  const int64_t* dinv_off;  // ns+1, offset of supernode s's first block in dinv
};

// Extend-add tasks: task t = (parent p, parent front column col);
// contributions [cptr[t], cptr[t+1]) are (child, child-local column) pairs.
void extend_add(const CholDev& c, int ntasks, const int* task_p, const int* task_col,
                const int* cptr, const int* contrib_child, const int* contrib_col,
                const Batch& b, hipStream_t st);
// Factor + invert the (<= 128 wide) diagonal block of step `step` of each
// supernode; max_b = largest block in the launch (selects the LDS footprint).
// dinv slots must be zero-initialized once (only the b x b lower part is written).
// Blocks wider than 64 (max_b > 64) leave only the inverses of their 16x16
// diagonal blocks in the slot: their TRSM is trsm(), the whole inverse comes
// from inv_blocks() after the level.
void potrf_inv(const CholDev& c, int nsn, const int* sns, int step, int max_b, const Batch& b,
               hipStream_t st);
// L21 = A21 L_kk^{-T} below the 128-wide diagonal blocks of step `step` by blocked
// substitution (tasks: supernode, tile row of TM = 32 (short_rows) or 128 rows).
void trsm(const CholDev& c, int ntasks, const int* task_sn, const int* task_tr, int step,
          const Batch& b, hipStream_t st, bool short_rows);
// The whole inverses of those blocks (tasks: supernode, step), same bits as formed
// inside potrf_inv.
void inv_blocks(const CholDev& c, int ntasks, const int* task_sn, const int* task_step,
                const Batch& b, hipStream_t st);

// MFMA tile kernel over (supernode, tile row, tile col) tasks.  mode 0: trailing
// panel update after block `step`; mode 1: U = -L21 L21^T; mode 2: TRSM of the
// rows below block `step` with its inverted diagonal block; mode m >= 3: as
// mode 0 by blocks step .. step+m-2.  short_rows: 32-row x 128-column tiles,
// tile rows counted in 32-row units (modes 0, 2, 3+).  done: an event the
// launch records on completion (nullptr: none).
void dense_nt(const CholDev& c, int ntasks, const int* task_sn, const int* task_tr,
              const int* task_tc, int mode, int step, const Batch& b, hipStream_t st,
              bool short_rows = false, hipEvent_t done = nullptr);

// Single-launch supernodal solves over the whole tree.  Work items are
// 128-row blocks of the fronts, dispatched in a dependency-respecting order;
// items synchronise on the values they exchange, whose buffers are filled
// with a sentinel (fill_sentinel) before each sweep.  With several systems,
// launch item i*count + k is item i of system k.
struct SolveDev {
  // Items of the small-supernode forest (fwd_items, bwd_items) and of the top
  // of the tree (*_top) are separate launches: forward forest then top,
  // backward top then forest.
  const int* fwd_items;     // 3 ints per item: supernode, kind (0 pivot, 1 update), block
  int n_fwd;
  const int* bwd_items;     // 3 ints per item: supernode, kind (0 row chunk, 1 pivot), index
                            // (row chunk: chunk * blocks + column block; pivot: block)
  int n_bwd;
  const int* fwd_items_top;
  int n_fwd_top;
  const int* bwd_items_top;
  int n_bwd_top;
  const int64_t* uvec_off;  // ns+1 offsets of the forward update vectors in ubuf
  const int* pull_ptr;      // front-row pull lists (children contributions)
  const int* pull_src;
  const int64_t* front_off; // ns+1: offset of s's front rows in the pull_ptr index
  double* ubuf;             // forward update vectors (sentinel-filled per sweep; per system)
  double* bpart;            // backward chunk partials: chunk u of s at bpart_off[s] + u*nc
  const int64_t* bpart_off;
  int* err;                 // 1: a wait hit its bound (sticky until the host clears it)
  int* ticket;              // 4 item counters (forward forest, forward top, backward top,
                            // backward forest), zeroed with the first sentinel fill of a solve
  int grid_fwd;             // workgroups of the persistent forward forest sweep (resident at once)
  int cap_top = 0;          // most workgroups of a top-of-tree sweep launch (0: one per item)
  int cap_forest = 0;       // most workgroups of a backward forest sweep launch (0: one per item)
};
// Workgroups of the forward forest sweep one compute unit holds at once.
int forest_blocks_per_cu();
// p[i] = the sentinel, i < n (per system); zero4 (when not null): four ints set to 0
void fill_sentinel(double* p, int64_t n, const Batch& b, hipStream_t st, int* zero4 = nullptr);
// The same for three buffers (p_i[0..n_i)) in one launch.
void fill_sentinel3(double* p0, int64_t n0, double* p1, int64_t n1, double* p2, int64_t n2,
                    const Batch& b, hipStream_t st, int* zero4 = nullptr);
// L y = b (b, y internal order; y sentinel-filled, ubuf sentinel-filled).
// part 0: the forest items, part 1: the top items.
void solve_forward(const CholDev& c, const SolveDev& sd, const double* rhs, double* y, int part,
                   const Batch& b, hipStream_t st);
// L^T x = y (x and bpart sentinel-filled).
void solve_backward(const CholDev& c, const SolveDev& sd, const double* y, double* x, int part,
                    const Batch& b, hipStream_t st);

}  // namespace dev
}  // namespace hykkt
