// device.hpp -- RAII device buffers and the supernodal Cholesky engine.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <utility>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "profiler.hpp"
#include "symbolic.hpp"

namespace hykkt {

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  // Exchange the buffers (no copies, no allocation).
  void swap(DevBuf& o) noexcept
  {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
  void alloc(size_t cnt)
  {
    release();
    if (cnt) HIPCHK(hipMalloc(reinterpret_cast<void**>(&p), cnt * sizeof(T)));
    n = cnt;
  }
  void ensure(size_t cnt)
  {
    if (cnt > n) alloc(cnt);
  }
  void upload(const std::vector<T>& v)
  {
    alloc(v.size());
    if (n) HIPCHK(hipMemcpy(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const T* v, size_t cnt)
  {
    alloc(cnt);
    if (n) HIPCHK(hipMemcpy(p, v, n * sizeof(T), hipMemcpyHostToDevice));
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Pinned host memory (copies into it are asynchronous).
template <class T>
struct HostBuf {
  T* p = nullptr;
  size_t n = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { release(); }
  void ensure(size_t cnt)
  {
    if (cnt <= n) return;
    release();
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p), cnt * sizeof(T), hipHostMallocDefault));
    n = cnt;
  }
  void release()
  {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

// Numeric supernodal multifrontal Cholesky on the device for one sparsity
// pattern (the analysis is done once; factor() and solve() run per solve),
// for up to `capacity` systems of that pattern at once.  The per-system
// storage (factor panels, update matrices, inverted diagonal blocks, sweep
// buffers) is either bound into a caller's arena (bind: system k's copy
// k*vs doubles from system 0's) or owned (own).
class CholEngine {
 public:
  // Doubles of per-system storage, by kind.
  struct Sizes {
    int64_t L, U, dinv, y, ubuf, bpart;
  };
  // Per-system storage of system 0 (pivot: one int per system).
  struct Storage {
    double *L, *U, *dinv, *y, *ubuf, *bpart;
    int* pivot;
  };
  // Launch shapes of the numeric factorization (tile counts are over all the
  // systems of a launch).  Every choice gives the same bits: a C entry takes
  // the same MFMA chain in the same K order whatever the tile or group.
  struct Tuning {
    int short_max = 128;   // TRSM tiles of 128 rows up to which a step uses 32-row tiles
    int group_max = 8;     // most steps sharing one trailing update
    int pair_min = 512;    // tiles a group of G steps needs: pair_min * G^2 / 4
  };
  // `a`: full n x n CSR whose lower triangle is factored.
  CholEngine(const Pattern& a, const CholOptions& opt, hipStream_t st);
  ~CholEngine();
  CholEngine(const CholEngine&) = delete;
  CholEngine& operator=(const CholEngine&) = delete;
  const CholPlan& plan() const { return pl_; }
  Sizes sizes() const;
  // Use storage for `cap` systems, vs doubles apart (the inverted diagonal
  // blocks are zeroed here: only their lower parts are ever written).
  void bind(const Storage& s, int64_t vs, int cap);
  // Allocate and bind storage for `cap` systems.
  void own(int cap);
  int64_t stride() const { return vs_; }
  // Rebuild the task lists with launch shapes t for launches of `count` systems.
  void set_tuning(const Tuning& t, int count);
  const Tuning& tuning() const { return tun_; }
  // Let each sweep launch hold at most 1/share of the device's workgroup
  // slots (share <= 1: no bound), leaving compute units to the kernels of
  // other solvers' streams; the sweeps' results do not depend on it.
  void set_sweep_share(int share);
  // Device words the solver reads back with its other per-pass results:
  // each system's first failing pivot (INT_MAX: none) and the sweeps' wait-
  // bound error word (0: none; clear_solve_error resets it).
  const int* pivot_dev() const { return cd_.pivot; }
  const int* solve_error_dev() const { return sd_.err; }
  void clear_solve_error();
  // Launch shapes fixed by HYKKT_PAIR_MIN / HYKKT_GROUP_MAX (not autotuned).
  bool tuning_pinned() const { return pinned_; }
  double* L() { return cd_.L; }
  // Zero the factor storage (before assembly of the matrix entries); where
  // prezero_L already zeroes it, only wait for that.
  void zero_L(const dev::Batch& b);
  // Zero the factor storage of `count` systems on the second stream, for the
  // next assembly, once the factor is no longer needed (all work on the
  // solver's stream that reads it has finished: the caller synchronized).
  void prezero_L(int count);
  // Wait for the second stream's work (before the storage is freed or rebound).
  void quiesce() { (void)hipStreamSynchronize(st2_); }
  // Scatter values (in a's CSR order) into the factor storage (zeroes first).
  void assemble(const double* vals_dev, const dev::Batch& b);
  // Add `delta` to every diagonal entry of the assembled matrix.
  void shift_diag(double delta, const dev::Batch& b);
  // Factor the assembled matrices of b.count systems; pivot failures are
  // recorded on the device.
  void factor(double pivot_tol, const dev::Batch& b);
  // First failing pivot (internal index) or -1 of each of `count` systems;
  // synchronizes the stream.
  std::vector<int> pivots(int count);
  int pivot() { return pivots(1)[0]; }
  // Nonzero if a triangular-solve sweep hit its wait bound since the last
  // call (synchronizes the stream, clears the error word; solves in between
  // leave it set).
  int solve_error();
  // Solve L L^T x = b in place for each system of b (its mask skips
  // systems), x in internal (postordered) order, per system at x + k*b.vs.
  void solve(double* x, const dev::Batch& b);
  void solve(double* x) { solve(x, dev::Batch{1, vs_, nullptr}); }
  // Device arrays of the assembly map (a's lower-triangle entries).
  const int* a_src() const { return a_src_.p; }
  const int64_t* a_dst() const { return a_dst_.p; }
  int n_assembled() const { return static_cast<int>(pl_.a_src.size()); }
  const int* post_dev() const { return post_.p; }
  // Time every factor/solve launch with `p` (nullptr: no timing).
  void set_profiler(Profiler* p) { prof_ = p; }

 private:
  struct Launch {
    int off = 0, cnt = 0, maxb = 0;
    bool short_rows = false;   // dense_nt tiles of 32 rows (short critical-path launches)
    double flops = 0.0, bytes = 0.0;   // algorithmic, for the profiler
  };
  struct Level {
    Launch ea_panel, ea_u, syrk, inv;
    std::vector<Launch> potrf, trsm, upd_next, upd_rest;
    // per step: its position in its group of consecutive steps and the
    // group's size (1: the step updates after its own block; see build_tasks)
    std::vector<int> gpos, gsize;
  };
  void build_tasks();
  void build_solve();

  CholPlan pl_;
  hipStream_t st_;
  int64_t vs_ = 0;   // per-system stride of the bound storage
  int cap_ = 0;
  Tuning tun_;
  bool pinned_ = false;
  int tcount_ = 1;   // systems per launch the task lists are shaped for
  int grid_fwd_full_ = 1;   // the forward forest sweep's workgroups at full share
  int share_ = 1;
  DevBuf<double> own_;   // owned storage (own)
  DevBuf<int> own_piv_;
  hipStream_t st2_ = nullptr;   // look-ahead: the rest of each trailing update
  hipEvent_t ev_main_ = nullptr, ev_rest_ = nullptr;
  hipEvent_t ev_zero_ = nullptr;   // prezero_L's fill done
  int clean_ = 0;                  // systems whose storage prezero_L zeroes, untouched since
  Profiler* prof_ = nullptr;
  dev::CholDev cd_{};
  dev::SolveDev sd_{};
  DevBuf<int> sn_first_, rows_ptr_, rows_, rel_, child_ptr_, child_list_, post_, sn_parent_;
  DevBuf<int64_t> l_off_, u_off_, a_dst_, diag_pos_, uvec_off_, dinv_off_, front_off_;
  DevBuf<int> a_src_;
  int64_t n_ubuf_ = 0, n_dinv_ = 0;
  // task lists
  std::vector<Level> lv_;
  DevBuf<int> t_ea_p_, t_ea_col_, t_ea_cptr_, t_ea_cc_, t_ea_ccol_;
  DevBuf<int> t_potrf_, t_trsm_sn_, t_trsm_tr_, t_trsm_tc_, t_upd_sn_, t_upd_tr_, t_upd_tc_;
  DevBuf<int> t_syrk_sn_, t_syrk_tr_, t_syrk_tc_;
  DevBuf<int> t_inv_sn_, t_inv_step_;
  int n_cu_ = 1;   // compute units: persistent-launch width
  // solve items
  DevBuf<int> fwd_items_, bwd_items_, pull_ptr_, pull_src_, flags_;
  DevBuf<int64_t> bpart_off_;
  DevBuf<int> fwd_top_items_, bwd_top_items_;
  double* ybuf_ = nullptr;   // forward-sweep output (bound storage)
  int64_t n_bpart_ = 0;
  double solve_lower_[2] = {0.0, 0.0};   // stored lower-trapezoid entries: forest, top
  int64_t solve_n_[2] = {0, 0};          // columns: forest, top
};

// The Schur-complement operator S = Jh A^{-1} Jh^T with A factored by `chol`
// (Jh: m x n, columns in the factor's internal order) and the CG workspace.
// Per-system vectors are at their system-0 address + k * vs (the Batch);
// the flag arrays hold one int per system.
struct SchurOps {
  int m = 0, n = 0;
  const int *jh_rp = nullptr, *jh_ci = nullptr, *jh_blk = nullptr;
  int jh_nblk = 0;
  const double* jh_v = nullptr;
  const int *jth_rp = nullptr, *jth_ci = nullptr, *jth_blk = nullptr;
  int jth_nblk = 0;
  const double* jth_v = nullptr;
  CholEngine* chol = nullptr;
  hipStream_t st = nullptr;
  double *t1 = nullptr;                                            // n
  double *x = nullptr, *r = nullptr, *w = nullptr, *p = nullptr, *s = nullptr;   // m
  double *sc = nullptr, *part = nullptr;   // 8 scalars; partials (max of both reductions)
  // optional (n each): z = A^{-1} Jh^T x carried along with the iterate x
  // (q the same image of the direction p), so the recovery needs no solve
  double *tq = nullptr, *tz = nullptr;
  int *live = nullptr, *iters = nullptr, *status = nullptr;   // per system
  double* gath = nullptr;   // 4 doubles per system: the results packed for the readback
  Profiler* prof = nullptr;  // optional launch timing of the SpMVs
  int nnz = 0;               // entries of Jc
  // optional pinned host memory of at least 4 doubles per system, for the
  // live-flag checks and the results (pageable vectors otherwise)
  double* pinned = nullptr;
  // with pinned memory: leave the results' copy in flight and return them
  // unfilled; cg_unpack(pinned, ...) fills them after the caller's next
  // synchronization of the stream (one host round trip less per pass)
  bool defer = false;
};
// Partials a SchurOps needs per system.
int schur_partials(int m, int jh_nblk);

struct CgResult {
  int iters = 0, status = 0;
  double final_norm = 0.0;
};

// SchurComplementConjugateGradient::solve (…ConjugateGradient.cpp:89-194)
// on the device for the systems of `b` (its mask: the systems taking part):
// ops.x holds x0 on entry and the solution on exit.  Every system's stopping
// test (sqrt(r.r) < tol, or itmax iterations) is taken on the device, which
// then leaves that system alone; the host reads the systems' live flags
// only after `chunk` iterations and then after every `chunk_next`, so the
// iterations in between run without a host round trip.  x0_zero: every x is
// zero (r = b without a solve).  Returns per system (count entries).
std::vector<CgResult> schur_cg_run(SchurOps& ops, const double* b, double tol, int itmax,
                                   const dev::Batch& bt, bool x0_zero, int chunk,
                                   int chunk_next = 1);
// The results of a deferred schur_cg_run from its pinned memory.
void cg_unpack(const double* pinned, std::vector<CgResult>& res);

}  // namespace hykkt
