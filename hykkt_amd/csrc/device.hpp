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

// Numeric supernodal multifrontal Cholesky on the device for one sparsity
// pattern (the analysis is done once; factor() and solve() run per solve).
class CholEngine {
 public:
  // `a`: full n x n CSR whose lower triangle is factored.
  CholEngine(const Pattern& a, const CholOptions& opt, hipStream_t st);
  ~CholEngine();
  CholEngine(const CholEngine&) = delete;
  CholEngine& operator=(const CholEngine&) = delete;
  const CholPlan& plan() const { return pl_; }
  double* L() { return L_.p; }
  // Zero the factor storage (before assembly of the matrix entries).
  void zero_L();
  // Scatter values (in a's CSR order) into the factor storage (zeroes first).
  void assemble(const double* vals_dev);
  // Add `delta` to every diagonal entry of the assembled matrix.
  void shift_diag(double delta);
  // Factor the assembled matrix; pivot failures are recorded on the device.
  void factor(double pivot_tol);
  // First failing pivot (internal index) or -1; synchronizes the stream.
  int pivot();
  // Nonzero if a triangular-solve sweep hit its wait bound since the last
  // call (synchronizes the stream, clears the error word; solves in between
  // leave it set).
  int solve_error();
  // Solve L L^T x = b in place, x in internal (postordered) order.
  void solve(double* x);
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
    Launch ea_panel, ea_u, syrk;
    std::vector<Launch> potrf, trsm, upd_next, upd_rest;
    // per step: its position in its group of consecutive steps and the
    // group's size (1: the step updates after its own block; see build_tasks)
    std::vector<int> gpos, gsize;
  };
  void build_tasks();
  void build_solve();

  CholPlan pl_;
  hipStream_t st_;
  hipStream_t st2_ = nullptr;   // look-ahead: the rest of each trailing update
  hipEvent_t ev_main_ = nullptr, ev_rest_ = nullptr;
  Profiler* prof_ = nullptr;
  dev::CholDev cd_{};
  dev::SolveDev sd_{};
  DevBuf<int> sn_first_, rows_ptr_, rows_, rel_, child_ptr_, child_list_, post_, sn_parent_;
  DevBuf<int64_t> l_off_, u_off_, a_dst_, diag_pos_, uvec_off_, dinv_off_, front_off_;
  DevBuf<int> a_src_;
  DevBuf<double> L_, U_, ubuf_, dinv_;
  DevBuf<int> pivot_;
  // task lists
  std::vector<Level> lv_;
  DevBuf<int> t_ea_p_, t_ea_col_, t_ea_cptr_, t_ea_cc_, t_ea_ccol_;
  DevBuf<int> t_potrf_, t_trsm_sn_, t_trsm_tr_, t_trsm_tc_, t_upd_sn_, t_upd_tr_, t_upd_tc_;
  DevBuf<int> t_syrk_sn_, t_syrk_tr_, t_syrk_tc_;
  int n_cu_ = 1;   // compute units: persistent-launch width
  // solve items
  DevBuf<int> fwd_items_, bwd_items_, pull_ptr_, pull_src_, flags_;
  DevBuf<int64_t> bpart_off_;
  DevBuf<double> bpart_;
  DevBuf<int> fwd_top_items_, bwd_top_items_;
  DevBuf<double> ybuf_;                   // forward-sweep output
  // development timeline of one solve (HYKKT_SOLVE_TRACE=file, taken at the
  // HYKKT_SOLVE_TRACE_AT-th solve, default 20)
  DevBuf<long long> trace_;
  std::vector<int> trace_items_;
  std::string trace_path_;
  long trace_at_ = -1, n_solve_calls_ = 0;
  int64_t n_bpart_ = 0;
  double solve_lower_[2] = {0.0, 0.0};   // stored lower-trapezoid entries: forest, top
  int64_t solve_n_[2] = {0, 0};          // columns: forest, top
};

// The Schur-complement operator S = Jh A^{-1} Jh^T with A factored by `chol`
// (Jh: m x n, columns in the factor's internal order) and the CG workspace.
struct SchurOps {
  int m = 0, n = 0;
  const int *jh_rp = nullptr, *jh_ci = nullptr;
  const double* jh_v = nullptr;
  const int *jth_rp = nullptr, *jth_ci = nullptr;
  const double* jth_v = nullptr;
  CholEngine* chol = nullptr;
  hipStream_t st = nullptr;
  double *t1 = nullptr;                                            // n
  double *x = nullptr, *r = nullptr, *w = nullptr, *p = nullptr, *s = nullptr;   // m
  double *sc = nullptr, *part = nullptr;
};

// SchurComplementConjugateGradient::solve (…ConjugateGradient.cpp:89-194) on
// the device: ops.x holds x0 on entry and the solution on exit.  The host
// reads ||r||^2 back every `check_every` iterations for the stopping test.
// Returns 0 if converged (sqrt(r.r) < tol), 1 after itmax iterations.
// x0_zero: ops.x is known to be all zeros (r = b without a solve).
int schur_cg_run(SchurOps& ops, const double* b, double tol, int itmax, int check_every,
                 int* iters, double* final_norm, bool x0_zero = false);

}  // namespace hykkt
