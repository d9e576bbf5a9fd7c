// solver.hpp -- the HykktSolver pipeline on one MI355X (host orchestration),
// for one system at a time (the drop-in) or a batch of systems that share
// one sparsity pattern.
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/hykkt.h"
#include "device.hpp"
#include "mmio.hpp"

namespace hykkt {

struct BlockIn {
  const int* rp = nullptr;
  const int* ci = nullptr;
  const double* v = nullptr;
  int n = 0, m = 0, nnz = 0;
  bool set = false;
};

enum { BLK_H = 0, BLK_DS, BLK_JC, BLK_JD };
enum { VEC_RX = 0, VEC_RS, VEC_RY, VEC_RYD, VEC_X, VEC_S, VEC_Y, VEC_YD };

// The value arrays of one system of a batch (device or host pointers, as
// the option inputs_on_device says): H, Ds, Jc, Jd values, rx, rs, ry, ryd,
// and the Schur CG start y (may be null: zero).
enum { IN_H = 0, IN_DS, IN_JC, IN_JD, IN_RX, IN_RS, IN_RY, IN_RYD, IN_Y0, IN_COUNT };
struct SystemIn {
  const double* v[IN_COUNT] = {};
  double gamma = 0.0;
  bool set = false;
};

class Solver {
 public:
  // capacity: systems of one pattern solved together by execute_batch
  explicit Solver(double gamma, int capacity = 1);
  ~Solver();
  void set_block(int which, const int* rp, const int* ci, const double* v, int n, int m, int nnz);
  void set_vec(int which, const double* v, int n);
  void set_gamma(double g) { gamma_ = g; }
  void set_perm(const int* perm, int n);
  void read_files(const char* const files[8], int skip);
  // The reference's execute() on the system bound by the setters.
  int execute();
  // Bind the values of system k of a batch (the pattern is the setters').
  void bind_system(int k, double gamma, const double* const vals[IN_COUNT]);
  // Solve systems 0 .. count-1 together; per-system status as execute()
  // returns it (0, 1, or a negative error code for that system).
  std::vector<int> execute_batch(int count);
  int capacity() const { return cap_; }
  const hykkt_stats& stats_of(int k) const;
  const hykkt_tuning& tuning() const { return tuning_; }
  void get_solution(int k, double* dx, double* ds, double* dyc, double* dyd);
  void get_solution(double* dx, double* ds, double* dyc, double* dyd)
  {
    get_solution(0, dx, ds, dyc, dyd);
  }
  void get_solution_device(int k, const double** dx, const double** ds, const double** dyc,
                           const double** dyd) const;
  void get_perm(int* perm, int n) const;
  long long debug_copy(int what, void* out, long long cap);

  hykkt_options opts;
  hykkt_stats stats{};   // system 0's (the drop-in's) statistics
  Profiler prof;   // per-kernel timing of the factor and solve launches

 private:
  struct Sym;    // once-per-pattern state
  struct Slots;  // per-system arena layout
  void analyze();
  bool pattern_matches() const;
  void upload_values(int count);
  // One pass over `count` systems: the setters' system alone (count 1,
  // from_setters) or the bound batch.
  std::vector<int> run(int count, bool from_setters);
  // H_gamma from the Ruiz-scaled blocks into the factor storage (+ delta).
  void assemble_hgamma(const dev::Batch& b);
  // Time the numeric factorization of the batch under candidate launch
  // shapes and keep the fastest (after formation, before the factorization).
  void autotune(const dev::Batch& b);
  // Work buffers of one solve pass (r^x permuted, Schur rhs, CG iterate):
  // arena offsets.
  struct Pass {
    int64_t rxp, schur, cx;
  };
  std::vector<CgResult> solve_phase(const Pass& w, int64_t rxt, int64_t ysc, int64_t rs,
                                    int64_t ryd, bool x0, int64_t dx, int64_t dsv, int64_t dyc,
                                    int64_t dyd, const dev::Batch& b, hipEvent_t after_cg,
                                    int iters_hint);
  // Residual rows of (dx, dsv, dyc, dyd) (arena offsets) for the systems of
  // b; rows: keep the signed rows for refinement.  Enqueues only; the eight
  // squared norms of each system are at its `res` slot.
  void residual(int64_t dx, int64_t dsv, int64_t dyc, int64_t dyd, bool rows, const dev::Batch& b);
  // The res slots of `count` systems to the host in one round trip
  // (synchronizes; throws if a sweep exceeded its wait bound); with zrow /
  // piv, also each system's zero-row flag (INT_MAX: none) and first failing
  // pivot (-1: none).
  std::vector<double> read_res(int count, std::vector<int>* zrow = nullptr,
                               std::vector<int>* piv = nullptr);
  double* P(int64_t off) const;   // system 0's address of an arena slot
  // Run the launch in fn bracketed for the profiler, charged with the work
  // of the b.count systems it covers.
  template <class F>
  void timed(int kind, double flops, double bytes, const dev::Batch& b, F&& fn)
  {
    prof.begin(st_);
    fn();
    prof.end(kind, st_, flops * b.count, bytes * b.count);
  }

  double gamma_;
  int cap_ = 1;
  hipStream_t st_ = nullptr;
  hipEvent_t ev_[7] = {};   // 5, 6: around the autotuning runs
  float tune_ms_ = 0.f;      // device time of the autotuning runs of this pass
  BlockIn blk_[4];
  const double* vecs_[8] = {};
  int vec_n_[8] = {};
  std::vector<SystemIn> sys_in_;
  std::vector<hykkt_stats> st_k_;
  // CG iterations of the last first pass and of the last refinement pass:
  // where the host first looks at the live flags in the next such pass
  int last_iters_ = 0;
  int last_ref_iters_ = 0;
  hykkt_tuning tuning_{};
  // file-path storage (read_matrix_files)
  CsrHost fblk_[4];
  std::vector<double> fvec_[8];
  bool file_path_ = false;
  bool status_ = true;     // HykktSolver::status_ (HykktSolver.cpp:142)
  bool jd_flag_ = false;
  std::vector<int> user_perm_;
  std::unique_ptr<Sym> sym_;
};

}  // namespace hykkt
