// solver.hpp -- the HykktSolver pipeline on one MI355X (host orchestration).
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

class Solver {
 public:
  explicit Solver(double gamma);
  ~Solver();
  void set_block(int which, const int* rp, const int* ci, const double* v, int n, int m, int nnz);
  void set_vec(int which, const double* v, int n);
  void set_gamma(double g) { gamma_ = g; }
  void set_perm(const int* perm, int n);
  void read_files(const char* const files[8], int skip);
  int execute();
  void get_solution(double* dx, double* ds, double* dyc, double* dyd);
  void get_solution_device(const double** dx, const double** ds, const double** dyc,
                           const double** dyd) const;
  void get_perm(int* perm, int n) const;
  long long debug_copy(int what, void* out, long long cap);

  hykkt_options opts;
  hykkt_stats stats{};
  Profiler prof;   // per-kernel timing of the factor and solve launches

 private:
  struct Sym;   // once-per-pattern state
  void analyze();
  bool pattern_matches() const;
  void upload_values();
  void chol_solve(double* x);
  // H_gamma from the Ruiz-scaled blocks into the factor storage (+ delta).
  void assemble_hgamma();
  // Work buffers of one solve pass (r^x permuted, Schur rhs, CG iterate).
  struct Pass {
    double *rxp, *schur, *cx;
  };
  void solve_phase(const Pass& w, const double* rxt, const double* ysc, const double* rs,
                   const double* ryd, const double* x0, double* dx, double* dsv, double* dyc,
                   double* dyd, int* iters, double* fnorm, int* cg_status, hipEvent_t after_cg);
  void residual(const double* dx, const double* dsv, const double* dyc, const double* dyd,
                bool rows, hipEvent_t done, double out[8]);
  const double* val(int which) const;
  const double* vec(int which) const;

  double gamma_;
  hipStream_t st_ = nullptr;
  hipEvent_t ev_[5] = {};
  BlockIn blk_[4];
  const double* vecs_[8] = {};
  int vec_n_[8] = {};
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
