// HykktSolver.hpp -- header-only C++ HykktSolver with the reference class's
// constructor, setters and execute() (src/HykktSolver.hpp:17-137), forwarding
// to the C ABI of libhykkt.so (hykkt.h).  A reference caller recompiles
// against this header and links -lhykkt; no other change.
//
// Ownership and call semantics are the reference's: setters keep borrowed
// host pointers that are read again at every execute() (HykktSolver.cpp:
// 635-666), so a later system is passed by overwriting the same buffers;
// read_matrix_files makes the solver own what it reads; execute() returns 0
// if the reference's residual test passes (HykktSolver.cpp:914-929), else 1.
//
// Errors: the reference ends the process on a CUDA or library error
// (cuda_check_errors.hpp:3-14).  This shim does the same (message on stderr,
// exit(EXIT_FAILURE)) unless HYKKT_SHIM_THROWS is defined before inclusion,
// in which case it throws std::runtime_error with the library's message.
//
// Additions (no reference counterpart): options(), stats(), get_solution(),
// and a move-only handle.
#ifndef HYKKT_SOLVER_HPP
#define HYKKT_SOLVER_HPP

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "hykkt.h"

class HykktSolver {
 public:
  explicit HykktSolver(double gamma) { check(hykkt_create(gamma, &h_), "HykktSolver"); }
  ~HykktSolver()
  {
    if (h_) hykkt_destroy(h_);
  }
  HykktSolver(const HykktSolver&) = delete;
  HykktSolver& operator=(const HykktSolver&) = delete;
  HykktSolver(HykktSolver&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  HykktSolver& operator=(HykktSolver&& o) noexcept
  {
    if (this != &o) {
      if (h_) hykkt_destroy(h_);
      h_ = o.h_;
      o.h_ = nullptr;
    }
    return *this;
  }

  // HykktSolver.hpp:25-33
  void read_matrix_files(char const* const h_file, char const* const ds_file,
                         char const* const jc_file, char const* const jd_file,
                         char const* const rx_file, char const* const rs_file,
                         char const* const ry_file, char const* const ryd_file, int skip_lines)
  {
    check(hykkt_read_matrix_files(h_, h_file, ds_file, jc_file, jd_file, rx_file, rs_file,
                                  ry_file, ryd_file, skip_lines),
          "read_matrix_files");
  }

  // HykktSolver.hpp:35-41 (CSR blocks: row pointers, column indices, values)
  void set_H(int* rowptrs, int* colidx, double* vals, int n, int m, int nnz)
  {
    check(hykkt_set_H(h_, rowptrs, colidx, vals, n, m, nnz), "set_H");
  }
  void set_Ds(int* rowptrs, int* colidx, double* vals, int n, int m, int nnz)
  {
    check(hykkt_set_Ds(h_, rowptrs, colidx, vals, n, m, nnz), "set_Ds");
  }
  void set_Jc(int* rowptrs, int* colidx, double* vals, int n, int m, int nnz)
  {
    check(hykkt_set_Jc(h_, rowptrs, colidx, vals, n, m, nnz), "set_Jc");
  }
  void set_Jd(int* rowptrs, int* colidx, double* vals, int n, int m, int nnz)
  {
    check(hykkt_set_Jd(h_, rowptrs, colidx, vals, n, m, nnz), "set_Jd");
  }

  // HykktSolver.hpp:43-54
  void set_rx(double* vals, int n) { check(hykkt_set_rx(h_, vals, n), "set_rx"); }
  void set_rs(double* vals, int n) { check(hykkt_set_rs(h_, vals, n), "set_rs"); }
  void set_ry(double* vals, int n) { check(hykkt_set_ry(h_, vals, n), "set_ry"); }
  void set_ryd(double* vals, int n) { check(hykkt_set_ryd(h_, vals, n), "set_ryd"); }
  void set_x_host(double* vals, int n) { check(hykkt_set_x_host(h_, vals, n), "set_x_host"); }
  void set_s_host(double* vals, int n) { check(hykkt_set_s_host(h_, vals, n), "set_s_host"); }
  void set_y_host(double* vals, int n) { check(hykkt_set_y_host(h_, vals, n), "set_y_host"); }
  void set_yd_host(double* vals, int n) { check(hykkt_set_yd_host(h_, vals, n), "set_yd_host"); }

  // HykktSolver.hpp:56
  void set_gamma(double gamma) { check(hykkt_set_gamma(h_, gamma), "set_gamma"); }

  // HykktSolver.hpp:58: 0 if the residual test passed, 1 otherwise
  int execute()
  {
    int r = hykkt_execute(h_);
    check(r, "execute");
    return r;
  }

  // ---- additions
  hykkt_options options() const
  {
    hykkt_options o;
    check(hykkt_get_options(h_, &o), "options");
    return o;
  }
  void set_options(const hykkt_options& o) { check(hykkt_set_options(h_, &o), "set_options"); }
  hykkt_stats stats() const
  {
    hykkt_stats s;
    check(hykkt_get_stats(h_, &s), "stats");
    return s;
  }
  // Copy the last solution to host arrays (any may be null).
  void get_solution(double* dx, double* ds, double* dyc, double* dyd) const
  {
    check(hykkt_get_solution(h_, dx, ds, dyc, dyd), "get_solution");
  }
  hykkt_solver handle() const { return h_; }

 private:
  static void check(int rc, const char* where)
  {
    if (rc >= 0) return;   // HYKKT_OK or HYKKT_RESIDUAL_FAIL
    std::string msg = std::string("hykkt: ") + where + ": " + hykkt_last_error();
#ifdef HYKKT_SHIM_THROWS
    throw std::runtime_error(msg);
#else
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
#endif
  }
  hykkt_solver h_ = nullptr;
};

#endif  // HYKKT_SOLVER_HPP
