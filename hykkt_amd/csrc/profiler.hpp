// profiler.hpp -- optional per-kernel timing with HIP events.
//
// When enabled, every launch of an instrumented kernel is bracketed by two
// events on the stream it runs on; collect() (stream idle) adds each launch's
// duration, and the algorithmic flops / HBM bytes the launch was charged with,
// to its kernel's totals.  Disabled, begin()/end() cost one branch.
#pragma once

#include <hip/hip_runtime_api.h>

#include <vector>

namespace hykkt {

enum ProfKind {
  PK_DENSE_NT = 0,       // 128 x 128 tiles
  PK_DENSE_NT_SHORT,     // 32 x 128 tiles (the short launches on the critical path)
  PK_POTRF,
  PK_EXTEND_ADD,
  PK_SOLVE_FWD_FOREST,
  PK_SOLVE_FWD_TOP,
  PK_SOLVE_BWD_TOP,
  PK_SOLVE_BWD_FOREST,
  PK_SPMV,
  PK_SPGEMM_VALUES,
  PK_SPGEMM_SCATTER,
  PK_TRSM,               // blocked-substitution TRSM, 128-row tiles
  PK_TRSM_SHORT,         // the same, 32-row tiles
  PK_INV_BLOCKS,
  PK_COUNT
};

// Algorithmic HBM bytes of one y = alpha A x + beta y_in over a CSR A (rows
// x cols, nnz entries): row pointers, column indices and values read once
// (and the value map, when the values are read through one), x and y_in
// once, y written once; with a fused dot, its other operand read once.
inline double spmv_bytes(int rows, int cols, double nnz, bool map, bool y_in, bool dot)
{
  return 4.0 * (rows + 1) + (map ? 16.0 : 12.0) * nnz + 8.0 * cols +
         8.0 * rows * (1 + (y_in ? 1 : 0) + (dot ? 1 : 0));
}

class Profiler {
 public:
  struct Total {
    long long launches = 0;
    double ms = 0.0, flops = 0.0, bytes = 0.0;
  };
  Profiler() = default;
  Profiler(const Profiler&) = delete;
  Profiler& operator=(const Profiler&) = delete;
  ~Profiler();
  void enable(bool on);   // also clears the totals
  bool on() const { return on_ && !paused_; }
  // Leave the launches between pause(true) and pause(false) out.
  void pause(bool p) { paused_ = p; }
  void begin(hipStream_t st);
  void end(int kind, hipStream_t st, double flops, double bytes);
  void collect();
  const Total& total(int k) const { return tot_[k]; }
  static const char* name(int k);

 private:
  struct Pend {
    int kind;
    hipEvent_t a, b;
    double flops, bytes;
  };
  hipEvent_t get();
  bool on_ = false;
  bool paused_ = false;
  hipEvent_t open_ = nullptr;
  std::vector<Pend> pend_;
  std::vector<hipEvent_t> free_, all_;
  Total tot_[PK_COUNT];
};

}  // namespace hykkt
