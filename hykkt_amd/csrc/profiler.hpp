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
  PK_DENSE_NT = 0,
  PK_POTRF,
  PK_EXTEND_ADD,
  PK_SOLVE_FWD_FOREST,
  PK_SOLVE_FWD_TOP,
  PK_SOLVE_BWD_TOP,
  PK_SOLVE_BWD_FOREST,
  PK_COUNT
};

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
  bool on() const { return on_; }
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
  hipEvent_t open_ = nullptr;
  std::vector<Pend> pend_;
  std::vector<hipEvent_t> free_, all_;
  Total tot_[PK_COUNT];
};

}  // namespace hykkt
