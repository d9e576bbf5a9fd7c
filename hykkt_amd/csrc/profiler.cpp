// profiler.cpp -- per-kernel timing with HIP events (see profiler.hpp).
#include "profiler.hpp"

#include "common.hpp"
#include "kernels.hpp"

namespace hykkt {

Profiler::~Profiler()
{
  for (hipEvent_t e : all_) (void)hipEventDestroy(e);
}

const char* Profiler::name(int k)
{
  // kernel names as rocprofv3 prints them (template instances of the sweeps)
  static const char* const names[PK_COUNT] = {
      "k_dense_nt<128>",      "k_dense_nt<32>",      "k_potrf_inv",         "k_extend_add",
      "k_fwd_forest",         "k_solve_fwd<1024, true>", "k_solve_bwd<1024, true>",
      "k_solve_bwd<256, false>", "k_spmv",             "k_spgemm_values",
      "k_spgemm_scatter",     "k_trsm<128>",         "k_trsm<32>",
      "k_inv_blocks"};
  return (k >= 0 && k < PK_COUNT) ? names[k] : "";
}

void Profiler::enable(bool on)
{
  on_ = on;
  paused_ = false;
  for (auto& t : tot_) t = Total{};
  for (const Pend& p : pend_) {
    free_.push_back(p.a);
    free_.push_back(p.b);
  }
  pend_.clear();
  if (open_) free_.push_back(open_);
  open_ = nullptr;
}

hipEvent_t Profiler::get()
{
  if (free_.empty()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    all_.push_back(e);
    return e;
  }
  hipEvent_t e = free_.back();
  free_.pop_back();
  return e;
}

void Profiler::begin(hipStream_t st)
{
  if (!on()) return;
  HYKKT_REQUIRE(open_ == nullptr, HYKKT_E_STATE, "profiler: nested begin");
  open_ = get();
  HIPCHK(hipEventRecord(open_, st));
}

void Profiler::end(int kind, hipStream_t st, double flops, double bytes)
{
  if (!on()) return;
  HYKKT_REQUIRE(open_ != nullptr, HYKKT_E_STATE, "profiler: end without begin");
  hipEvent_t b = get();
  HIPCHK(hipEventRecord(b, st));
  pend_.push_back({kind, open_, b, flops, bytes});
  open_ = nullptr;
}

void Profiler::collect()
{
  for (const Pend& p : pend_) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
    Total& t = tot_[p.kind];
    t.launches++;
    t.ms += ms;
    t.flops += p.flops;
    t.bytes += p.bytes;
    free_.push_back(p.a);
    free_.push_back(p.b);
  }
  pend_.clear();
}

}  // namespace hykkt
