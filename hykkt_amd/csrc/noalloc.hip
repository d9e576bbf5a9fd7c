// noalloc.hip -- the device scratch of the allocation-free entry point
// (hykkt_ruiz_scale_noalloc, capi.cpp): a file-scope device array that every
// GPU the library runs on holds until the process ends, and its address on
// the current device.
//
// The array's name is a host-side stand-in in host code, not a device
// address: the address comes from hipGetSymbolAddress, taken once per device
// (with that device current) and kept.  Every call on a device shares that
// device's array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace hykkt {
namespace dev {

namespace {

constexpr size_t SCRATCH_BYTES = size_t{32} << 20;   // 32 MiB per GPU

static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return SCRATCH_BYTES; }

void* noalloc_scratch()
{
  int d = 0;
  HIPCHK(hipGetDevice(&d));
  HYKKT_REQUIRE(d >= 0 && d < MAX_DEVICES, HYKKT_E_STATE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[d]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HYKKT_REQUIRE(p != nullptr, HYKKT_E_HIP, "no device address for the scratch array");
    g_ptr[d] = p;
  }
  return g_ptr[d];
}

}  // namespace dev
}  // namespace hykkt
