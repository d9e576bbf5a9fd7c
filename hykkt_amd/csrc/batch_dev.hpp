// batch_dev.hpp -- device helpers for launches over a batch of same-pattern
// systems (see Batch in kernels.hpp): a workgroup's system and its copy of a
// per-system buffer.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace hykkt {
namespace dev {

// The system of this workgroup (blockIdx.y), or -1 when the mask skips it.
__device__ __forceinline__ int batch_sys(const Batch& b)
{
  const int s = blockIdx.y;
  return (b.mask && !b.mask[s]) ? -1 : s;
}

// System s's copy of the per-system buffer p (null stays null).
template <class T>
__device__ __forceinline__ T* batch_at(T* p, int s, const Batch& b)
{
  return p ? p + static_cast<int64_t>(s) * b.vs : p;
}

}  // namespace dev
}  // namespace hykkt
