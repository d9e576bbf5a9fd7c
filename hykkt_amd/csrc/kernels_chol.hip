// kernels_chol.hip -- extend-add of the supernodal multifrontal Cholesky (gfx950).
//
// Storage (see CholPlan): supernode s owns columns [a, a+nc) and the sorted
// off-diagonal-block rows `rows(s)` (nr of them); its front index list is
// [cols; rows] (f = nc + nr).  The factor panel is f x nc column-major at
// L + l_off[s]; the update matrix is the nr x nr lower triangle packed by
// columns at U + u_off[s].
#include <hip/hip_runtime.h>

#include <climits>

#include "batch_dev.hpp"
#include "kernels.hpp"

namespace hykkt {
namespace dev {

namespace {

struct SnInfo {
  int a, nc, nr, f;
  double* Ls;
};

__device__ __forceinline__ SnInfo sninfo(const CholDev& c, double* L, int s)
{
  SnInfo i;
  i.a = c.sn_first[s];
  i.nc = c.sn_first[s + 1] - i.a;
  i.nr = c.rows_ptr[s + 1] - c.rows_ptr[s];
  i.f = i.nc + i.nr;
  i.Ls = L + c.l_off[s];
  return i;
}

// ----------------------------------------------------------------- extend-add
// One workgroup per parent front column and system.  Children are added one
// after the other (fixed order => deterministic sums); within one child the
// relative index map is injective.
__global__ void __launch_bounds__(256) k_extend_add(CholDev c, const int* __restrict__ task_p,
                                                    const int* __restrict__ task_col,
                                                    const int* __restrict__ cptr,
                                                    const int* __restrict__ cc,
                                                    const int* __restrict__ ccol, Batch b)
{
  const int sy = batch_sys(b);
  if (sy < 0) return;
  double* const L = batch_at(c.L, sy, b);
  double* const U = batch_at(c.U, sy, b);
  const int t = blockIdx.x;
  const int p = task_p[t], col = task_col[t];
  const SnInfo ip = sninfo(c, L, p);
  double* base;
  int shift;
  if (col < ip.nc) {
    base = ip.Ls + static_cast<int64_t>(col) * ip.f;
    shift = 0;
  } else {
    base = U + c.u_off[p] + packed_lower(0, col - ip.nc, ip.nr);   // rows >= col only
    shift = -ip.nc;
  }
  for (int k = cptr[t]; k < cptr[t + 1]; k++) {
    const int ch = cc[k], r2 = ccol[k];
    const int nrc = c.rows_ptr[ch + 1] - c.rows_ptr[ch];
    const double* src = U + c.u_off[ch] + packed_lower(0, r2, nrc);   // rows r1 >= r2
    const int* relc = c.rel + c.rows_ptr[ch];
    // four entries per thread at a time, every load issued before the
    // read-modify-writes (one child's rows map to distinct parent rows, so
    // the order among them is free); the sums are the same bits
    constexpr int UNR = 4, BS = 256;
    int r1 = r2 + threadIdx.x;
    for (; r1 + (UNR - 1) * BS < nrc; r1 += UNR * BS) {
      int ri[UNR];
      double sv[UNR], bv[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        ri[u] = relc[r1 + u * BS] + shift;
        sv[u] = src[r1 + u * BS];
      }
#pragma unroll
      for (int u = 0; u < UNR; u++) bv[u] = base[ri[u]];
#pragma unroll
      for (int u = 0; u < UNR; u++) base[ri[u]] = bv[u] + sv[u];
    }
    for (; r1 < nrc; r1 += BS) base[relc[r1] + shift] += src[r1];
    __syncthreads();
  }
}

}  // namespace

// =================================================================== launchers
void extend_add(const CholDev& c, int ntasks, const int* task_p, const int* task_col,
                const int* cptr, const int* cc, const int* ccol, const Batch& b, hipStream_t st)
{
  if (ntasks <= 0 || b.count <= 0) return;
  k_extend_add<<<dim3(ntasks, b.count), 256, 0, st>>>(c, task_p, task_col, cptr, cc, ccol, b);
  HIPCHK(hipGetLastError());
}

}  // namespace dev
}  // namespace hykkt
