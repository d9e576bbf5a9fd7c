// kernels_trsv.hip -- supernodal triangular solves L y = b and L^T x = y as
// two single-launch sweeps over the whole elimination tree (gfx950).
//
// Work items are 128-row blocks of the supernodal fronts.  In the forward
// sweep a pivot block of supernode s (rows of its own columns) waits for the
// update vectors of s's children and for the earlier pivot blocks of s,
// accumulates L(blk, j) * y_j as those become ready, and finishes with the
// inverted diagonal block: y_blk = Dinv_blk (v - acc).  An update block
// (rows of s below its columns) accumulates the same products into s's
// update vector, which the parent pulls.  The backward sweep runs the
// transpose from the root down.
//
// Items are dispatched in an order where every item depends only on items
// with a smaller index.  The GPU dispatches the workgroups of each XCD in
// increasing index order, so the smallest unfinished item is always resident
// and never blocked: the sweep cannot deadlock.  Every wait is additionally
// bounded (SPIN_LIMIT); on expiry the item sets an error flag the host
// reports, and carries on, so the kernel always terminates.
//
// Cross-workgroup visibility: the producer makes its stores visible with a
// device-scope fence before publishing its flag with a device-scope atomic;
// the consumer fences after observing the flag.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace hykkt {
namespace dev {

namespace {

constexpr int SB = 128;                   // item rows / diagonal block size
constexpr int SPIN_LIMIT = 1 << 22;   // polls of ~100 cycles: ~0.2 s per wait

struct SnInfo {
  int a, nc, nr, f, np, nu;
  const double* Ls;
};

__device__ __forceinline__ SnInfo sninfo(const CholDev& c, int s)
{
  SnInfo i;
  i.a = c.sn_first[s];
  i.nc = c.sn_first[s + 1] - i.a;
  i.nr = c.rows_ptr[s + 1] - c.rows_ptr[s];
  i.f = i.nc + i.nr;
  i.np = (i.nc + SB - 1) / SB;
  i.nu = (i.nr + SB - 1) / SB;
  i.Ls = c.L + c.l_off[s];
  return i;
}

// Block until *p >= target: thread 0 spins on a device-scope atomic load,
// then every thread acquires at device scope.  Gives up (and sets *err) after
// SPIN_LIMIT polls, or at once if another item already gave up.
__device__ void wait_geq(int* p, int target, int* err)
{
  if (threadIdx.x == 0) {
    int it = 0;
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++it > SPIN_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Publish: every thread releases its stores at device scope, then one thread
// increments the flag.
__device__ void signal_add(int* p, int* p2 = nullptr)
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p2) __hip_atomic_fetch_add(p2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ __forceinline__ double wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__global__ void k_solve_reset(int* flags, int n)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = 0;
}

// ---------------------------------------------------------------- forward
__global__ void __launch_bounds__(256) k_solve_fwd(CholDev c, SolveDev sd, double* __restrict__ x)
{
  __shared__ double yv[SB];
  __shared__ double part[SB];
  const int s = sd.fwd_items[3 * blockIdx.x];
  const int kind = sd.fwd_items[3 * blockIdx.x + 1];
  const int blk = sd.fwd_items[3 * blockIdx.x + 2];
  const SnInfo si = sninfo(c, s);
  int* fflag = sd.flags;                              // forward pivot-block flags
  int* fcount = sd.flags + 2 * sd.n_pblk;             // forward update-block counters
  int* err = sd.flags + 2 * sd.n_pblk + 2 * c.ns;
  const int r0 = (kind == 0) ? blk * SB : si.nc + blk * SB;
  const int rend = (kind == 0) ? min(si.nc, r0 + SB) : min(si.f, r0 + SB);
  const int nrows = rend - r0;
  const int row = threadIdx.x & (SB - 1), half = threadIdx.x >> 7;
  const int r = r0 + row;

  // children's update vectors must be complete before they are pulled
  for (int t = c.child_ptr[s]; t < c.child_ptr[s + 1]; t++) {
    const int ch = c.child_list[t];
    const int nu_ch = (c.rows_ptr[ch + 1] - c.rows_ptr[ch] + SB - 1) / SB;
    wait_geq(fcount + ch, nu_ch, err);
  }
  double v = 0.0;
  if (half == 0 && row < nrows) {
    v = (kind == 0) ? x[si.a + r] : 0.0;
    const int64_t fr = sd.front_off[s] + r;
    for (int q = sd.pull_ptr[fr]; q < sd.pull_ptr[fr + 1]; q++) v += sd.ubuf[sd.pull_src[q]];
  }
  // v -= L(r, B_j) y_j for the earlier pivot blocks
  const int jmax = (kind == 0) ? blk : si.np;
  for (int j = 0; j < jmax; j++) {
    wait_geq(fflag + sd.pblk_off[s] + j, 1, err);
    const int bj = min(SB, si.nc - j * SB);
    if (threadIdx.x < SB) yv[threadIdx.x] = (threadIdx.x < bj) ? x[si.a + j * SB + threadIdx.x] : 0.0;
    __syncthreads();
    double tp = 0.0;
    if (row < nrows) {
      // 64 independent loads per thread (one column of L each), then the FMAs
      const double* col = si.Ls + static_cast<int64_t>(j * SB + half * 64) * si.f + r;
      const int cn = min(64, bj - half * 64);
      double lv[64];
#pragma unroll
      for (int q = 0; q < 64; q++) lv[q] = (q < cn) ? col[static_cast<int64_t>(q) * si.f] : 0.0;
#pragma unroll
      for (int q = 0; q < 64; q++) tp += lv[q] * yv[half * 64 + q];
    }
    if (half == 1) part[row] = tp;
    __syncthreads();
    if (half == 0 && row < nrows) v -= tp + part[row];
    __syncthreads();
  }
  if (kind == 0) {
    // y = Dinv_blk * v  (Dinv lower triangular, column-major 128 x 128)
    if (half == 0) part[row] = (row < nrows) ? v : 0.0;
    __syncthreads();
    const double* W = c.dinv + c.dinv_off[s] + static_cast<int64_t>(blk) * SB * SB;
    // y(row) = sum_{cc <= row} W(row, cc) v(cc): the two halves of the
    // threads take the columns [0,64) and [64,128); loads issued 32 at a time
    double y = 0.0;
    if (row < nrows) {
      const double* wr = W + row + static_cast<int64_t>(half * 64) * SB;
#pragma unroll
      for (int q0 = 0; q0 < 64; q0 += 32) {
        double w[32];
#pragma unroll
        for (int q = 0; q < 32; q++) w[q] = wr[static_cast<int64_t>(q0 + q) * SB];
#pragma unroll
        for (int q = 0; q < 32; q++) y += w[q] * part[half * 64 + q0 + q];
      }
    }
    __syncthreads();
    if (half == 1) yv[row] = y;
    __syncthreads();
    if (half == 0 && row < nrows) x[si.a + r] = y + yv[row];
    signal_add(fflag + sd.pblk_off[s] + blk);
  } else {
    if (half == 0 && row < nrows) sd.ubuf[sd.uvec_off[s] + (r - si.nc)] = v;
    signal_add(fcount + s);
  }
}

// ---------------------------------------------------------------- backward
// Partial of sum_{k < nrow} L(row0 + k, c0 + cc) * xs[k] for this thread's
// column cc = tid % 128 over its half of the rows.  Lp points at (row0, c0)
// of a column-major panel with leading dimension f.  Each thread walks a
// contiguous stretch of its column, 16 independent loads at a time (one
// 128-byte line), so many lines are in flight per workgroup.
__device__ __forceinline__ double gemv_t_cols(const double* Lp, int64_t f, int nrow, int ncol,
                                              const double* xs)
{
  const int cc = threadIdx.x & (SB - 1), half = threadIdx.x >> 7;
  if (cc >= ncol) return 0.0;
  const int h = ((nrow + 31) / 32) * 16;          // split on a 16-row boundary
  const int rb = half ? min(h, nrow) : 0, re = half ? nrow : min(h, nrow);
  const double* col = Lp + static_cast<int64_t>(cc) * f;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  int k = rb;
  for (; k + 16 <= re; k += 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = col[k + q];
#pragma unroll
    for (int q = 0; q < 16; q++) a[q & 3] += v[q] * xs[k + q];
  }
  for (; k < re; k++) a[0] += col[k] * xs[k];
  return (a[0] + a[1]) + (a[2] + a[3]);
}

__global__ void __launch_bounds__(256) k_solve_bwd(CholDev c, SolveDev sd, double* __restrict__ x)
{
  __shared__ double t[SB];
  __shared__ double part[SB];
  __shared__ double xs[1024];
  const int s = sd.bwd_items[2 * blockIdx.x];
  const int blk = sd.bwd_items[2 * blockIdx.x + 1];
  const SnInfo si = sninfo(c, s);
  int* bflag = sd.flags + sd.n_pblk;                  // backward pivot-block flags
  int* bcount = sd.flags + 2 * sd.n_pblk + c.ns;      // backward per-supernode counters
  int* err = sd.flags + 2 * sd.n_pblk + 2 * c.ns;
  const int c0 = blk * SB;
  const int nb = min(SB, si.nc - c0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cc = threadIdx.x & (SB - 1), half = threadIdx.x >> 7;
  const int* rows = c.rows + c.rows_ptr[s];

  const int p = c.sn_parent[s];
  if (p >= 0) {
    const int np_p = (c.sn_first[p + 1] - c.sn_first[p] + SB - 1) / SB;
    wait_geq(bcount + p, np_p, err);
  }
  if (threadIdx.x < SB) t[threadIdx.x] = (threadIdx.x < nb) ? x[si.a + c0 + threadIdx.x] : 0.0;
  double acc = 0.0;
  // L21(:, block)^T x(rows), x(rows) gathered through LDS 1024 at a time
  for (int k0 = 0; k0 < si.nr; k0 += 1024) {
    const int kn = min(1024, si.nr - k0);
    __syncthreads();
    for (int k = threadIdx.x; k < kn; k += blockDim.x) xs[k] = x[rows[k0 + k]];
    __syncthreads();
    acc += gemv_t_cols(si.Ls + static_cast<int64_t>(c0) * si.f + si.nc + k0, si.f, kn, nb, xs);
  }
  // L(B_j, block)^T x_j for the later pivot blocks of s
  for (int j = si.np - 1; j > blk; j--) {
    wait_geq(bflag + sd.pblk_off[s] + j, 1, err);
    const int bj = min(SB, si.nc - j * SB);
    for (int k = threadIdx.x; k < bj; k += blockDim.x) xs[k] = x[si.a + j * SB + k];
    __syncthreads();
    acc += gemv_t_cols(si.Ls + static_cast<int64_t>(c0) * si.f + j * SB, si.f, bj, nb, xs);
    __syncthreads();
  }
  if (half == 1) part[cc] = acc;
  __syncthreads();
  if (half == 0) t[cc] -= acc + part[cc];
  __syncthreads();
  // x_block = Dinv^T t: column q of W (contiguous) dotted with t, two threads
  // per column, 16 independent loads at a time
  const double* W = c.dinv + c.dinv_off[s] + static_cast<int64_t>(blk) * SB * SB;
  {
    const double* wc = W + static_cast<int64_t>(cc) * SB + half * 64;
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int q0 = 0; q0 < 64; q0 += 16) {
      double w[16];
#pragma unroll
      for (int q = 0; q < 16; q++) w[q] = wc[q0 + q];
#pragma unroll
      for (int q = 0; q < 16; q += 2) {
        a0 += w[q] * t[half * 64 + q0 + q];
        a1 += w[q + 1] * t[half * 64 + q0 + q + 1];
      }
    }
    if (half == 1) part[cc] = a0 + a1;
    __syncthreads();
    if (half == 0 && cc < nb) x[si.a + c0 + cc] = a0 + a1 + part[cc];
  }
  signal_add(bflag + sd.pblk_off[s] + blk, bcount + s);
}

}  // namespace

void solve_reset(const SolveDev& sd, int ns, hipStream_t st)
{
  const int n = 2 * sd.n_pblk + 2 * ns + 1;
  k_solve_reset<<<(n + 255) / 256, 256, 0, st>>>(sd.flags, n);
  HIPCHK(hipGetLastError());
}

void solve_forward(const CholDev& c, const SolveDev& sd, double* x, hipStream_t st)
{
  if (sd.n_fwd <= 0) return;
  k_solve_fwd<<<sd.n_fwd, 256, 0, st>>>(c, sd, x);
  HIPCHK(hipGetLastError());
}

void solve_backward(const CholDev& c, const SolveDev& sd, double* x, hipStream_t st)
{
  if (sd.n_bwd <= 0) return;
  k_solve_bwd<<<sd.n_bwd, 256, 0, st>>>(c, sd, x);
  HIPCHK(hipGetLastError());
}

}  // namespace dev
}  // namespace hykkt
