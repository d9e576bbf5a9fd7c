// kernels_trsv.hip -- supernodal triangular solves L y = b and L^T x = y as
// two single-launch sweeps over the whole elimination tree (gfx950).
//
// Work items are 128-row blocks of the supernodal fronts, 1024 threads each:
// thread (row = tid % 128, slice q = tid / 128) owns one row (forward) or one
// column (backward) of a 128 x 128 block and a 16-wide slice of the other
// index, so a whole block and an inverted diagonal block fit in registers.
//
// Forward sweep, per supernode s in postorder:
//   pivot item (s, b): waits for s's children's update vectors and for the
//     earlier pivot blocks of s, accumulates L(b, j) y_j as each y_j becomes
//     ready, and finishes with the inverted diagonal block:
//     y_b = Dinv_b (v - acc);
//   update item (s, u): the same products for 128 rows below s's columns,
//     written to s's update vector, which the parent pulls.
// Backward sweep, per supernode from the root down:
//   chunk item (s, u, b): the rows of chunk u of s are ancestors' columns, so
//     their x is final once computed there; the item computes
//     L(rows of chunk u, column block b)^T x(rows) into a partial-sum buffer
//     (all chunks of all column blocks run in parallel);
//   pivot item (s, b): sums the chunk partials in a fixed order, then
//     L(j, b)^T x_j over the later pivot blocks j > b as they finish, and
//     x_b = Dinv_b^T (t - acc).
// The L block of the next step, and the inverted diagonal block before the
// last wait, are loaded into registers before the item waits, so a link of
// the dependency chain costs the flag round trip and arithmetic only.
//
// Items are dispatched in an order where every item depends only on items
// with a smaller index.  Each XCD dispatches its share of the workgroups in
// increasing index order, so the smallest unfinished item is always resident
// and runnable: the sweep cannot deadlock.  Every wait is additionally
// bounded (SPIN_LIMIT); on expiry the item sets an error word the host
// reports, and carries on, so the kernel always terminates.
//
// Synchronisation is by the data itself.  Every value passed between items
// (the forward y blocks, update vectors, backward chunk partials, backward x
// blocks) lives in a buffer that is filled with a sentinel NaN bit pattern
// before the sweep -- a pattern arithmetic never produces -- and is written
// once, with a device-scope atomic store.  A consumer spins on each value it
// needs with device-scope atomic loads until the sentinel is gone.  Each
// value is thus published and observed on its own location, with no flags,
// no fences and no cross-location ordering (the eight XCDs' L2 caches are
// not coherent with each other; device-scope atomics go to the coherent
// level).  Factor values are plain loads: earlier kernels wrote them.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace hykkt {
namespace dev {

namespace {

constexpr int SB = 128;                  // item rows / diagonal block size
// Items of the top of the tree (supernodes wider than one block, fronts
// taller than TOP_F rows, and their ancestors) use NT_TOP threads with
// register prefetch; the small-supernode forest below them uses NT_FOREST.
constexpr int NT_TOP = 1024;
constexpr int NT_FOREST = 256;
constexpr int SPIN_LIMIT = 1 << 22;      // polls of ~100 cycles: ~0.2 s per wait

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

// Sentinel: a signalling-NaN payload no arithmetic result carries.
constexpr uint64_t SENTINEL = 0x7ff4dead0badf00dull;

__device__ __forceinline__ double xload(const double* p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void xstore(double* p, double v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool pending(double v)
{
  return __builtin_bit_cast(uint64_t, v) == SENTINEL;
}

// The value at p once its producer has written it.  Gives up (sets *err,
// returns 0) after SPIN_LIMIT polls, or soon after another wait gave up.
__device__ double spin_load(const double* p, int* err)
{
  double v = xload(p);
  int it = 0;
  while (pending(v)) {
    __builtin_amdgcn_s_sleep(2);
    ++it;
    if (it > SPIN_LIMIT ||
        ((it & 1023) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return 0.0;
    }
    v = xload(p);
  }
  return v;
}

// Fixed-order sum of the NQ slice partials of one row (column).
template <int NQ>
__device__ __forceinline__ double sum_slices(const double (&part)[NQ][SB], int i)
{
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int k = 0; k < NQ; k += 2) {
    a += part[k][i];
    b += part[k + 1][i];
  }
  return a + b;
}

// Buffer loads: a uniform base in a buffer descriptor (SGPRs), a per-lane
// byte offset and a uniform byte offset, so a stretch of strided loads costs
// scalar arithmetic and one offset register instead of one address pair per
// load.  Word 3 of the descriptor is the gfx9-family raw-buffer setting.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int RSRC_W3 = 0x00020000;

__device__ __forceinline__ rsrc_t view(const double* base, uint32_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), 0, bytes, RSRC_W3);
}

__device__ __forceinline__ double ld(rsrc_t r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__global__ void k_fill_sentinel(double* __restrict__ p, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = __builtin_bit_cast(double, SENTINEL);
}

// A slice of QW values, value k at byte offset vo + so0 + k*sstride of the
// view rs, for k < kvalid (uniform); other values and dead lanes read 0.
template <int QW>
__device__ __forceinline__ void slice_load(double (&v)[QW], rsrc_t rs, uint32_t vo, uint32_t so0,
                                           uint32_t sstride, int kvalid, bool live)
{
  const uint32_t vl = live ? vo : 0u;
#pragma unroll
  for (int k = 0; k < QW; k++) {
    const bool in = k < kvalid;
    const double t = ld(rs, vl, in ? so0 + static_cast<uint32_t>(k) * sstride : 0u);
    v[k] = (live && in) ? t : 0.0;
  }
}

template <int QW>
__device__ __forceinline__ double slice_dot(const double (&v)[QW], const double* vec)
{
  double a = 0.0;
#pragma unroll
  for (int k = 0; k < QW; k++) a += v[k] * vec[k];
  return a;
}

// The same dot product with the slice streamed through registers 16 values
// at a time (no prefetch; few registers).
template <int QW>
__device__ __forceinline__ double slice_dot_stream(rsrc_t rs, uint32_t vo, uint32_t so0,
                                                   uint32_t sstride, int kvalid, bool live,
                                                   const double* vec)
{
  constexpr int CH = QW < 16 ? QW : 16;
  double a = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < QW; k0 += CH) {
    double t[CH];
    slice_load<CH>(t, rs, vo, so0 + static_cast<uint32_t>(k0) * sstride, sstride, kvalid - k0,
                   live);
#pragma unroll
    for (int k = 0; k < CH; k++) a += t[k] * vec[k0 + k];
  }
  return a;
}

// ---------------------------------------------------------------- forward
// NT threads per item; PF: the L slice of the next step and the inverted
// diagonal block are prefetched into registers before the waits (top of the
// tree); otherwise slices are streamed after each wait (small supernodes,
// high occupancy).
template <int NT, bool PF>
__global__ void __launch_bounds__(NT) k_solve_fwd(CholDev c, SolveDev sd,
                                                  const int* __restrict__ items,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ y)
{
  constexpr int NQ = NT / SB, QW = SB / NQ;
  __shared__ double yv[SB];
  __shared__ double part[NQ][SB];
  const int s = items[3 * blockIdx.x];
  const int kind = items[3 * blockIdx.x + 1];
  const int blk = items[3 * blockIdx.x + 2];
  const SnInfo si = sninfo(c, s);
  const int r0 = (kind == 0) ? blk * SB : si.nc + blk * SB;
  const int rend = (kind == 0) ? min(si.nc, r0 + SB) : min(si.f, r0 + SB);
  const int nrows = rend - r0;
  const int row = threadIdx.x & (SB - 1);
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);   // whole waves per slice
  const bool live = row < nrows;
  const int r = r0 + row;
  const int jmax = (kind == 0) ? blk : si.np;
  const uint32_t f8 = static_cast<uint32_t>(si.f) * 8u;

  // row r of column block j: L(r, j*128 + q*QW + k)
  auto lview = [&](int j) {
    return view(si.Ls + static_cast<int64_t>(j) * SB * si.f,
                static_cast<uint32_t>(min(SB, si.nc - j * SB)) * f8);
  };
  auto lvalid = [&](int j) { return min(SB, si.nc - j * SB) - q * QW; };
  const uint32_t vo = static_cast<uint32_t>(r) * 8u;
  const uint32_t so = static_cast<uint32_t>(q * QW) * f8;
  // row `row` of the inverted diagonal block (upper part stored as zeros)
  const rsrc_t wv = view(c.dinv + c.dinv_off[s] + static_cast<int64_t>(blk) * SB * SB,
                         SB * SB * 8u);
  const uint32_t wvo = static_cast<uint32_t>(row) * 8u, wso = static_cast<uint32_t>(q * QW) * SB * 8u;

  double l[PF ? QW : 1], w[PF ? QW : 1];
  if constexpr (PF) {
    if (jmax > 0) slice_load<QW>(l, lview(0), vo, so, f8, lvalid(0), live);
    if (kind == 0 && jmax == 0) slice_load<QW>(w, wv, wvo, wso, SB * 8u, nrows - q * QW, live);
  }

  // right-hand side plus the children's update-vector entries for this row
  // (each waited for as it is read)
  double v = 0.0;
  if (q == 0 && live) {
    v = (kind == 0) ? b[si.a + r] : 0.0;
    const int64_t fr = sd.front_off[s] + r;
    for (int p = sd.pull_ptr[fr]; p < sd.pull_ptr[fr + 1]; p++)
      v += spin_load(sd.ubuf + sd.pull_src[p], sd.err);
  }
  // v -= L(r, B_j) y_j over the earlier pivot blocks, as each y_j is ready
  for (int j = 0; j < jmax; j++) {
    if constexpr (PF) {
      if (kind == 0 && j == jmax - 1)   // in flight during the last wait
        slice_load<QW>(w, wv, wvo, wso, SB * 8u, nrows - q * QW, live);
    }
    const int bj = min(SB, si.nc - j * SB);
    if (threadIdx.x < SB)
      yv[threadIdx.x] = (threadIdx.x < bj) ? spin_load(y + si.a + j * SB + threadIdx.x, sd.err) : 0.0;
    __syncthreads();
    double tp;
    if constexpr (PF) tp = slice_dot<QW>(l, yv + q * QW);
    else tp = slice_dot_stream<QW>(lview(j), vo, so, f8, lvalid(j), live, yv + q * QW);
    part[q][row] = tp;
    __syncthreads();
    // the barrier keeps the next block's loads after the last use of l
    if constexpr (PF) {
      if (j + 1 < jmax) slice_load<QW>(l, lview(j + 1), vo, so, f8, lvalid(j + 1), live);
    }
    if (q == 0) v -= sum_slices<NQ>(part, row);
    __syncthreads();
  }
  if (kind == 0) {
    // y = Dinv_blk v
    if (q == 0) yv[row] = live ? v : 0.0;
    __syncthreads();
    double res;
    if constexpr (PF) res = slice_dot<QW>(w, yv + q * QW);
    else res = slice_dot_stream<QW>(wv, wvo, wso, SB * 8u, nrows - q * QW, live, yv + q * QW);
    part[q][row] = res;
    __syncthreads();
    if (q == 0 && live) xstore(y + si.a + r, sum_slices<NQ>(part, row));
  } else {
    if (q == 0 && live) xstore(sd.ubuf + sd.uvec_off[s] + (r - si.nc), v);
  }
}

// ---------------------------------------------------------------- backward
template <int NT, bool PF>
__global__ void __launch_bounds__(NT) k_solve_bwd(CholDev c, SolveDev sd,
                                                  const int* __restrict__ items,
                                                  const double* __restrict__ y,
                                                  double* __restrict__ x)
{
  constexpr int NQ = NT / SB, QW = SB / NQ;
  __shared__ double xs[SB];
  __shared__ double part[NQ][SB];
  const int s = items[3 * blockIdx.x];
  const int kind = items[3 * blockIdx.x + 1];
  const int idx = items[3 * blockIdx.x + 2];
  const SnInfo si = sninfo(c, s);
  const int col = threadIdx.x & (SB - 1);
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);   // whole waves per slice
  const uint32_t f8 = static_cast<uint32_t>(si.f) * 8u;
  // column `col` of column block starting at c0 (nb wide): L(row, c0 + col)
  auto cview = [&](int c0, int nb) {
    return view(si.Ls + static_cast<int64_t>(c0) * si.f, static_cast<uint32_t>(nb) * f8);
  };
  const uint32_t cvo = static_cast<uint32_t>(col) * f8;

  if (kind == 0) {
    // ---- chunk item (u, b): rows [nc + u*128, ...) of s against column
    // block b: partial L(rows, b-block)^T x(rows)
    const int u = idx / si.np, b = idx - u * si.np;
    const int rr0 = u * SB;
    const int nrows = min(SB, si.nr - rr0);
    const int bb = min(SB, si.nc - b * SB);
    const int* rows = c.rows + c.rows_ptr[s] + rr0;
    const uint32_t so = static_cast<uint32_t>(si.nc + rr0 + q * QW) * 8u;
    const rsrc_t lv = cview(b * SB, bb);
    double l[PF ? QW : 1];
    if constexpr (PF) slice_load<QW>(l, lv, cvo, so, 8u, nrows - q * QW, col < bb);
    // x of these rows belongs to ancestors: waited for as it is read
    if (threadIdx.x < SB)
      xs[threadIdx.x] = (threadIdx.x < nrows) ? spin_load(x + rows[threadIdx.x], sd.err) : 0.0;
    __syncthreads();
    double tp;
    if constexpr (PF) tp = slice_dot<QW>(l, xs + q * QW);
    else tp = slice_dot_stream<QW>(lv, cvo, so, 8u, nrows - q * QW, col < bb, xs + q * QW);
    part[q][col] = tp;
    __syncthreads();
    double* out = sd.bpart + sd.bpart_off[s] + static_cast<int64_t>(u) * si.nc + b * SB;
    if (q == 0 && col < bb) xstore(out + col, sum_slices<NQ>(part, col));
    return;
  }

  // ---- pivot item: columns [c0, c0 + nb) of s
  const int blk = idx;
  const int c0 = blk * SB;
  const int nb = min(SB, si.nc - c0);
  const bool live = col < nb;
  const rsrc_t lv = cview(c0, nb);
  auto lso = [&](int j) { return static_cast<uint32_t>(j * SB + q * QW) * 8u; };
  auto lvalid = [&](int j) { return min(SB, si.nc - j * SB) - q * QW; };
  // Dinv_blk^T: x(col) = sum_{row >= col} W(row, col) t(row): column `col` of
  // W (its upper part is stored as zeros)
  const rsrc_t wv = view(c.dinv + c.dinv_off[s] + static_cast<int64_t>(blk) * SB * SB,
                         SB * SB * 8u);
  const uint32_t wvo = static_cast<uint32_t>(col) * SB * 8u, wso = static_cast<uint32_t>(q * QW) * 8u;
  const int jtop = si.np - 1;   // later pivot blocks j = jtop .. blk+1
  double l[PF ? QW : 1], w[PF ? QW : 1];
  if constexpr (PF) {
    if (jtop > blk) slice_load<QW>(l, lv, cvo, lso(jtop), 8u, lvalid(jtop), live);
    else slice_load<QW>(w, wv, wvo, wso, 8u, nb - q * QW, live);
  }

  double t = 0.0;
  if (q == 0 && live) {
    t = y[si.a + c0 + col];
    const double* bp = sd.bpart + sd.bpart_off[s] + c0 + col;
    for (int u = 0; u < si.nu; u++) t -= spin_load(bp + static_cast<int64_t>(u) * si.nc, sd.err);
  }
  for (int j = jtop; j > blk; j--) {
    if constexpr (PF) {
      if (j == blk + 1) slice_load<QW>(w, wv, wvo, wso, 8u, nb - q * QW, live);   // last wait
    }
    const int bj = min(SB, si.nc - j * SB);
    if (threadIdx.x < SB)
      xs[threadIdx.x] = (threadIdx.x < bj) ? spin_load(x + si.a + j * SB + threadIdx.x, sd.err) : 0.0;
    __syncthreads();
    double tp;
    if constexpr (PF) tp = slice_dot<QW>(l, xs + q * QW);
    else tp = slice_dot_stream<QW>(lv, cvo, lso(j), 8u, lvalid(j), live, xs + q * QW);
    part[q][col] = tp;
    __syncthreads();
    if constexpr (PF) {
      if (j - 1 > blk) slice_load<QW>(l, lv, cvo, lso(j - 1), 8u, lvalid(j - 1), live);
    }
    if (q == 0) t -= sum_slices<NQ>(part, col);
    __syncthreads();
  }
  if (q == 0) xs[col] = live ? t : 0.0;
  __syncthreads();
  double res;
  if constexpr (PF) res = slice_dot<QW>(w, xs + q * QW);
  else res = slice_dot_stream<QW>(wv, wvo, wso, 8u, nb - q * QW, live, xs + q * QW);
  part[q][col] = res;
  __syncthreads();
  if (q == 0 && live) xstore(x + si.a + c0 + col, sum_slices<NQ>(part, col));
}

}  // namespace

void fill_sentinel(double* p, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  k_fill_sentinel<<<static_cast<unsigned>(g), 256, 0, st>>>(p, n);
  HIPCHK(hipGetLastError());
}

void solve_forward(const CholDev& c, const SolveDev& sd, const double* b, double* y, int part,
                   hipStream_t st)
{
  if (part == 0 && sd.n_fwd > 0)
    k_solve_fwd<NT_FOREST, false><<<sd.n_fwd, NT_FOREST, 0, st>>>(c, sd, sd.fwd_items, b, y);
  if (part == 1 && sd.n_fwd_top > 0)
    k_solve_fwd<NT_TOP, true><<<sd.n_fwd_top, NT_TOP, 0, st>>>(c, sd, sd.fwd_items_top, b, y);
  HIPCHK(hipGetLastError());
}

void solve_backward(const CholDev& c, const SolveDev& sd, const double* y, double* x, int part,
                    hipStream_t st)
{
  if (part == 1 && sd.n_bwd_top > 0)
    k_solve_bwd<NT_TOP, true><<<sd.n_bwd_top, NT_TOP, 0, st>>>(c, sd, sd.bwd_items_top, y, x);
  if (part == 0 && sd.n_bwd > 0)
    k_solve_bwd<NT_FOREST, false><<<sd.n_bwd, NT_FOREST, 0, st>>>(c, sd, sd.bwd_items, y, x);
  HIPCHK(hipGetLastError());
}

}  // namespace dev
}  // namespace hykkt
