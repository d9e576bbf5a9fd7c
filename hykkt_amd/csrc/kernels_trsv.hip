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
//     x_b = Dinv_b^T (t - acc);
//   a one-block supernode of the small-supernode forest with at most two row
//     chunks has no chunk items: its pivot item forms the chunks' products
//     itself, in the same order and arithmetic, so the backward forest
//     carries one dependency link per tree level instead of two.
// The L block of the next step, and the inverted diagonal block before the
// last wait, are loaded into registers before the item waits, so a link of
// the dependency chain costs the flag round trip and arithmetic only.
//
// Items are numbered so that every item depends only on items with a
// smaller number, and a workgroup takes its item from a counter when it
// starts running (take_item), so the smallest unfinished item is always
// held by a running workgroup: the sweep cannot deadlock, also while other
// streams' kernels share the device.  Every wait is additionally bounded
// (SPIN_LIMIT); on expiry the item sets an error word the host reports, and
// carries on, so the kernel always terminates.
//
// Several systems of one pattern share a launch: launch item i*count + k is
// item i of system k, whose buffers (factor, right-hand side, exchanged
// values) sit k*vs doubles from system 0's.  The order property above holds
// for the combined numbering, and a system the mask skips returns at once.
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

#include <algorithm>

#include "batch_dev.hpp"
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
constexpr int SPIN_LIMIT = 1 << 22;      // polls; each sleeps >= 128 clocks and waits one
                                         // device-scope load: seconds per wait at most

struct SnInfo {
  int a, nc, nr, f, np, nu;
  const double* Ls;
};

// Wave-uniform copies (scalar registers) of values every lane loaded alike.
// Buffer descriptors built from them stay scalar; a descriptor the compiler
// cannot prove uniform turns every buffer load into a serialising loop.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v)
{
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(v) >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Supernode s (wave-uniform) of the factor L.
__device__ __forceinline__ SnInfo sninfo(const CholDev& c, const double* L, int s)
{
  SnInfo i;
  i.a = uni(c.sn_first[s]);
  i.nc = uni(c.sn_first[s + 1]) - i.a;
  i.nr = uni(c.rows_ptr[s + 1]) - uni(c.rows_ptr[s]);
  i.f = i.nc + i.nr;
  i.np = (i.nc + SB - 1) / SB;
  i.nu = (i.nr + SB - 1) / SB;
  i.Ls = L + uni64(c.l_off[s]);
  return i;
}

// One system's buffers in a sweep (wave-uniform pointers).
struct SysBufs {
  const double* L;
  const double* dinv;
  double* ubuf;
  double* bpart;
  const double* in;    // right-hand side (forward) or y (backward)
  double* out;         // y (forward) or x (backward)
};

__device__ __forceinline__ SysBufs sys_bufs(const CholDev& c, const SolveDev& sd, int sy,
                                            const Batch& bt, const double* in, double* out)
{
  const int64_t o = static_cast<int64_t>(sy) * bt.vs;
  return SysBufs{c.L + o, c.dinv + o, sd.ubuf + o, sd.bpart + o, in + o, out + o};
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

// The launch's next item, in the order workgroups start running.  Item i
// goes to a workgroup only after items 0..i-1 went to workgroups that were
// already running, so the smallest unfinished item is always held by a
// running workgroup whose producers have all finished: the sweep progresses
// whatever else shares the device (kernels of other streams may hold the
// compute units, and workgroup dispatch order alone would not guarantee it).
__device__ __forceinline__ int take_item(int* ticket)
{
  __shared__ int item;
  if (threadIdx.x == 0)
    item = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return uni(item);
}

// Fixed-order sum of the NQ slice partials of one row (column).
template <int NQ>
__device__ __forceinline__ double sum_slices(const double (*part)[SB], int i)
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

// Up to three buffers in one launch: element i of the concatenation
// p0[0..n0) p1[0..n1) p2[0..n2) takes the sentinel.
__global__ void k_fill_sentinel(double* __restrict__ p0, int64_t n0, double* __restrict__ p1,
                                int64_t n1, double* __restrict__ p2, int64_t n2,
                                int* __restrict__ zero4, Batch bt)
{
  if (zero4 && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4)
    __hip_atomic_store(zero4 + threadIdx.x, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int sy = batch_sys(bt);
  if (sy < 0) return;
  p0 = batch_at(p0, sy, bt);
  p1 = batch_at(p1, sy, bt);
  p2 = batch_at(p2, sy, bt);
  const int64_t n = n0 + n1 + n2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double* p = i < n0 ? p0 + i : (i < n0 + n1 ? p1 + (i - n0) : p2 + (i - n0 - n1));
    *p = __builtin_bit_cast(double, SENTINEL);
  }
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
// diagonal block are prefetched into registers before the waits (the top of
// the tree, NT_TOP); otherwise slices are streamed after each wait.  The
// small-supernode forest has its own persistent kernel (k_fwd_forest).
template <int NT, bool PF>
__device__ __forceinline__ void fwd_item(const CholDev& c, const SolveDev& sd,
                                         const int* __restrict__ items,
                                         const double* __restrict__ rhs,
                                         double* __restrict__ yout, const Batch& bt, int itc,
                                         double* yv, double (*part)[SB])
{
  constexpr int NQ = NT / SB, QW = SB / NQ;
  const int sy = itc % bt.count, it = itc / bt.count;
  if (bt.mask && !bt.mask[sy]) return;
  const SysBufs B = sys_bufs(c, sd, sy, bt, rhs, yout);
  const double* b = B.in;
  double* y = B.out;
  const int s = uni(items[3 * it]);
  const int kind = uni(items[3 * it + 1]);
  const int blk = uni(items[3 * it + 2]);
  const SnInfo si = sninfo(c, B.L, s);
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
  const rsrc_t wv = view(B.dinv + uni64(c.dinv_off[s]) + static_cast<int64_t>(blk) * SB * SB,
                         SB * SB * 8u);
  const uint32_t wvo = static_cast<uint32_t>(row) * 8u, wso = static_cast<uint32_t>(q * QW) * SB * 8u;

  // Both slices are loaded up front: a load still in flight when the wave
  // polls for a value delays the poll (loads complete in order).
  double l[PF ? QW : 1], w[PF ? QW : 1];
  if constexpr (PF) {
    if (jmax > 0) slice_load<QW>(l, lview(0), vo, so, f8, lvalid(0), live);
    if (kind == 0) slice_load<QW>(w, wv, wvo, wso, SB * 8u, nrows - q * QW, live);
  }

  // right-hand side plus the children's update-vector entries for this row
  // (each waited for as it is read)
  double v = 0.0;
  if (q == 0 && live) {
    v = (kind == 0) ? b[si.a + r] : 0.0;
    const int64_t fr = sd.front_off[s] + r;
    for (int p = sd.pull_ptr[fr]; p < sd.pull_ptr[fr + 1]; p++)
      v += spin_load(B.ubuf + sd.pull_src[p], sd.err);
  }
  // v -= L(r, B_j) y_j over the earlier pivot blocks, as each y_j is ready.
  // Step j's slice products go to part[j & 1]: it is written again two steps
  // on, after the next step's barriers, so a step needs two barriers.
  for (int j = 0; j < jmax; j++) {
    const int bj = min(SB, si.nc - j * SB);
    double (*pj)[SB] = part + (j & 1) * NQ;
    if (threadIdx.x < SB)
      yv[threadIdx.x] = (threadIdx.x < bj) ? spin_load(y + si.a + j * SB + threadIdx.x, sd.err) : 0.0;
    __syncthreads();
    double tp;
    if constexpr (PF) tp = slice_dot<QW>(l, yv + q * QW);
    else tp = slice_dot_stream<QW>(lview(j), vo, so, f8, lvalid(j), live, yv + q * QW);
    pj[q][row] = tp;
    __syncthreads();
    // the barrier keeps the next block's loads after the last use of l
    if constexpr (PF) {
      if (j + 1 < jmax) slice_load<QW>(l, lview(j + 1), vo, so, f8, lvalid(j + 1), live);
    }
    if (q == 0) v -= sum_slices<NQ>(pj, row);
  }
  __syncthreads();   // part and yv are written again below
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
    if (q == 0 && live) xstore(B.ubuf + sd.uvec_off[s] + (r - si.nc), v);
  }
}

// LOOP: persistent; a workgroup takes items from the sweep's counter until
// they run out, so the grid may be smaller than the item list (fewer resident
// workgroups leave compute units to other streams); the order argument above
// holds for any grid size.  Otherwise one item per workgroup, one workgroup
// per item (fewer registers on the item's path).
template <int NT, bool PF, bool LOOP>
__global__ void __launch_bounds__(NT) k_solve_fwd(CholDev c, SolveDev sd,
                                                  const int* __restrict__ items,
                                                  const double* __restrict__ rhs,
                                                  double* __restrict__ yout, Batch bt, int total)
{
  constexpr int NQ = NT / SB;
  __shared__ double yv[SB];
  __shared__ double part[2 * NQ][SB];
  if constexpr (!LOOP) {
    fwd_item<NT, PF>(c, sd, items, rhs, yout, bt, take_item(sd.ticket + (PF ? 1 : 0)), yv, part);
    return;
  }
  for (;;) {
    const int itc = take_item(sd.ticket + (PF ? 1 : 0));
    if (itc >= total) break;
    fwd_item<NT, PF>(c, sd, items, rhs, yout, bt, itc, yv, part);
    __syncthreads();   // the item's last reads of yv and part
  }
}

// ------------------------------------------------------- forest, persistent
// The forward sweep of the small-supernode forest runs on resident
// workgroups that loop over items.  A workgroup takes its next item as soon as the current one has
// all its inputs (past its waits): it never holds an item it cannot start, so
// the smallest unfinished item is still always being worked on, and the
// counter's round trip overlaps the item's arithmetic.
__device__ __forceinline__ int ticket_add(int* ticket)
{
  return __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Forest supernodes are one block wide (nc <= 128): a pivot item applies the
// inverted diagonal block, an update item the one L block.  The item's slice
// is loaded before its waits.
__global__ void __launch_bounds__(NT_FOREST) k_fwd_forest(CholDev c, SolveDev sd,
                                                          const int* __restrict__ items,
                                                          const double* __restrict__ rhs,
                                                          double* __restrict__ yout, Batch bt)
{
  constexpr int NQ = NT_FOREST / SB, QW = SB / NQ;
  __shared__ double yv[SB];
  __shared__ double part[NQ][SB];
  __shared__ int nxt;
  int* const ticket = sd.ticket + 0;
  const int total = sd.n_fwd * bt.count;
  if (threadIdx.x == 0) nxt = ticket_add(ticket);
  __syncthreads();
  int itc = uni(nxt);
  while (itc < total) {
    const int sy = itc % bt.count, it = itc / bt.count;
    if (bt.mask && !bt.mask[sy]) {
      // a skipped system: take the next item
      __syncthreads();   // every thread has read nxt
      if (threadIdx.x == 0) nxt = ticket_add(ticket);
      __syncthreads();
      itc = uni(nxt);
      continue;
    }
    const SysBufs B = sys_bufs(c, sd, sy, bt, rhs, yout);
    const double* b = B.in;
    double* y = B.out;
    const int s = uni(items[3 * it]);
    const int kind = uni(items[3 * it + 1]);
    const int blk = uni(items[3 * it + 2]);
    const SnInfo si = sninfo(c, B.L, s);
    const int r0 = (kind == 0) ? blk * SB : si.nc + blk * SB;
    const int rend = (kind == 0) ? min(si.nc, r0 + SB) : min(si.f, r0 + SB);
    const int nrows = rend - r0;
    const int row = threadIdx.x & (SB - 1);
    const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);   // whole waves per slice
    const bool live = row < nrows;
    const int r = r0 + row;
    const uint32_t f8 = static_cast<uint32_t>(si.f) * 8u;
    double buf[QW];
    if (kind == 0) {
      // row `row` of the inverted diagonal block (upper part stored as zeros)
      const rsrc_t wv = view(B.dinv + uni64(c.dinv_off[s]) + static_cast<int64_t>(blk) * SB * SB,
                             SB * SB * 8u);
      slice_load<QW>(buf, wv, static_cast<uint32_t>(row) * 8u,
                     static_cast<uint32_t>(q * QW) * SB * 8u, SB * 8u, nrows - q * QW, live);
    } else {
      // row r of the one column block: L(r, q*QW + k)
      const rsrc_t lv = view(si.Ls, static_cast<uint32_t>(min(SB, si.nc)) * f8);
      slice_load<QW>(buf, lv, static_cast<uint32_t>(r) * 8u, static_cast<uint32_t>(q * QW) * f8, f8,
                     min(SB, si.nc) - q * QW, live);
    }
    double v = 0.0;
    if (q == 0 && live) {
      v = (kind == 0) ? b[si.a + r] : 0.0;
      const int64_t fr = sd.front_off[s] + r;
      for (int p = sd.pull_ptr[fr]; p < sd.pull_ptr[fr + 1]; p++)
        v += spin_load(B.ubuf + sd.pull_src[p], sd.err);
    }
    if (kind == 1) {
      if (threadIdx.x < SB)
        yv[threadIdx.x] = (threadIdx.x < si.nc) ? spin_load(y + si.a + threadIdx.x, sd.err) : 0.0;
    } else if (q == 0) {
      yv[row] = live ? v : 0.0;
    }
    __syncthreads();
    // every input of this item is in: take the next item now
    int nx = 0;
    if (threadIdx.x == 0) nx = ticket_add(ticket);
    part[q][row] = slice_dot<QW>(buf, yv + q * QW);
    __syncthreads();
    if (q == 0 && live) {
      const double t = sum_slices<NQ>(part, row);
      if (kind == 0) xstore(y + si.a + r, t);
      else xstore(B.ubuf + sd.uvec_off[s] + (r - si.nc), v - t);
    }
    if (threadIdx.x == 0) nxt = nx;
    __syncthreads();
    itc = uni(nxt);
  }
}

// ---------------------------------------------------------------- backward
// Column sums of a 128 x 128 block product over 1024 threads: thread
// (row r = tid % 128, column slice q = tid / 128) holds acc[k], its row's
// term for column 16 q + k.  Within each wave the 16 values are
// transpose-reduced over the 64 lanes (each exchange halves the values a
// lane keeps: the lanes with bit M of the lane index set keep the upper
// half, the others the lower half, and each adds its partner's, lane
// l ^ M), then the two row halves of a slice meet in LDS.  The exchanges run
// on the VALU: the lane-half swaps of gfx950 (v_permlane32_swap,
// v_permlane16_swap) for M = 32 and 16, DPP row rotations and shifts for
// M = 8 and 4, DPP quad permutations for the last two full sums.
__device__ __forceinline__ uint32_t lo32(double v)
{
  return static_cast<uint32_t>(__builtin_bit_cast(uint64_t, v));
}
__device__ __forceinline__ uint32_t hi32(double v)
{
  return static_cast<uint32_t>(__builtin_bit_cast(uint64_t, v) >> 32);
}
__device__ __forceinline__ double mk64(uint32_t lo, uint32_t hi)
{
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// One exchange for M = 32 (W32) or 16 over the pairs (acc[i], acc[i + N/2]):
// the swap puts a lane's kept value and its partner's sent value in the
// two registers, whose sum is keep + partner's value (addition commutes).
template <int N, bool W32>
__device__ __forceinline__ void halve_swap(double (&acc)[16])
{
#pragma unroll
  for (int i = 0; i < N / 2; i++) {
    const double x = acc[i], y = acc[i + N / 2];
    const auto lo = W32 ? __builtin_amdgcn_permlane32_swap(lo32(x), lo32(y), false, false)
                        : __builtin_amdgcn_permlane16_swap(lo32(x), lo32(y), false, false);
    const auto hi = W32 ? __builtin_amdgcn_permlane32_swap(hi32(x), hi32(y), false, false)
                        : __builtin_amdgcn_permlane16_swap(hi32(x), hi32(y), false, false);
    acc[i] = mk64(lo[0], hi[0]) + mk64(lo[1], hi[1]);
  }
}

// The value of lane l ^ M (M = 8, 4, 2 or 1) by DPP: row_ror:8 for 8;
// row_shr:4 / row_shl:4 by the lane's bit 2 for 4; quad_perm for 2 and 1.
template <int M>
__device__ __forceinline__ double xor_dpp(double v, int lane)
{
  const uint32_t a = lo32(v), b = hi32(v);
  if constexpr (M == 8) {
    return mk64(__builtin_amdgcn_update_dpp(0u, a, 0x128, 0xf, 0xf, false),
                __builtin_amdgcn_update_dpp(0u, b, 0x128, 0xf, 0xf, false));
  } else if constexpr (M == 4) {
    const uint32_t dl = __builtin_amdgcn_update_dpp(0u, a, 0x114, 0xf, 0xf, false);
    const uint32_t dh = __builtin_amdgcn_update_dpp(0u, b, 0x114, 0xf, 0xf, false);
    const uint32_t ul = __builtin_amdgcn_update_dpp(0u, a, 0x104, 0xf, 0xf, false);
    const uint32_t uh = __builtin_amdgcn_update_dpp(0u, b, 0x104, 0xf, 0xf, false);
    return (lane & 4) ? mk64(dl, dh) : mk64(ul, uh);
  } else {
    constexpr int QP = M == 2 ? 0x4E : 0xB1;   // quad_perm [2,3,0,1] / [1,0,3,2]
    return mk64(__builtin_amdgcn_update_dpp(0u, a, QP, 0xf, 0xf, false),
                __builtin_amdgcn_update_dpp(0u, b, QP, 0xf, 0xf, false));
  }
}

// One exchange for M = 8 or 4 over the pairs (acc[i], acc[i + N/2]).
template <int N, int M>
__device__ __forceinline__ void halve(double (&acc)[16], int lane)
{
  const bool hi = (lane & M) != 0;
#pragma unroll
  for (int i = 0; i < N / 2; i++) {
    const double send = hi ? acc[i] : acc[i + N / 2];
    const double keep = hi ? acc[i + N / 2] : acc[i];
    acc[i] = keep + xor_dpp<M>(send, lane);
  }
}

// The two halves of the column sums: half[h][c] for the rows of wave
// parity h; all threads past a barrier.
__device__ __forceinline__ void column_halves(double (&acc)[16], double (*half)[SB])
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  halve_swap<16, true>(acc);
  halve_swap<8, false>(acc);
  halve<4, 8>(acc, lane);
  halve<2, 4>(acc, lane);
  acc[0] += xor_dpp<2>(acc[0], lane);
  acc[0] += xor_dpp<1>(acc[0], lane);
  // lane l now holds column k(l) = 8 b5 + 4 b4 + 2 b3 + b2 of its slice
  if ((lane & 3) == 0) {
    const int k = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                  ((lane >> 2) & 1);
    half[wave & 1][(wave >> 1) * 16 + k] = acc[0];
  }
  __syncthreads();
}

// out[c] = the sum over the 128 rows for column c (all threads past a barrier).
__device__ __forceinline__ void column_sums(double (&acc)[16], double (*half)[SB], double* out)
{
  column_halves(acc, half);
  if (threadIdx.x < SB) out[threadIdx.x] = half[0][threadIdx.x] + half[1][threadIdx.x];
  __syncthreads();
}

// Top of the tree (1024 threads).  Loads run along rows (lanes = rows of L,
// coalesced); a thread keeps its row's terms for 16 columns across all the
// blocks it consumes and the column sums are formed once at the end.
//   chunk item (u, b): rows [nc + 128u, ...) of s against column block b:
//     bpart = L(rows, b-block)^T x(rows)
//   pivot item: columns [128 blk, ...): x = Dinv^T (y - sum of its chunks'
//     partials - L(later blocks, blk-block)^T x(later blocks))
__device__ __forceinline__ void solve_bwd_top(const CholDev& c, const SolveDev& sd,
                                              const SysBufs& B, int s, int kind, int idx,
                                              double* wl)
{
  constexpr int QW = 16;
  __shared__ double xs[2][SB];
  __shared__ double half[2][SB];
  __shared__ double cs[SB];
  __shared__ double ts[SB];
  __shared__ double part[8][SB];
  const double* __restrict__ y = B.in;
  double* __restrict__ x = B.out;
  const SnInfo si = sninfo(c, B.L, s);
  const int r = threadIdx.x & (SB - 1);
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);
  const uint32_t f8 = static_cast<uint32_t>(si.f) * 8u;
  double acc[QW];
#pragma unroll
  for (int k = 0; k < QW; k++) acc[k] = 0.0;
  // row r of the block of rows starting at r0 (nrows valid) and columns
  // starting at c0 (ncols valid): L(r0 + r, c0 + 16 q + k)
  auto row_slice = [&](double (&l)[QW], int r0, int nrows, int c0, int ncols) {
    const rsrc_t lv = view(si.Ls + static_cast<int64_t>(c0) * si.f,
                           static_cast<uint32_t>(ncols) * f8);
    slice_load<QW>(l, lv, static_cast<uint32_t>(r0 + r) * 8u, static_cast<uint32_t>(q * QW) * f8,
                   f8, ncols - q * QW, r < nrows);
  };

  if (kind == 0) {
    const int u = idx / si.np, b = idx - u * si.np;
    const int rr0 = u * SB;
    const int nrows = min(SB, si.nr - rr0);
    const int bb = min(SB, si.nc - b * SB);
    double l[QW];
    row_slice(l, si.nc + rr0, nrows, b * SB, bb);
    // x of these rows belongs to ancestors: waited for as it is read
    if (threadIdx.x < SB) {
      const int* rows = c.rows + c.rows_ptr[s] + rr0;
      xs[0][threadIdx.x] = (threadIdx.x < nrows) ? spin_load(x + rows[threadIdx.x], sd.err) : 0.0;
    }
    __syncthreads();
    const double xr = xs[0][r];
#pragma unroll
    for (int k = 0; k < QW; k++) acc[k] = l[k] * xr;
    column_sums(acc, half, cs);
    double* out = B.bpart + sd.bpart_off[s] + static_cast<int64_t>(u) * si.nc + b * SB;
    if (threadIdx.x < bb) xstore(out + threadIdx.x, cs[threadIdx.x]);
    return;
  }

  const int blk = idx;
  const int c0 = blk * SB;
  const int nb = min(SB, si.nc - c0);
  const int jtop = si.np - 1;   // later pivot blocks j = jtop .. blk+1
  // x = Dinv^T t: x(i) = sum_j Dinv(j, i) t(j), thread (i = r, slice q of j)
  const rsrc_t wv = view(B.dinv + uni64(c.dinv_off[s]) + static_cast<int64_t>(blk) * SB * SB,
                         SB * SB * 8u);
  // The inverse goes to LDS up front (wl[j * 128 + i] = Dinv(j, i), read
  // along Dinv's column i), so the last link reads no global memory before
  // its result.
  {
    double wt[QW];
    slice_load<QW>(wt, wv, static_cast<uint32_t>(r) * SB * 8u, static_cast<uint32_t>(q * QW) * 8u,
                   8u, nb - q * QW, r < nb);
#pragma unroll
    for (int k = 0; k < QW; k++) wl[(q * QW + k) * SB + r] = wt[k];
  }
  auto rows_of = [&](int j) { return min(SB, si.nc - j * SB); };
  double l[QW];
  if (jtop > blk) row_slice(l, jtop * SB, rows_of(jtop), c0, nb);

  if (threadIdx.x < SB) {
    double t = 0.0;
    if (threadIdx.x < nb) {
      t = y[si.a + c0 + threadIdx.x];
      const double* bp = B.bpart + sd.bpart_off[s] + c0 + threadIdx.x;
      for (int uu = 0; uu < si.nu; uu++) t -= spin_load(bp + static_cast<int64_t>(uu) * si.nc, sd.err);
    }
    ts[threadIdx.x] = t;
  }
  for (int j = jtop; j > blk; j--) {
    double* xb = xs[j & 1];
    if (threadIdx.x < SB)
      xb[threadIdx.x] = (threadIdx.x < rows_of(j)) ? spin_load(x + si.a + j * SB + threadIdx.x, sd.err) : 0.0;
    __syncthreads();   // xs[j & 1] is written again two blocks on, after the next barrier
    const double xr = xb[r];
#pragma unroll
    for (int k = 0; k < QW; k++) acc[k] += l[k] * xr;
    if (j - 1 > blk) row_slice(l, (j - 1) * SB, rows_of(j - 1), c0, nb);
  }
  // x = Dinv^T (t - column sums), the subtraction formed where it is used
  column_halves(acc, half);
  {
    double res = 0.0;
#pragma unroll
    for (int k = 0; k < QW; k++) {
      const int cc = q * QW + k;
      res += wl[cc * SB + r] * (ts[cc] - (half[0][cc] + half[1][cc]));
    }
    part[q][r] = res;
  }
  __syncthreads();
  if (threadIdx.x < nb) xstore(x + si.a + c0 + threadIdx.x, sum_slices<8>(part, threadIdx.x));
}

template <int NT, bool PF>
__device__ __forceinline__ void bwd_item(const CholDev& c, const SolveDev& sd,
                                         const int* __restrict__ items,
                                         const double* __restrict__ yin,
                                         double* __restrict__ xout, const Batch& bt, int itc,
                                         double* xs, double (*part)[SB], double* wl)
{
  constexpr int NQ = NT / SB, QW = SB / NQ;
  const int sy = itc % bt.count, it = itc / bt.count;
  if (bt.mask && !bt.mask[sy]) return;
  const SysBufs B = sys_bufs(c, sd, sy, bt, yin, xout);
  const int s = uni(items[3 * it]);
  const int kind = uni(items[3 * it + 1]);
  const int idx = uni(items[3 * it + 2]);
  if constexpr (PF) {
    solve_bwd_top(c, sd, B, s, kind, idx, wl);
    return;
  }
  const double* __restrict__ y = B.in;
  double* __restrict__ x = B.out;
  const SnInfo si = sninfo(c, B.L, s);
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
    double* out = B.bpart + sd.bpart_off[s] + static_cast<int64_t>(u) * si.nc + b * SB;
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
  const rsrc_t wv = view(B.dinv + uni64(c.dinv_off[s]) + static_cast<int64_t>(blk) * SB * SB,
                         SB * SB * 8u);
  const uint32_t wvo = static_cast<uint32_t>(col) * SB * 8u, wso = static_cast<uint32_t>(q * QW) * 8u;
  const int jtop = si.np - 1;   // later pivot blocks j = jtop .. blk+1
  // both slices up front (a load in flight delays the wave's polls)
  double l[PF ? QW : 1], w[PF ? QW : 1];
  if constexpr (PF) {
    if (jtop > blk) slice_load<QW>(l, lv, cvo, lso(jtop), 8u, lvalid(jtop), live);
    slice_load<QW>(w, wv, wvo, wso, 8u, nb - q * QW, live);
  }

  double t = 0.0;
  if (kind == 2) {
    // a one-block supernode whose item forms its row chunks' products
    // L(rows, :)^T x(rows) itself, chunk by chunk in the order (and with the
    // arithmetic) of the chunk items it replaces
    if (q == 0 && live) t = y[si.a + col];
    for (int u = 0; u < si.nu; u++) {
      const int rr0 = u * SB;
      const int nrows = min(SB, si.nr - rr0);
      const int* rows = c.rows + c.rows_ptr[s] + rr0;
      __syncthreads();   // the previous chunk's reads of xs and part are done
      if (threadIdx.x < SB)
        xs[threadIdx.x] = (threadIdx.x < nrows) ? spin_load(x + rows[threadIdx.x], sd.err) : 0.0;
      __syncthreads();
      const uint32_t so = static_cast<uint32_t>(si.nc + rr0 + q * QW) * 8u;
      part[q][col] = slice_dot_stream<QW>(lv, cvo, so, 8u, nrows - q * QW, live, xs + q * QW);
      __syncthreads();
      if (q == 0) t -= sum_slices<NQ>(part, col);
    }
    __syncthreads();   // every read of part and xs before they are written below
  } else if (q == 0 && live) {
    t = y[si.a + c0 + col];
    const double* bp = B.bpart + sd.bpart_off[s] + c0 + col;
    for (int u = 0; u < si.nu; u++) t -= spin_load(bp + static_cast<int64_t>(u) * si.nc, sd.err);
  }
  for (int j = jtop; j > blk; j--) {
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

template <int NT, bool PF, bool LOOP>
__global__ void __launch_bounds__(NT) k_solve_bwd(CholDev c, SolveDev sd,
                                                  const int* __restrict__ items,
                                                  const double* __restrict__ yin,
                                                  double* __restrict__ xout, Batch bt, int total)
{
  constexpr int NQ = NT / SB;
  __shared__ double xs[SB];
  __shared__ double part[NQ][SB];
  extern __shared__ double wl[];   // 128 x 128 inverse block (dynamic LDS; top only)
  if constexpr (!LOOP) {
    bwd_item<NT, PF>(c, sd, items, yin, xout, bt, take_item(sd.ticket + (PF ? 2 : 3)), xs, part,
                     wl);
    return;
  }
  for (;;) {
    const int itc = take_item(sd.ticket + (PF ? 2 : 3));
    if (itc >= total) break;
    bwd_item<NT, PF>(c, sd, items, yin, xout, bt, itc, xs, part, wl);
    __syncthreads();   // the item's last reads of the shared arrays
  }
}

}  // namespace

void fill_sentinel(double* p, int64_t n, const Batch& b, hipStream_t st, int* zero4)
{
  fill_sentinel3(p, n, nullptr, 0, nullptr, 0, b, st, zero4);
}

void fill_sentinel3(double* p0, int64_t n0, double* p1, int64_t n1, double* p2, int64_t n2,
                    const Batch& b, hipStream_t st, int* zero4)
{
  n0 = std::max<int64_t>(n0, 0);
  n1 = std::max<int64_t>(n1, 0);
  n2 = std::max<int64_t>(n2, 0);
  const int64_t n = n0 + n1 + n2;
  if ((n <= 0 && !zero4) || b.count <= 0) return;
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  k_fill_sentinel<<<dim3(static_cast<unsigned>(g), b.count), 256, 0, st>>>(p0, n0, p1, n1, p2,
                                                                             n2, zero4, b);
  HIPCHK(hipGetLastError());
}

// The top backward sweep stages one 128 x 128 inverse block in dynamic LDS.
constexpr size_t WL = static_cast<size_t>(SB) * SB * sizeof(double);

// Raise kernel K's dynamic LDS limit to WL, once per process and kernel (a
// function-local static initialises thread-safely).
template <auto K>
static void allow_top_lds()
{
  static const bool done = [] {
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(K),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(WL)));
    return true;
  }();
  (void)done;
}

// Workgroups of a persistent sweep launch: one per item, at most `cap` (0:
// no bound).
static unsigned sweep_grid(int items, int cap)
{
  const int g = cap > 0 ? std::min(items, cap) : items;
  return static_cast<unsigned>(std::max(g, 1));
}

int forest_blocks_per_cu()
{
  int nb = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fwd_forest, NT_FOREST, 0));
  return std::max(nb, 1);
}

void solve_forward(const CholDev& c, const SolveDev& sd, const double* b, double* y, int part,
                   const Batch& bt, hipStream_t st)
{
  if (bt.count <= 0) return;
  if (part == 0 && sd.n_fwd > 0)
    k_fwd_forest<<<std::min(sd.n_fwd * bt.count, sd.grid_fwd), NT_FOREST, 0, st>>>(
        c, sd, sd.fwd_items, b, y, bt);
  if (part == 1 && sd.n_fwd_top > 0) {
    const int total = sd.n_fwd_top * bt.count;
    if (sd.cap_top > 0)
      k_solve_fwd<NT_TOP, true, true><<<sweep_grid(total, sd.cap_top), NT_TOP, 0, st>>>(
          c, sd, sd.fwd_items_top, b, y, bt, total);
    else
      k_solve_fwd<NT_TOP, true, false><<<total, NT_TOP, 0, st>>>(c, sd, sd.fwd_items_top, b, y,
                                                                 bt, total);
  }
  HIPCHK(hipGetLastError());
}


void solve_backward(const CholDev& c, const SolveDev& sd, const double* y, double* x, int part,
                    const Batch& bt, hipStream_t st)
{
  if (bt.count <= 0) return;
  if (part == 1 && sd.n_bwd_top > 0) {
    const int total = sd.n_bwd_top * bt.count;
    if (sd.cap_top > 0) {
      allow_top_lds<k_solve_bwd<NT_TOP, true, true>>();
      k_solve_bwd<NT_TOP, true, true><<<sweep_grid(total, sd.cap_top), NT_TOP, WL, st>>>(
          c, sd, sd.bwd_items_top, y, x, bt, total);
    } else {
      allow_top_lds<k_solve_bwd<NT_TOP, true, false>>();
      k_solve_bwd<NT_TOP, true, false><<<total, NT_TOP, WL, st>>>(c, sd, sd.bwd_items_top, y, x,
                                                                  bt, total);
    }
  }
  if (part == 0 && sd.n_bwd > 0) {
    const int total = sd.n_bwd * bt.count;
    if (sd.cap_forest > 0)
      k_solve_bwd<NT_FOREST, false, true><<<sweep_grid(total, sd.cap_forest), NT_FOREST, 0,
                                            st>>>(c, sd, sd.bwd_items, y, x, bt, total);
    else
      k_solve_bwd<NT_FOREST, false, false><<<total, NT_FOREST, 0, st>>>(c, sd, sd.bwd_items, y, x,
                                                                        bt, total);
  }
  HIPCHK(hipGetLastError());
}

}  // namespace dev
}  // namespace hykkt
