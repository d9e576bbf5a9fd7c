// kernels_basic.hip -- gfx950 kernels for the formation, Ruiz, SpMV, CG and
// recovery steps of the HyKKT solve, over a batch of same-pattern systems
// (system blockIdx.y; see Batch in kernels.hpp).
//
// Built with -ffp-contract=off: every fused multiply-add here is an explicit
// fma(), so the SpGEMM/Ruiz/rhs values are bit-identical to the CPU oracle,
// which uses the same operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <string>

#include "common.hpp"
#include "kernels.hpp"

namespace hykkt {
namespace dev {

void check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    int code = (e == hipErrorOutOfMemory) ? HYKKT_E_NOMEM : HYKKT_E_HIP;
    throw Error(code, std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
  }
}

namespace {

constexpr int TPB = 256;
constexpr int RED_MAX = 256;    // most partials of one reduction (per system)
constexpr int SPMV_E = 1024;    // entries of a row block staged in LDS

inline unsigned nblocks(int64_t n, int b = TPB)
{
  int64_t g = (n + b - 1) / b;
  if (g > 65536) g = 65536;
  return static_cast<unsigned>(g > 0 ? g : 1);
}

inline dim3 grid2(int64_t n, const Batch& b) { return dim3(nblocks(n), b.count); }

// This workgroup's system, or -1 when the mask skips it.
__device__ __forceinline__ int sys_of(const Batch& b)
{
  const int s = blockIdx.y;
  return (b.mask && !b.mask[s]) ? -1 : s;
}
template <class T>
__device__ __forceinline__ T* at(T* p, int s, const Batch& b)
{
  return p ? p + static_cast<int64_t>(s) * b.vs : p;
}

// Deterministic tree sum of one value per thread (TPB threads); all threads
// get the result.
__device__ double block_sum(double v)
{
  __shared__ double sh[TPB];
  __syncthreads();   // sh may still be read by a previous call
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int k = TPB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  return sh[0];
}

// ------------------------------------------------------------- vectors
__global__ void k_fill(double* __restrict__ p, double v, int64_t n, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  p = at(p, s, b);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i(int* __restrict__ p, int v, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_copy(double* __restrict__ d, const double* __restrict__ src, int64_t n,
                       const int* __restrict__ keep, Batch b)
{
  const int s = sys_of(b);
  if (s < 0 || (keep && !keep[s])) return;
  d = at(d, s, b);
  src = at(src, s, b);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = src[i];
}

struct CopyDesc {
  int nsrc;
  double* dst[12];
  int64_t len[12];
};

// grid (blocks, systems, arrays): array j of system s from the caller's
// pointer tab[s * nsrc + j] into its arena slot
__global__ void k_copy_in(CopyDesc c, const double* const* __restrict__ tab, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  const int j = blockIdx.z;
  const double* src = tab[static_cast<int64_t>(s) * c.nsrc + j];
  double* d = at(c.dst[j], s, b);
  const int64_t n = c.len[j];
  if (!src) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = src[i];
}

__global__ void k_collect(double* __restrict__ out, const double* __restrict__ base, int nslot,
                          Batch b)
{
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(nslot) * b.count) return;
  const int64_t k = t / nslot, j = t - k * nslot;
  out[t] = base[k * b.vs + j];
}

// Per system k: the CG's ||r||^2 at the start and at exit (sc slots 0, 1),
// its iterations and status, as four doubles for one readback.
__global__ void k_cg_pack(double* __restrict__ out, const double* __restrict__ sc,
                          const int* __restrict__ iters, const int* __restrict__ status, Batch b)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.count) return;
  const double* c = sc + static_cast<int64_t>(k) * b.vs;
  out[4 * k + 0] = c[0];
  out[4 * k + 1] = c[1];
  out[4 * k + 2] = iters[k];
  out[4 * k + 3] = status[k];
}

// out[0] = *w, then a[0..count) and b[0..count) where given.
__global__ void k_collect_ints(int* __restrict__ out, const int* __restrict__ w,
                               const int* __restrict__ a, const int* __restrict__ bv, int count)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) out[0] = *w;
  if (t < count) {
    int at = 1;
    if (a) {
      out[at + t] = a[t];
      at += count;
    }
    if (bv) out[at + t] = bv[t];
  }
}

__global__ void k_gather(double* __restrict__ d, const double* __restrict__ src,
                         const int* __restrict__ map, int64_t n, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  d = at(d, s, b);
  src = at(src, s, b);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = src[map[i]];
}

__global__ void k_gather64(double* __restrict__ d, const double* __restrict__ s,
                           const int64_t* __restrict__ map, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[map[i]];
}

__global__ void k_gather_scale(double* __restrict__ d, const double* __restrict__ src,
                               const int* __restrict__ map, const double* __restrict__ sc, int n,
                               Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  d = at(d, s, b);
  src = at(src, s, b);
  sc = at(sc, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = src[map[i]] * sc[i];
}

// The recovery's scaled results in one launch: dx[i] = (h0 - cz)[map[i]] *
// max_d[i] (i < n) and dyc[i] = cx[i] * max_d[n + i] (i < m), each the
// rounding of its own sub / gather_scale / vec_scale.
__global__ void k_recover_xy(double* __restrict__ dx, const double* __restrict__ h0,
                             const double* __restrict__ cz, const int* __restrict__ map,
                             const double* __restrict__ max_d, int n, double* __restrict__ dyc,
                             const double* __restrict__ cx, int m, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  dx = at(dx, s, b);
  h0 = at(h0, s, b);
  cz = at(cz, s, b);
  max_d = at(max_d, s, b);
  dyc = at(dyc, s, b);
  cx = at(cx, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n || i < m; i += gridDim.x * blockDim.x) {
    if (i < n) {
      const int j = map[i];
      const double t = h0[j] - cz[j];
      dx[i] = t * max_d[i];
    }
    if (i < m) dyc[i] = cx[i] * max_d[n + i];
  }
}

__global__ void k_vec_scale(double* __restrict__ y, const double* __restrict__ sv, int n, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  y = at(y, s, b);
  sv = at(sv, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] *= sv[i];
}

__global__ void k_sub(double* __restrict__ out, const double* __restrict__ a,
                      const double* __restrict__ bv, int64_t n, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  out = at(out, s, b);
  a = at(a, s, b);
  bv = at(bv, s, b);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - bv[i];
}

__global__ void k_add(double* __restrict__ out, const double* __restrict__ a,
                      const double* __restrict__ bv, int64_t n, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  out = at(out, s, b);
  a = at(a, s, b);
  bv = at(bv, s, b);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] + bv[i];
}

// ------------------------------------------------------------- sparse
__global__ void k_row_scale(int d, const int* __restrict__ rp, const double* __restrict__ v,
                            const double* __restrict__ ds, const double* __restrict__ ryd,
                            const double* __restrict__ rs, double* __restrict__ vs,
                            double* __restrict__ ryd_s, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  v = at(v, s, b);
  ds = at(ds, s, b);
  ryd = at(ryd, s, b);
  rs = at(rs, s, b);
  vs = at(vs, s, b);
  ryd_s = at(ryd_s, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d; i += gridDim.x * blockDim.x) {
    double di = ds[i];
    if (vs)
      for (int p = rp[i]; p < rp[i + 1]; p++) vs[p] = v[p] * di;
    double t = ryd[i] * di;
    ryd_s[i] = t + 1.0 * rs[i];
  }
}

// CSR-stream SpMV (Greathouse & Daga, SC'14): one workgroup per row block;
// the block's values and the x entries they multiply are staged through LDS
// with coalesced loads, then each row is summed by one thread in column
// order, so the result equals the sequential fma chain bit for bit.
__global__ void __launch_bounds__(TPB) k_spmv(const int* __restrict__ rp,
                                              const int* __restrict__ ci,
                                              const int* __restrict__ map,
                                              const double* __restrict__ v, double alpha,
                                              const double* __restrict__ alpha_ps,
                                              const double* __restrict__ x, double beta,
                                              const double* __restrict__ y_in,
                                              double* __restrict__ y_out,
                                              const int* __restrict__ blocks,
                                              const double* __restrict__ dot_w,
                                              double* __restrict__ dot_part, Batch b)
{
  __shared__ double lv[SPMV_E];
  __shared__ double lx[SPMV_E];
  const int s = sys_of(b);
  if (s < 0) return;
  v = at(v, s, b);
  x = at(x, s, b);
  y_in = at(y_in, s, b);
  y_out = at(y_out, s, b);
  if (alpha_ps) alpha = *at(alpha_ps, s, b);
  const int r0 = blocks[blockIdx.x], r1 = blocks[blockIdx.x + 1];
  const int e0 = rp[r0], e1 = rp[r1];
  const int tid = threadIdx.x;
  double yi = 0.0;
  const int r = r0 + tid;
  if (e1 - e0 <= SPMV_E) {
    for (int k = tid; k < e1 - e0; k += TPB) {
      const int q = e0 + k;
      lv[k] = v[map ? map[q] : q];
      lx[k] = x[ci[q]];
    }
    __syncthreads();
    if (r < r1) {
      double t = 0.0;
      for (int q = rp[r] - e0; q < rp[r + 1] - e0; q++) t = fma(lv[q], lx[q], t);
      yi = (beta == 0.0) ? alpha * t : fma(alpha, t, beta * y_in[r]);
      y_out[r] = yi;
    }
  } else {
    // one long row (the block has only this row): staged in chunks, summed
    // by thread 0 in order
    double t = 0.0;
    for (int c0 = e0; c0 < e1; c0 += SPMV_E) {
      const int cn = min(SPMV_E, e1 - c0);
      __syncthreads();
      for (int k = tid; k < cn; k += TPB) {
        const int q = c0 + k;
        lv[k] = v[map ? map[q] : q];
        lx[k] = x[ci[q]];
      }
      __syncthreads();
      if (tid == 0)
        for (int k = 0; k < cn; k++) t = fma(lv[k], lx[k], t);
    }
    if (tid == 0) {
      yi = (beta == 0.0) ? alpha * t : fma(alpha, t, beta * y_in[r0]);
      y_out[r0] = yi;
    }
  }
  if (dot_part) {
    const double w = (r < r1) ? yi * at(dot_w, s, b)[r] : 0.0;
    const double tot = block_sum(w);
    if (tid == 0) at(dot_part, s, b)[blockIdx.x] = tot;
  }
}

__device__ __forceinline__ double spgemm_entry(int p, const int* __restrict__ d_src,
                                               const double* __restrict__ d_v,
                                               const int* __restrict__ tp,
                                               const int* __restrict__ ta,
                                               const int* __restrict__ tb,
                                               const double* __restrict__ av,
                                               const double* __restrict__ bv, double alpha_p)
{
  double acc = 0.0;
  for (int t = tp[p]; t < tp[p + 1]; t++) acc = fma(av[ta[t]], bv[tb[t]], acc);
  int s = d_src[p];
  double d = (s >= 0) ? d_v[s] : 0.0;
  double c = alpha_p * acc;
  return 1.0 * d + 1.0 * c;
}

__global__ void k_spgemm_values(int nnz, const int* __restrict__ d_src,
                                const double* __restrict__ d_v, const int* __restrict__ tp,
                                const int* __restrict__ ta, const int* __restrict__ tb,
                                const double* __restrict__ av, const double* __restrict__ bv,
                                double alpha_p, double* __restrict__ out, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  d_v = at(d_v, s, b);
  av = at(av, s, b);
  bv = at(bv, s, b);
  out = at(out, s, b);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += gridDim.x * blockDim.x)
    out[p] = spgemm_entry(p, d_src, d_v, tp, ta, tb, av, bv, alpha_p);
}

__global__ void k_spgemm_scatter(int nent, const int* __restrict__ src,
                                 const int64_t* __restrict__ dst,
                                 const int* __restrict__ d_src, const double* __restrict__ d_v,
                                 const int* __restrict__ tp, const int* __restrict__ ta,
                                 const int* __restrict__ tb, const double* __restrict__ av,
                                 const double* __restrict__ bv,
                                 const double* __restrict__ alpha_ps, double* __restrict__ out,
                                 Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  d_v = at(d_v, s, b);
  av = at(av, s, b);
  bv = at(bv, s, b);
  out = at(out, s, b);
  const double alpha_p = *at(alpha_ps, s, b);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nent; e += gridDim.x * blockDim.x)
    out[dst[e]] = spgemm_entry(src[e], d_src, d_v, tp, ta, tb, av, bv, alpha_p);
}

__global__ void k_scatter(double* __restrict__ out, const int64_t* __restrict__ dst,
                          const double* __restrict__ vals, const int* __restrict__ src, int n,
                          Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  out = at(out, s, b);
  vals = at(vals, s, b);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    out[dst[e]] = vals[src[e]];
}

__global__ void k_add_at(double* __restrict__ out, const int64_t* __restrict__ pos, int n,
                         double delta, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  out = at(out, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[pos[i]] += delta;
}

// ------------------------------------------------------------- Ruiz
__global__ void k_ruiz_rowmax(int n, int totn, const int* __restrict__ h_rp,
                              const double* __restrict__ h_v, const int* __restrict__ jt_rp,
                              const int* __restrict__ jt_map, const double* __restrict__ jt_v,
                              const int* __restrict__ j_rp, const double* __restrict__ j_v,
                              double* __restrict__ scale, int* __restrict__ zero_row, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  h_v = at(h_v, s, b);
  jt_v = at(jt_v, s, b);
  j_v = at(j_v, s, b);
  scale = at(scale, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < totn; i += gridDim.x * blockDim.x) {
    double max_l = 0.0, max_u = 0.0;
    if (i < n) {
      for (int p = h_rp[i]; p < h_rp[i + 1]; p++) {
        double e = fabs(h_v[p]);
        if (e > max_l) max_l = e;
      }
      for (int q = jt_rp[i]; q < jt_rp[i + 1]; q++) {
        double e = fabs(jt_v[jt_map ? jt_map[q] : q]);
        if (e > max_u) max_u = e;
      }
      scale[i] = (max_l > max_u) ? 1.0 / sqrt(max_l) : 1.0 / sqrt(max_u);
    } else {
      for (int p = j_rp[i - n]; p < j_rp[i - n + 1]; p++) {
        double e = fabs(j_v[p]);
        if (e > max_l) max_l = e;
      }
      scale[i] = 1.0 / sqrt(max_l);
    }
    // a row without a nonzero value (or with a NaN) gets the scale 1/sqrt(0)
    if (zero_row && !(max_l > 0.0 || max_u > 0.0)) atomicMin(zero_row + s, i);
  }
}

__global__ void k_ruiz_scale(int n, int totn, const int* __restrict__ h_rp,
                             const int* __restrict__ h_ci, double* __restrict__ h_v,
                             const int* __restrict__ jt_rp, const int* __restrict__ jt_ci,
                             double* __restrict__ jt_v, const int* __restrict__ j_rp,
                             const int* __restrict__ j_ci, double* __restrict__ j_v,
                             const double* __restrict__ scale, double* __restrict__ rhs1,
                             double* __restrict__ rhs2, double* __restrict__ max_d, int first,
                             Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  h_v = at(h_v, s, b);
  jt_v = at(jt_v, s, b);
  j_v = at(j_v, s, b);
  scale = at(scale, s, b);
  rhs1 = at(rhs1, s, b);
  rhs2 = at(rhs2, s, b);
  max_d = at(max_d, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < totn; i += gridDim.x * blockDim.x) {
    double si = scale[i];
    if (i < n) {
      for (int p = h_rp[i]; p < h_rp[i + 1]; p++) h_v[p] *= si * scale[h_ci[p]];
      rhs1[i] *= si;
      max_d[i] = first ? 1.0 * si : max_d[i] * si;
      if (jt_v)
        for (int q = jt_rp[i]; q < jt_rp[i + 1]; q++) jt_v[q] *= si * scale[n + jt_ci[q]];
    } else {
      for (int p = j_rp[i - n]; p < j_rp[i - n + 1]; p++) j_v[p] *= si * scale[j_ci[p]];
      rhs2[i - n] *= si;
      max_d[i] = first ? 1.0 * si : max_d[i] * si;
    }
  }
}

// ------------------------------------------------------------- reductions
// grid (reduce_blocks(n), systems): block j sums i = j*TPB + t + k*(blocks*TPB)
__global__ void k_dot_partial(int n, const double* __restrict__ a, const double* __restrict__ bv,
                              double* __restrict__ part, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  a = at(a, s, b);
  bv = at(bv, s, b);
  double t = 0.0;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    t = bv ? fma(a[i], bv[i], t) : t + a[i];
  const double tot = block_sum(t);
  if (threadIdx.x == 0) at(part, s, b)[blockIdx.x] = tot;
}

// Several sums in one launch: block j of sum k (off[k] <= j < off[k+1])
// takes the block j - off[k] of the g = off[k+1] - off[k] blocks that
// k_dot_partial would launch for it, in the same order, so the partials are
// the same bits.
__global__ void k_dots_partial(DotSet d, double* __restrict__ part, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  int k = 0;
  while (k + 1 < d.count && d.off[k + 1] <= static_cast<int>(blockIdx.x)) k++;
  const int jb = blockIdx.x - d.off[k], g = d.off[k + 1] - d.off[k];
  const double* a = at(d.a[k], s, b);
  const double* bv = d.b[k] ? at(d.b[k], s, b) : nullptr;
  const int n = d.n[k];
  double t = 0.0;
  for (int i = jb * TPB + threadIdx.x; i < n; i += g * TPB)
    t = bv ? fma(a[i], bv[i], t) : t + a[i];
  const double tot = block_sum(t);
  if (threadIdx.x == 0) at(part, s, b)[blockIdx.x] = tot;
}

// one workgroup per system: each sum's partials added as k_dot_final adds them
__global__ void k_dots_final(DotSet d, const double* __restrict__ part, double* __restrict__ out,
                             Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  part = at(part, s, b);
  out = at(out, s, b);
  for (int k = 0; k < d.count; k++) {
    const int np = d.off[k + 1] - d.off[k];
    double t = 0.0;
    for (int i = threadIdx.x; i < np; i += TPB) t += part[d.off[k] + i];
    const double tot = block_sum(t);
    if (threadIdx.x == 0) out[k] = tot;
  }
}

// one workgroup per system: the sum of its partials, then the CG step's
// scalar work (see dot_final in kernels.hpp)
__global__ void k_dot_final(int np, const double* __restrict__ part, double* __restrict__ sc,
                            int slot, int check, int scalars, int it, double tol, int itmax,
                            CgFlags f, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  part = at(part, s, b);
  double t = 0.0;
  for (int i = threadIdx.x; i < np; i += TPB) t += part[i];
  const double tot = block_sum(t);
  if (threadIdx.x != 0) return;
  double* c = at(sc, s, b);
  c[slot] = tot;
  if (check) {
    if (it < 0) {
      if (tot == 0.0) {   // x0 already solves the system (the reference divides 0/0 here)
        f.live[s] = 0;
        f.iters[s] = 0;
        f.status[s] = 0;
      }
    } else if (sqrt(tot) < tol) {
      f.live[s] = 0;
      f.iters[s] = it + 1;
      f.status[s] = 0;
    } else if (it + 1 >= itmax) {
      f.live[s] = 0;
      f.iters[s] = itmax;
      f.status[s] = 1;
    }
  }
  if (scalars) {
    if (it < 0) {
      // alpha = gam / delta (SchurComplementConjugateGradient.cpp:143), beta = 0
      c[3] = c[0] / c[2];
      c[4] = 0.0;
    } else {
      // beta = gam1/gam; gam = gam1; alpha = gam/(delta - beta*gam/alpha) (:181-183)
      const double beta = c[1] / c[0];
      const double gam = c[1];
      c[4] = beta;
      c[0] = gam;
      c[3] = gam / (c[2] - beta * gam / c[3]);
    }
  }
}

// ------------------------------------------------------------- CG
__global__ void k_cg_update_dot(int m, const double* __restrict__ sc, double* __restrict__ p,
                                double* __restrict__ sv, double* __restrict__ x,
                                double* __restrict__ r, const double* __restrict__ w,
                                double* __restrict__ part, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  const double* c = at(sc, s, b);
  p = at(p, s, b);
  sv = at(sv, s, b);
  x = at(x, s, b);
  r = at(r, s, b);
  w = at(w, s, b);
  const double alpha = c[3], beta = c[4];
  double t = 0.0;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < m; i += gridDim.x * TPB) {
    const double pi = beta * p[i] + r[i];
    const double si = beta * sv[i] + w[i];
    p[i] = pi;
    sv[i] = si;
    x[i] = fma(alpha, pi, x[i]);
    const double ri = fma(-alpha, si, r[i]);
    r[i] = ri;
    t = fma(ri, ri, t);
  }
  const double tot = block_sum(t);
  if (threadIdx.x == 0) at(part, s, b)[blockIdx.x] = tot;
}

// The CG's image of its iterate under A^{-1} Jh^T, carried along with it:
// q = beta q + t (t = A^{-1} Jh^T r, formed by the last Schur product),
// z += alpha q, the same coefficients in the same forms as p and x, so z =
// A^{-1} Jh^T x without a solve.
__global__ void k_cg_track(int n, const double* __restrict__ sc, double* __restrict__ q,
                           double* __restrict__ z, const double* __restrict__ t, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  const double* c = at(sc, s, b);
  q = at(q, s, b);
  z = at(z, s, b);
  t = at(t, s, b);
  const double alpha = c[3], beta = c[4];
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
    const double qi = beta * q[i] + t[i];
    q[i] = qi;
    z[i] = fma(alpha, qi, z[i]);
  }
}

// ------------------------------------------------------------- recover / residual
__global__ void k_resid_n(int n, const int* __restrict__ h_rp, const int* __restrict__ h_ci,
                          const double* __restrict__ h_v, const int* __restrict__ ct_rp,
                          const int* __restrict__ ct_ci, const int* __restrict__ ct_map,
                          const double* __restrict__ jc_v, const int* __restrict__ dt_rp,
                          const int* __restrict__ dt_ci, const int* __restrict__ dt_map,
                          const double* __restrict__ jd_v, const double* __restrict__ rx,
                          const double* __restrict__ dx, const double* __restrict__ dyc,
                          const double* __restrict__ dyd, double* __restrict__ sq,
                          double* __restrict__ e, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  h_v = at(h_v, s, b);
  jc_v = at(jc_v, s, b);
  jd_v = at(jd_v, s, b);
  rx = at(rx, s, b);
  dx = at(dx, s, b);
  dyc = at(dyc, s, b);
  dyd = at(dyd, s, b);
  sq = at(sq, s, b);
  e = at(e, s, b);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double t = rx[i];
    for (int p = h_rp[i]; p < h_rp[i + 1]; p++) t = fma(-h_v[p], dx[h_ci[p]], t);
    for (int q = ct_rp[i]; q < ct_rp[i + 1]; q++) t = fma(-jc_v[ct_map[q]], dyc[ct_ci[q]], t);
    if (dt_rp)
      for (int q = dt_rp[i]; q < dt_rp[i + 1]; q++) t = fma(-jd_v[dt_map[q]], dyd[dt_ci[q]], t);
    sq[i] = t * t;
    if (e) e[i] = t;
  }
}

__global__ void k_resid_m(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                          const double* __restrict__ v, const double* __restrict__ ry,
                          const double* __restrict__ dx, double* __restrict__ sq,
                          double* __restrict__ e, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  v = at(v, s, b);
  ry = at(ry, s, b);
  dx = at(dx, s, b);
  sq = at(sq, s, b);
  e = at(e, s, b);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
    double t = ry[k];
    for (int p = rp[k]; p < rp[k + 1]; p++) t = fma(-v[p], dx[ci[p]], t);
    sq[k] = t * t;
    if (e) e[k] = t;
  }
}

__global__ void k_resid_d(int d, const int* __restrict__ rp, const int* __restrict__ ci,
                          const double* __restrict__ v, int has_jd,
                          const double* __restrict__ ds, const double* __restrict__ rs,
                          const double* __restrict__ ryd, const double* __restrict__ dx,
                          const double* __restrict__ dsv, const double* __restrict__ dyd,
                          double* __restrict__ sq2, double* __restrict__ sq4,
                          double* __restrict__ e2, double* __restrict__ e4, Batch b)
{
  const int s = sys_of(b);
  if (s < 0) return;
  v = at(v, s, b);
  ds = at(ds, s, b);
  rs = at(rs, s, b);
  ryd = at(ryd, s, b);
  dx = at(dx, s, b);
  dsv = at(dsv, s, b);
  dyd = at(dyd, s, b);
  sq2 = at(sq2, s, b);
  sq4 = at(sq4, s, b);
  e2 = at(e2, s, b);
  e4 = at(e4, s, b);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < d; k += gridDim.x * blockDim.x) {
    double a2 = rs[k] - (ds[k] * dsv[k] - dyd[k]);
    double jx = 0.0;
    if (has_jd)
      for (int p = rp[k]; p < rp[k + 1]; p++) jx = fma(v[p], dx[ci[p]], jx);
    double a4 = ryd[k] - (jx - dsv[k]);
    sq2[k] = a2 * a2;
    sq4[k] = a4 * a4;
    if (e2) e2[k] = a2;
    if (e4) e4[k] = a4;
  }
}

__global__ void k_recover_d(int d, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ v, int has_jd,
                            const double* __restrict__ dx, const double* __restrict__ ryd,
                            const double* __restrict__ ds, const double* __restrict__ rs,
                            double* __restrict__ dsv, double* __restrict__ dyd, Batch b)
{
  const int sy = sys_of(b);
  if (sy < 0) return;
  v = at(v, sy, b);
  dx = at(dx, sy, b);
  ryd = at(ryd, sy, b);
  ds = at(ds, sy, b);
  rs = at(rs, sy, b);
  dsv = at(dsv, sy, b);
  dyd = at(dyd, sy, b);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < d; k += gridDim.x * blockDim.x) {
    double s;
    if (has_jd) {
      double t = 0.0;
      for (int p = rp[k]; p < rp[k + 1]; p++) t = fma(v[p], dx[ci[p]], t);
      s = fma(1.0, t, -1.0 * ryd[k]);          // Jd*dx - ryd (HykktSolver.cpp:429-440)
    } else {
      s = ryd[k] * -1.0;                       // mult_const(-1) (HykktSolver.cpp:443)
    }
    dsv[k] = s;
    double y = s * ds[k];                      // vec_scale (HykktSolver.cpp:447)
    dyd[k] = y + -1.0 * rs[k];                 // add_vecs(-1) (HykktSolver.cpp:448)
  }
}

// ok[s] = 1 when system s's factorization met no failing pivot, else 0,
// and then its solution (dx, dsv, dyc, dyd) is zero: such a system takes no
// part in the solve phase.
__global__ void k_pivot_split(const int* __restrict__ pivot, int* __restrict__ ok, int n, int m,
                              int d, double* __restrict__ dx, double* __restrict__ dsv,
                              double* __restrict__ dyc, double* __restrict__ dyd, Batch b)
{
  const int sy = sys_of(b);
  if (sy < 0) return;
  const bool failed = pivot[sy] != INT_MAX;
  if (blockIdx.x == 0 && threadIdx.x == 0) ok[sy] = failed ? 0 : 1;
  if (!failed) return;
  dx = at(dx, sy, b);
  dsv = at(dsv, sy, b);
  dyc = at(dyc, sy, b);
  dyd = at(dyd, sy, b);
  const int len = max(n, max(m, d));
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
    if (i < n) dx[i] = 0.0;
    if (i < m) dyc[i] = 0.0;
    if (i < d) {
      dsv[i] = 0.0;
      dyd[i] = 0.0;
    }
  }
}

}  // namespace

// =================================================================== launchers
#define LAUNCH_CHECK() HIPCHK(hipGetLastError())

void fill(double* p, double v, int64_t n, const Batch& b, hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_fill<<<grid2(n, b), TPB, 0, st>>>(p, v, n, b);
  LAUNCH_CHECK();
}

void fill_i(int* p, int v, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  k_fill_i<<<nblocks(n), TPB, 0, st>>>(p, v, n);
  LAUNCH_CHECK();
}

void copy(double* d, const double* s, int64_t n, const Batch& b, hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_copy<<<grid2(n, b), TPB, 0, st>>>(d, s, n, nullptr, b);
  LAUNCH_CHECK();
}

void copy_where(double* d, const double* s, int64_t n, const int* keep, const Batch& b,
                hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_copy<<<grid2(n, b), TPB, 0, st>>>(d, s, n, keep, b);
  LAUNCH_CHECK();
}

void copy_in(int nsrc, double* const* dst, const double* const* src_tab, const int64_t* len,
             const Batch& b, hipStream_t st)
{
  if (nsrc <= 0 || b.count <= 0) return;
  HYKKT_REQUIRE(nsrc <= 12, HYKKT_E_ARG, "copy_in: too many arrays");
  CopyDesc c{};
  c.nsrc = nsrc;
  int64_t mx = 1;
  for (int j = 0; j < nsrc; j++) {
    c.dst[j] = dst[j];
    c.len[j] = len[j];
    mx = std::max(mx, len[j]);
  }
  k_copy_in<<<dim3(nblocks(mx), b.count, nsrc), TPB, 0, st>>>(c, src_tab, b);
  LAUNCH_CHECK();
}

void collect(double* out, const double* base, int nslot, const Batch& b, hipStream_t st)
{
  const int64_t n = static_cast<int64_t>(nslot) * b.count;
  if (n <= 0) return;
  k_collect<<<nblocks(n), TPB, 0, st>>>(out, base, nslot, b);
  LAUNCH_CHECK();
}

void gather(double* d, const double* s, const int* map, int64_t n, const Batch& b,
            hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_gather<<<grid2(n, b), TPB, 0, st>>>(d, s, map, n, b);
  LAUNCH_CHECK();
}

void gather64(double* d, const double* s, const int64_t* map, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  k_gather64<<<nblocks(n), TPB, 0, st>>>(d, s, map, n);
  LAUNCH_CHECK();
}

void gather_scale(double* d, const double* s, const int* map, const double* sc, int n,
                  const Batch& b, hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_gather_scale<<<grid2(n, b), TPB, 0, st>>>(d, s, map, sc, n, b);
  LAUNCH_CHECK();
}

void recover_xy(double* dx, const double* h0, const double* cz, const int* map,
                const double* max_d, int n, double* dyc, const double* cx, int m, const Batch& b,
                hipStream_t st)
{
  if (b.count <= 0 || (n <= 0 && m <= 0)) return;
  k_recover_xy<<<grid2(std::max(n, m), b), TPB, 0, st>>>(dx, h0, cz, map, max_d, n, dyc, cx, m,
                                                         b);
  LAUNCH_CHECK();
}

void vec_scale(double* y, const double* s, int n, const Batch& b, hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_vec_scale<<<grid2(n, b), TPB, 0, st>>>(y, s, n, b);
  LAUNCH_CHECK();
}

void add(double* out, const double* a, const double* bv, int64_t n, const Batch& b,
         hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_add<<<grid2(n, b), TPB, 0, st>>>(out, a, bv, n, b);
  LAUNCH_CHECK();
}

void row_scale(int d, const int* rp, const double* v, const double* ds, const double* ryd,
               const double* rs, double* vs, double* ryd_s, const Batch& b, hipStream_t st)
{
  if (d <= 0 || b.count <= 0) return;
  k_row_scale<<<grid2(d, b), TPB, 0, st>>>(d, rp, v, ds, ryd, rs, vs, ryd_s, b);
  LAUNCH_CHECK();
}

int spmv_lds_entries() { return SPMV_E; }

void spmv(int nrows, const int* rp, const int* ci, const int* map, const double* v, double alpha,
          const double* alpha_ps, const double* x, double beta, const double* y_in,
          double* y_out, const int* blocks, int nblk, const Batch& b, hipStream_t st,
          const double* dot_w, double* dot_part)
{
  if (nrows <= 0 || nblk <= 0 || b.count <= 0) return;
  k_spmv<<<dim3(nblk, b.count), TPB, 0, st>>>(rp, ci, map, v, alpha, alpha_ps, x, beta, y_in,
                                               y_out, blocks, dot_w, dot_part, b);
  LAUNCH_CHECK();
}

void spgemm_values(int nnz, const int* d_src, const double* d_v, const int* tp, const int* ta,
                   const int* tb, const double* av, const double* bv, double alpha_p,
                   double* out, const Batch& b, hipStream_t st)
{
  if (nnz <= 0 || b.count <= 0) return;
  k_spgemm_values<<<grid2(nnz, b), TPB, 0, st>>>(nnz, d_src, d_v, tp, ta, tb, av, bv, alpha_p,
                                                 out, b);
  LAUNCH_CHECK();
}

void spgemm_scatter(int nent, const int* src, const int64_t* dst, const int* d_src,
                    const double* d_v, const int* tp, const int* ta, const int* tb,
                    const double* av, const double* bv, const double* alpha_ps, double* out,
                    const Batch& b, hipStream_t st)
{
  if (nent <= 0 || b.count <= 0) return;
  k_spgemm_scatter<<<grid2(nent, b), TPB, 0, st>>>(nent, src, dst, d_src, d_v, tp, ta, tb, av,
                                                   bv, alpha_ps, out, b);
  LAUNCH_CHECK();
}

void scatter(double* out, const int64_t* dst, const double* vals, const int* src, int n,
             const Batch& b, hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_scatter<<<grid2(n, b), TPB, 0, st>>>(out, dst, vals, src, n, b);
  LAUNCH_CHECK();
}

void add_at(double* out, const int64_t* pos, int n, double delta, const Batch& b,
            hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_add_at<<<grid2(n, b), TPB, 0, st>>>(out, pos, n, delta, b);
  LAUNCH_CHECK();
}

void ruiz_rowmax(int n, int totn, const int* h_rp, const double* h_v, const int* jt_rp,
                 const int* jt_map, const double* jt_v, const int* j_rp, const double* j_v,
                 double* scale, int* zero_row, const Batch& b, hipStream_t st)
{
  if (totn <= 0 || b.count <= 0) return;
  k_ruiz_rowmax<<<grid2(totn, b), TPB, 0, st>>>(n, totn, h_rp, h_v, jt_rp, jt_map, jt_v, j_rp,
                                                j_v, scale, zero_row, b);
  LAUNCH_CHECK();
}

void ruiz_scale(int n, int totn, const int* h_rp, const int* h_ci, double* h_v,
                const int* jt_rp, const int* jt_ci, double* jt_v, const int* j_rp,
                const int* j_ci, double* j_v, const double* scale, double* rhs1, double* rhs2,
                double* max_d, int first, const Batch& b, hipStream_t st)
{
  if (totn <= 0 || b.count <= 0) return;
  k_ruiz_scale<<<grid2(totn, b), TPB, 0, st>>>(n, totn, h_rp, h_ci, h_v, jt_rp, jt_ci, jt_v,
                                               j_rp, j_ci, j_v, scale, rhs1, rhs2, max_d, first,
                                               b);
  LAUNCH_CHECK();
}

int reduce_blocks(int64_t n)
{
  int64_t g = (n + TPB - 1) / TPB;
  return static_cast<int>(std::clamp<int64_t>(g, 1, RED_MAX));
}

void dot(int n, const double* a, const double* bv, double* part, double* out, const Batch& b,
         hipStream_t st)
{
  if (b.count <= 0) return;
  const int g = reduce_blocks(n);
  k_dot_partial<<<dim3(g, b.count), TPB, 0, st>>>(n, a, bv, part, b);
  LAUNCH_CHECK();
  // out is a per-system slot: sc = out, slot 0
  k_dot_final<<<dim3(1, b.count), TPB, 0, st>>>(g, part, out, 0, 0, 0, 0, 0.0, 0,
                                                CgFlags{nullptr, nullptr, nullptr}, b);
  LAUNCH_CHECK();
}

void dots(DotSet d, double* part, double* out, const Batch& b, hipStream_t st)
{
  if (b.count <= 0 || d.count <= 0) return;
  d.off[0] = 0;
  for (int k = 0; k < d.count; k++) d.off[k + 1] = d.off[k] + reduce_blocks(d.n[k]);
  k_dots_partial<<<dim3(d.off[d.count], b.count), TPB, 0, st>>>(d, part, b);
  LAUNCH_CHECK();
  k_dots_final<<<dim3(1, b.count), TPB, 0, st>>>(d, part, out, b);
  LAUNCH_CHECK();
}

void cg_pack(double* out, const double* sc, const int* iters, const int* status, const Batch& b,
             hipStream_t st)
{
  if (b.count <= 0) return;
  k_cg_pack<<<nblocks(b.count), TPB, 0, st>>>(out, sc, iters, status, b);
  LAUNCH_CHECK();
}

void collect_ints(int* out, const int* w, const int* a, const int* bv, int count, hipStream_t st)
{
  k_collect_ints<<<nblocks(std::max(count, 1)), TPB, 0, st>>>(out, w, a, bv, count);
  LAUNCH_CHECK();
}

void dot_partial(int n, const double* a, const double* bv, double* part, const Batch& b,
                 hipStream_t st)
{
  if (b.count <= 0) return;
  k_dot_partial<<<dim3(reduce_blocks(n), b.count), TPB, 0, st>>>(n, a, bv, part, b);
  LAUNCH_CHECK();
}

void sub(double* out, const double* a, const double* bv, int64_t n, const Batch& b,
         hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_sub<<<grid2(n, b), TPB, 0, st>>>(out, a, bv, n, b);
  LAUNCH_CHECK();
}

void cg_track(int n, const double* sc, double* q, double* z, const double* t, const Batch& b,
              hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_cg_track<<<dim3(reduce_blocks(n), b.count), TPB, 0, st>>>(n, sc, q, z, t, b);
  LAUNCH_CHECK();
}

void cg_update_dot(int m, const double* sc, double* p, double* s, double* x, double* r,
                   const double* w, double* part, const Batch& b, hipStream_t st)
{
  if (m <= 0 || b.count <= 0) return;
  k_cg_update_dot<<<dim3(reduce_blocks(m), b.count), TPB, 0, st>>>(m, sc, p, s, x, r, w, part,
                                                                    b);
  LAUNCH_CHECK();
}

void dot_final(int np, double* part, double* sc, int slot, bool check, bool scalars, int it,
               double tol, int itmax, const CgFlags& f, const Batch& b, hipStream_t st)
{
  if (b.count <= 0) return;
  k_dot_final<<<dim3(1, b.count), TPB, 0, st>>>(np, part, sc, slot, check ? 1 : 0,
                                                scalars ? 1 : 0, it, tol, itmax, f, b);
  LAUNCH_CHECK();
}

void resid_rows_n(int n, const int* h_rp, const int* h_ci, const double* h_v, const int* ct_rp,
                  const int* ct_ci, const int* ct_map, const double* jc_v, const int* dt_rp,
                  const int* dt_ci, const int* dt_map, const double* jd_v, const double* rx,
                  const double* dx, const double* dyc, const double* dyd, double* sq,
                  double* e, const Batch& b, hipStream_t st)
{
  if (n <= 0 || b.count <= 0) return;
  k_resid_n<<<grid2(n, b), TPB, 0, st>>>(n, h_rp, h_ci, h_v, ct_rp, ct_ci, ct_map, jc_v, dt_rp,
                                         dt_ci, dt_map, jd_v, rx, dx, dyc, dyd, sq, e, b);
  LAUNCH_CHECK();
}

void resid_rows_m(int m, const int* rp, const int* ci, const double* v, const double* ry,
                  const double* dx, double* sq, double* e, const Batch& b, hipStream_t st)
{
  if (m <= 0 || b.count <= 0) return;
  k_resid_m<<<grid2(m, b), TPB, 0, st>>>(m, rp, ci, v, ry, dx, sq, e, b);
  LAUNCH_CHECK();
}

void resid_rows_d(int d, const int* rp, const int* ci, const double* v, int has_jd,
                  const double* ds, const double* rs, const double* ryd, const double* dx,
                  const double* dsv, const double* dyd, double* sq2, double* sq4, double* e2,
                  double* e4, const Batch& b, hipStream_t st)
{
  if (d <= 0 || b.count <= 0) return;
  k_resid_d<<<grid2(d, b), TPB, 0, st>>>(d, rp, ci, v, has_jd, ds, rs, ryd, dx, dsv, dyd, sq2,
                                         sq4, e2, e4, b);
  LAUNCH_CHECK();
}

void pivot_split(const int* pivot, int* ok, int n, int m, int d, double* dx, double* dsv,
                 double* dyc, double* dyd, const Batch& b, hipStream_t st)
{
  if (b.count <= 0) return;
  const int len = std::max(1, std::max(n, std::max(m, d)));
  k_pivot_split<<<grid2(len, b), TPB, 0, st>>>(pivot, ok, n, m, d, dx, dsv, dyc, dyd, b);
  LAUNCH_CHECK();
}

void recover_d(int d, const int* rp, const int* ci, const double* v, int has_jd,
               const double* dx, const double* ryd, const double* ds, const double* rs,
               double* dsv, double* dyd, const Batch& b, hipStream_t st)
{
  if (d <= 0 || b.count <= 0) return;
  k_recover_d<<<grid2(d, b), TPB, 0, st>>>(d, rp, ci, v, has_jd, dx, ryd, ds, rs, dsv, dyd, b);
  LAUNCH_CHECK();
}

}  // namespace dev
}  // namespace hykkt
