// kernels_basic.hip -- gfx950 kernels for the formation, Ruiz, SpMV, CG and
// recovery steps of the HyKKT solve.
//
// Built with -ffp-contract=off: every fused multiply-add here is an explicit
// fma(), so the SpGEMM/Ruiz/rhs values are bit-identical to the CPU oracle,
// which uses the same operation order.
#include <hip/hip_runtime.h>

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

inline unsigned nblocks(int64_t n, int b = TPB)
{
  int64_t g = (n + b - 1) / b;
  return static_cast<unsigned>(g > 0 ? g : 1);
}

// ------------------------------------------------------------- vectors
__global__ void k_fill(double* __restrict__ p, double v, int64_t n)
{
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

__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

__global__ void k_gather(double* __restrict__ d, const double* __restrict__ s,
                         const int* __restrict__ map, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[map[i]];
}

__global__ void k_gather64(double* __restrict__ d, const double* __restrict__ s,
                           const int64_t* __restrict__ map, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[map[i]];
}

__global__ void k_gather_scale(double* __restrict__ d, const double* __restrict__ s,
                               const int* __restrict__ map, const double* __restrict__ sc, int n)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[map[i]] * sc[i];
}

__global__ void k_vec_scale(double* __restrict__ y, const double* __restrict__ s, int n)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] *= s[i];
}

// ------------------------------------------------------------- sparse
__global__ void k_row_scale(int d, const int* __restrict__ rp, const double* __restrict__ v,
                            const double* __restrict__ ds, const double* __restrict__ ryd,
                            const double* __restrict__ rs, double* __restrict__ vs,
                            double* __restrict__ ryd_s)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d) return;
  double di = ds[i];
  if (vs)
    for (int p = rp[i]; p < rp[i + 1]; p++) vs[p] = v[p] * di;
  double t = ryd[i] * di;
  ryd_s[i] = t + 1.0 * rs[i];
}

__global__ void k_spmv(int nrows, const int* __restrict__ rp, const int* __restrict__ ci,
                       const int* __restrict__ map, const double* __restrict__ v, double alpha,
                       const double* __restrict__ x, double beta,
                       const double* __restrict__ y_in, double* __restrict__ y_out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  double t = 0.0;
  if (map) {
    for (int q = rp[i]; q < rp[i + 1]; q++) t = fma(v[map[q]], x[ci[q]], t);
  } else {
    for (int q = rp[i]; q < rp[i + 1]; q++) t = fma(v[q], x[ci[q]], t);
  }
  y_out[i] = (beta == 0.0) ? alpha * t : fma(alpha, t, beta * y_in[i]);
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
                                double alpha_p, double* __restrict__ out)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nnz) out[p] = spgemm_entry(p, d_src, d_v, tp, ta, tb, av, bv, alpha_p);
}

__global__ void k_spgemm_scatter(int nent, const int* __restrict__ src,
                                 const int64_t* __restrict__ dst,
                                 const int* __restrict__ d_src, const double* __restrict__ d_v,
                                 const int* __restrict__ tp, const int* __restrict__ ta,
                                 const int* __restrict__ tb, const double* __restrict__ av,
                                 const double* __restrict__ bv, double alpha_p,
                                 double* __restrict__ out)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nent) out[dst[e]] = spgemm_entry(src[e], d_src, d_v, tp, ta, tb, av, bv, alpha_p);
}

__global__ void k_scatter(double* __restrict__ out, const int64_t* __restrict__ dst,
                          const double* __restrict__ vals, const int* __restrict__ src, int n)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[dst[e]] = vals[src[e]];
}

__global__ void k_add_at(double* __restrict__ out, const int64_t* __restrict__ pos, int n,
                         double delta)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[pos[i]] += delta;
}

// ------------------------------------------------------------- Ruiz
__global__ void k_ruiz_rowmax(int n, int totn, const int* __restrict__ h_rp,
                              const double* __restrict__ h_v, const int* __restrict__ jt_rp,
                              const int* __restrict__ jt_map, const double* __restrict__ jt_v,
                              const int* __restrict__ j_rp, const double* __restrict__ j_v,
                              double* __restrict__ scale, int* __restrict__ zero_row)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= totn) return;
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
  if (zero_row && !(max_l > 0.0 || max_u > 0.0)) atomicMin(zero_row, i);
}

__global__ void k_ruiz_scale(int n, int totn, const int* __restrict__ h_rp,
                             const int* __restrict__ h_ci, double* __restrict__ h_v,
                             const int* __restrict__ jt_rp, const int* __restrict__ jt_ci,
                             double* __restrict__ jt_v, const int* __restrict__ j_rp,
                             const int* __restrict__ j_ci, double* __restrict__ j_v,
                             const double* __restrict__ scale, double* __restrict__ rhs1,
                             double* __restrict__ rhs2, double* __restrict__ max_d, int first)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= totn) return;
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

// ------------------------------------------------------------- reductions
constexpr int RED_BLOCKS = 256;

__device__ double block_sum(double v)
{
  __shared__ double sh[TPB];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

__global__ void k_dot_partial(int n, const double* __restrict__ a, const double* __restrict__ b,
                              double* __restrict__ part)
{
  double t = 0.0;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    t = b ? fma(a[i], b[i], t) : t + a[i];
  double s = block_sum(t);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_dot_final(int g, const double* __restrict__ part, double* __restrict__ out)
{
  double t = 0.0;
  for (int i = threadIdx.x; i < g; i += TPB) t += part[i];
  double s = block_sum(t);
  if (threadIdx.x == 0) *out = s;
}

// ------------------------------------------------------------- CG
__global__ void k_cg_update(int m, const double* __restrict__ sc, double* __restrict__ p,
                            double* __restrict__ s, double* __restrict__ x,
                            double* __restrict__ r, const double* __restrict__ w)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double alpha = sc[3], beta = sc[4];
  double pi = beta * p[i] + r[i];
  double si = beta * s[i] + w[i];
  p[i] = pi;
  s[i] = si;
  x[i] = fma(alpha, pi, x[i]);
  r[i] = fma(-alpha, si, r[i]);
}

__global__ void k_cg_scalars(double* sc, int init)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (init) {
    // alpha = gam / delta (SchurComplementConjugateGradient.cpp:143), beta = 0
    sc[3] = sc[0] / sc[2];
    sc[4] = 0.0;
  } else {
    // beta = gam1/gam; gam = gam1; alpha = gam/(delta - beta*gam/alpha) (:181-183)
    double beta = sc[1] / sc[0];
    double gam = sc[1];
    sc[4] = beta;
    sc[0] = gam;
    sc[3] = gam / (sc[2] - beta * gam / sc[3]);
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
                          double* __restrict__ e)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t = rx[i];
  for (int p = h_rp[i]; p < h_rp[i + 1]; p++) t = fma(-h_v[p], dx[h_ci[p]], t);
  for (int q = ct_rp[i]; q < ct_rp[i + 1]; q++) t = fma(-jc_v[ct_map[q]], dyc[ct_ci[q]], t);
  if (dt_rp)
    for (int q = dt_rp[i]; q < dt_rp[i + 1]; q++) t = fma(-jd_v[dt_map[q]], dyd[dt_ci[q]], t);
  sq[i] = t * t;
  if (e) e[i] = t;
}

__global__ void k_resid_m(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                          const double* __restrict__ v, const double* __restrict__ ry,
                          const double* __restrict__ dx, double* __restrict__ sq,
                          double* __restrict__ e)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  double t = ry[k];
  for (int p = rp[k]; p < rp[k + 1]; p++) t = fma(-v[p], dx[ci[p]], t);
  sq[k] = t * t;
  if (e) e[k] = t;
}

__global__ void k_resid_d(int d, const int* __restrict__ rp, const int* __restrict__ ci,
                          const double* __restrict__ v, int has_jd,
                          const double* __restrict__ ds, const double* __restrict__ rs,
                          const double* __restrict__ ryd, const double* __restrict__ dx,
                          const double* __restrict__ dsv, const double* __restrict__ dyd,
                          double* __restrict__ sq2, double* __restrict__ sq4,
                          double* __restrict__ e2, double* __restrict__ e4)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d) return;
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

__global__ void k_add(double* __restrict__ out, const double* __restrict__ a,
                      const double* __restrict__ b, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] + b[i];
}

__global__ void k_recover_d(int d, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ v, int has_jd,
                            const double* __restrict__ dx, const double* __restrict__ ryd,
                            const double* __restrict__ ds, const double* __restrict__ rs,
                            double* __restrict__ dsv, double* __restrict__ dyd)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d) return;
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

}  // namespace

// =================================================================== launchers
void fill(double* p, double v, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  unsigned g = nblocks(n);
  if (g > 65536) g = 65536;
  k_fill<<<g, TPB, 0, st>>>(p, v, n);
  HIPCHK(hipGetLastError());
}

void fill_i(int* p, int v, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  unsigned g = nblocks(n);
  if (g > 65536) g = 65536;
  k_fill_i<<<g, TPB, 0, st>>>(p, v, n);
  HIPCHK(hipGetLastError());
}

void copy(double* d, const double* s, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  unsigned g = nblocks(n);
  if (g > 65536) g = 65536;
  k_copy<<<g, TPB, 0, st>>>(d, s, n);
  HIPCHK(hipGetLastError());
}

void gather(double* d, const double* s, const int* map, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  unsigned g = nblocks(n);
  if (g > 65536) g = 65536;
  k_gather<<<g, TPB, 0, st>>>(d, s, map, n);
  HIPCHK(hipGetLastError());
}

void gather64(double* d, const double* s, const int64_t* map, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  unsigned g = nblocks(n);
  if (g > 65536) g = 65536;
  k_gather64<<<g, TPB, 0, st>>>(d, s, map, n);
  HIPCHK(hipGetLastError());
}

void gather_scale(double* d, const double* s, const int* map, const double* sc, int n,
                  hipStream_t st)
{
  if (n <= 0) return;
  k_gather_scale<<<nblocks(n), TPB, 0, st>>>(d, s, map, sc, n);
  HIPCHK(hipGetLastError());
}

void vec_scale(double* y, const double* s, int n, hipStream_t st)
{
  if (n <= 0) return;
  k_vec_scale<<<nblocks(n), TPB, 0, st>>>(y, s, n);
  HIPCHK(hipGetLastError());
}

void row_scale(int d, const int* rp, const double* v, const double* ds, const double* ryd,
               const double* rs, double* vs, double* ryd_s, hipStream_t st)
{
  if (d <= 0) return;
  k_row_scale<<<nblocks(d), TPB, 0, st>>>(d, rp, v, ds, ryd, rs, vs, ryd_s);
  HIPCHK(hipGetLastError());
}

void spmv(int nrows, const int* rp, const int* ci, const int* map, const double* v, double alpha,
          const double* x, double beta, const double* y_in, double* y_out, hipStream_t st)
{
  if (nrows <= 0) return;
  k_spmv<<<nblocks(nrows), TPB, 0, st>>>(nrows, rp, ci, map, v, alpha, x, beta, y_in, y_out);
  HIPCHK(hipGetLastError());
}

void spgemm_values(int nnz, const int* d_src, const double* d_v, const int* tp, const int* ta,
                   const int* tb, const double* av, const double* bv, double alpha_p,
                   double* out, hipStream_t st)
{
  if (nnz <= 0) return;
  k_spgemm_values<<<nblocks(nnz), TPB, 0, st>>>(nnz, d_src, d_v, tp, ta, tb, av, bv, alpha_p,
                                                out);
  HIPCHK(hipGetLastError());
}

void spgemm_scatter(int nent, const int* src, const int64_t* dst, const int* d_src,
                    const double* d_v, const int* tp, const int* ta, const int* tb,
                    const double* av, const double* bv, double alpha_p, double* out,
                    hipStream_t st)
{
  if (nent <= 0) return;
  k_spgemm_scatter<<<nblocks(nent), TPB, 0, st>>>(nent, src, dst, d_src, d_v, tp, ta, tb, av,
                                                  bv, alpha_p, out);
  HIPCHK(hipGetLastError());
}

void scatter(double* out, const int64_t* dst, const double* vals, const int* src, int n,
             hipStream_t st)
{
  if (n <= 0) return;
  k_scatter<<<nblocks(n), TPB, 0, st>>>(out, dst, vals, src, n);
  HIPCHK(hipGetLastError());
}

void add_at(double* out, const int64_t* pos, int n, double delta, hipStream_t st)
{
  if (n <= 0) return;
  k_add_at<<<nblocks(n), TPB, 0, st>>>(out, pos, n, delta);
  HIPCHK(hipGetLastError());
}

void ruiz_rowmax(int n, int totn, const int* h_rp, const double* h_v, const int* jt_rp,
                 const int* jt_map, const double* jt_v, const int* j_rp, const double* j_v,
                 double* scale, hipStream_t st, int* zero_row)
{
  if (totn <= 0) return;
  k_ruiz_rowmax<<<nblocks(totn), TPB, 0, st>>>(n, totn, h_rp, h_v, jt_rp, jt_map, jt_v, j_rp,
                                               j_v, scale, zero_row);
  HIPCHK(hipGetLastError());
}

void ruiz_scale(int n, int totn, const int* h_rp, const int* h_ci, double* h_v,
                const int* jt_rp, const int* jt_ci, double* jt_v, const int* j_rp,
                const int* j_ci, double* j_v, const double* scale, double* rhs1, double* rhs2,
                double* max_d, int first, hipStream_t st)
{
  if (totn <= 0) return;
  k_ruiz_scale<<<nblocks(totn), TPB, 0, st>>>(n, totn, h_rp, h_ci, h_v, jt_rp, jt_ci, jt_v,
                                              j_rp, j_ci, j_v, scale, rhs1, rhs2, max_d, first);
  HIPCHK(hipGetLastError());
}

int reduce_partials() { return RED_BLOCKS; }

void dot(int n, const double* a, const double* b, double* part, double* out, hipStream_t st)
{
  int g = static_cast<int>(nblocks(n));
  if (g > RED_BLOCKS) g = RED_BLOCKS;
  k_dot_partial<<<g, TPB, 0, st>>>(n, a, b, part);
  HIPCHK(hipGetLastError());
  k_dot_final<<<1, TPB, 0, st>>>(g, part, out);
  HIPCHK(hipGetLastError());
}

void cg_update(int m, const double* sc, double* p, double* s, double* x, double* r,
               const double* w, hipStream_t st)
{
  if (m <= 0) return;
  k_cg_update<<<nblocks(m), TPB, 0, st>>>(m, sc, p, s, x, r, w);
  HIPCHK(hipGetLastError());
}

void cg_scalars(double* sc, int init, hipStream_t st)
{
  k_cg_scalars<<<1, 64, 0, st>>>(sc, init);
  HIPCHK(hipGetLastError());
}

void resid_rows_n(int n, const int* h_rp, const int* h_ci, const double* h_v, const int* ct_rp,
                  const int* ct_ci, const int* ct_map, const double* jc_v, const int* dt_rp,
                  const int* dt_ci, const int* dt_map, const double* jd_v, const double* rx,
                  const double* dx, const double* dyc, const double* dyd, double* sq,
                  double* e, hipStream_t st)
{
  if (n <= 0) return;
  k_resid_n<<<nblocks(n), TPB, 0, st>>>(n, h_rp, h_ci, h_v, ct_rp, ct_ci, ct_map, jc_v, dt_rp,
                                        dt_ci, dt_map, jd_v, rx, dx, dyc, dyd, sq, e);
  HIPCHK(hipGetLastError());
}

void resid_rows_m(int m, const int* rp, const int* ci, const double* v, const double* ry,
                  const double* dx, double* sq, double* e, hipStream_t st)
{
  if (m <= 0) return;
  k_resid_m<<<nblocks(m), TPB, 0, st>>>(m, rp, ci, v, ry, dx, sq, e);
  HIPCHK(hipGetLastError());
}

void resid_rows_d(int d, const int* rp, const int* ci, const double* v, int has_jd,
                  const double* ds, const double* rs, const double* ryd, const double* dx,
                  const double* dsv, const double* dyd, double* sq2, double* sq4, double* e2,
                  double* e4, hipStream_t st)
{
  if (d <= 0) return;
  k_resid_d<<<nblocks(d), TPB, 0, st>>>(d, rp, ci, v, has_jd, ds, rs, ryd, dx, dsv, dyd, sq2,
                                        sq4, e2, e4);
  HIPCHK(hipGetLastError());
}

void add(double* out, const double* a, const double* b, int64_t n, hipStream_t st)
{
  if (n <= 0) return;
  unsigned g = nblocks(n);
  if (g > 65536) g = 65536;
  k_add<<<g, TPB, 0, st>>>(out, a, b, n);
  HIPCHK(hipGetLastError());
}

void recover_d(int d, const int* rp, const int* ci, const double* v, int has_jd,
               const double* dx, const double* ryd, const double* ds, const double* rs,
               double* dsv, double* dyd, hipStream_t st)
{
  if (d <= 0) return;
  k_recover_d<<<nblocks(d), TPB, 0, st>>>(d, rp, ci, v, has_jd, dx, ryd, ds, rs, dsv, dyd);
  HIPCHK(hipGetLastError());
}

}  // namespace dev
}  // namespace hykkt
