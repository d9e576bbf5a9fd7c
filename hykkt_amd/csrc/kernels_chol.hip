// kernels_chol.hip -- gfx950 kernels of the supernodal multifrontal Cholesky
// (numeric factorization and level-scheduled triangular solves).
//
// Storage (see CholPlan): supernode s owns columns [a, a+nc) and the sorted
// off-diagonal-block rows `rows(s)` (nr of them); its front index list is
// [cols; rows] (f = nc + nr).  The factor panel is f x nc column-major at
// L + l_off[s]; the update matrix is nr x nr column-major at U + u_off[s].
#include <hip/hip_runtime.h>

#include <climits>

#include "kernels.hpp"

namespace hykkt {
namespace dev {

namespace {

constexpr int NB = 64;     // panel block width (columns per factorization step)
constexpr int TS = 64;     // output tile edge of the update kernels

struct SnInfo {
  int a, nc, nr, f;
  double* Ls;
};

__device__ __forceinline__ SnInfo sninfo(const CholDev& c, int s)
{
  SnInfo i;
  i.a = c.sn_first[s];
  i.nc = c.sn_first[s + 1] - i.a;
  i.nr = c.rows_ptr[s + 1] - c.rows_ptr[s];
  i.f = i.nc + i.nr;
  i.Ls = c.L + c.l_off[s];
  return i;
}

// ----------------------------------------------------------------- extend-add
// One workgroup per parent front column.  Children are added one after the
// other (fixed order => deterministic sums); within one child the relative
// index map is injective.
__global__ void __launch_bounds__(256) k_extend_add(CholDev c, const int* __restrict__ task_p,
                                                    const int* __restrict__ task_col,
                                                    const int* __restrict__ cptr,
                                                    const int* __restrict__ cc,
                                                    const int* __restrict__ ccol)
{
  const int t = blockIdx.x;
  const int p = task_p[t], col = task_col[t];
  const SnInfo ip = sninfo(c, p);
  double* base;
  int shift;
  if (col < ip.nc) {
    base = ip.Ls + static_cast<int64_t>(col) * ip.f;
    shift = 0;
  } else {
    base = c.U + c.u_off[p] + static_cast<int64_t>(col - ip.nc) * ip.nr;
    shift = -ip.nc;
  }
  for (int k = cptr[t]; k < cptr[t + 1]; k++) {
    const int ch = cc[k], r2 = ccol[k];
    const int nrc = c.rows_ptr[ch + 1] - c.rows_ptr[ch];
    const double* src = c.U + c.u_off[ch] + static_cast<int64_t>(r2) * nrc;
    const int* relc = c.rel + c.rows_ptr[ch];
    for (int r1 = r2 + threadIdx.x; r1 < nrc; r1 += blockDim.x) base[relc[r1] + shift] += src[r1];
    __syncthreads();
  }
}

// ----------------------------------------------------------------- POTRF
// Factor the diagonal NB x NB block of step `step` of each listed supernode.
__global__ void __launch_bounds__(256) k_potrf(CholDev c, const int* __restrict__ sns, int step)
{
  const SnInfo si = sninfo(c, sns[blockIdx.x]);
  const int k0 = step * NB;
  if (k0 >= si.nc) return;
  const int b = min(NB, si.nc - k0);
  __shared__ double A[NB][NB + 1];
  __shared__ int bad;
  double* D = si.Ls + static_cast<int64_t>(k0) * si.f + k0;   // D[r + col*f]
  for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
    int col = idx / b, r = idx % b;
    A[r][col] = (r >= col) ? D[r + static_cast<int64_t>(col) * si.f] : 0.0;
  }
  if (threadIdx.x == 0) bad = INT_MAX;
  __syncthreads();
  for (int k = 0; k < b; k++) {
    if (threadIdx.x == 0) {
      double d = A[k][k];
      if (!(d > 0.0) || !(sqrt(d) > c.pivot_tol)) {
        if (bad == INT_MAX) bad = si.a + k0 + k;
        d = 1.0;   // keep the arithmetic finite; the failure is reported
      }
      A[k][k] = sqrt(d);
    }
    __syncthreads();
    const double dk = A[k][k];
    for (int r = k + 1 + threadIdx.x; r < b; r += blockDim.x) A[r][k] /= dk;
    __syncthreads();
    const int m = b - k - 1;
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
      int r = k + 1 + idx / m, col = k + 1 + idx % m;
      if (col <= r) A[r][col] -= A[r][k] * A[col][k];
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
    int col = idx / b, r = idx % b;
    if (r >= col) D[r + static_cast<int64_t>(col) * si.f] = A[r][col];
  }
  if (threadIdx.x == 0 && bad != INT_MAX) atomicMin(c.pivot, bad);
}

// ----------------------------------------------------------------- TRSM
// X = A * L_kk^{-T} for a 64-row tile below the diagonal block; one row per lane.
__global__ void __launch_bounds__(64) k_trsm(CholDev c, const int* __restrict__ task_sn,
                                             const int* __restrict__ task_tile, int step)
{
  const SnInfo si = sninfo(c, task_sn[blockIdx.x]);
  const int k0 = step * NB;
  const int b = min(NB, si.nc - k0);
  const int r = k0 + b + task_tile[blockIdx.x] * 64 + threadIdx.x;
  __shared__ double Lk[NB][NB + 1];
  const double* D = si.Ls + static_cast<int64_t>(k0) * si.f + k0;
  for (int idx = threadIdx.x; idx < NB * NB; idx += 64) {
    int col = idx / NB, rr = idx % NB;
    double v;
    if (rr < b && col < b) v = (rr >= col) ? D[rr + static_cast<int64_t>(col) * si.f] : 0.0;
    else v = (rr == col) ? 1.0 : 0.0;
    Lk[rr][col] = v;
  }
  __syncthreads();
  if (r >= si.f) return;
  double* row = si.Ls + static_cast<int64_t>(k0) * si.f + r;   // row[col*f]
  double x[NB];
#pragma unroll
  for (int col = 0; col < NB; col++) x[col] = (col < b) ? row[static_cast<int64_t>(col) * si.f] : 0.0;
#pragma unroll
  for (int col = 0; col < NB; col++) {
    double t = x[col];
#pragma unroll
    for (int l = 0; l < col; l++) t -= x[l] * Lk[col][l];
    x[col] = t / Lk[col][col];
  }
#pragma unroll
  for (int col = 0; col < NB; col++)
    if (col < b) row[static_cast<int64_t>(col) * si.f] = x[col];
}

// ----------------------------------------------------------------- GEMM tiles
// acc[4][4] += sum_k A(rowsA, k) * B(rowsB, k) over a K chunk staged in LDS.
// Thread (ty, tx) owns rows ty + 16 i and columns tx + 16 j of the 64x64 tile.
__device__ __forceinline__ void tile_mac(const double (*As)[TS + 1], const double (*Bs)[TS + 1],
                                         int kc, int ty, int tx, double acc[4][4])
{
  for (int k = 0; k < kc; k++) {
    double a[4], bb[4];
#pragma unroll
    for (int i = 0; i < 4; i++) a[i] = As[k][ty + 16 * i];
#pragma unroll
    for (int j = 0; j < 4; j++) bb[j] = Bs[k][tx + 16 * j];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
  }
}

// Stage rows [r0, r0+64) x columns [k0, k0+kc) of a column-major matrix with
// leading dimension ld into S[k][row]; rows beyond rmax are zero.
__device__ __forceinline__ void stage(double (*S)[TS + 1], const double* M, int64_t ld, int r0,
                                      int rmax, int k0, int kc)
{
  for (int idx = threadIdx.x; idx < TS * kc; idx += blockDim.x) {
    int k = idx / TS, r = idx % TS;
    int rr = r0 + r;
    S[k][r] = (rr < rmax) ? M[static_cast<int64_t>(k0 + k) * ld + rr] : 0.0;
  }
}

// Trailing update of the panel after block `step`: columns cs.. nc, rows >= col.
__global__ void __launch_bounds__(256) k_panel_update(CholDev c, const int* __restrict__ tsn,
                                                      const int* __restrict__ ttr,
                                                      const int* __restrict__ ttc, int step)
{
  const SnInfo si = sninfo(c, tsn[blockIdx.x]);
  const int k0 = step * NB;
  const int b = min(NB, si.nc - k0);
  const int cs = k0 + b;
  const int R0 = cs + ttr[blockIdx.x] * TS;
  const int C0 = cs + ttc[blockIdx.x] * TS;
  __shared__ double As[NB][TS + 1];
  __shared__ double Bs[NB][TS + 1];
  stage(As, si.Ls, si.f, R0, si.f, k0, b);
  stage(Bs, si.Ls, si.f, C0, si.nc, k0, b);
  __syncthreads();
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  double acc[4][4] = {};
  tile_mac(As, Bs, b, ty, tx, acc);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int r = R0 + ty + 16 * i, col = C0 + tx + 16 * j;
      if (r < si.f && col < si.nc && r >= col) si.Ls[static_cast<int64_t>(col) * si.f + r] -= acc[i][j];
    }
}

// U(r, col) = -sum_k L21(r,k) L21(col,k), K = nc, for one 64x64 tile.
__global__ void __launch_bounds__(256) k_syrk(CholDev c, const int* __restrict__ tsn,
                                              const int* __restrict__ ttr,
                                              const int* __restrict__ ttc)
{
  const int s = tsn[blockIdx.x];
  const SnInfo si = sninfo(c, s);
  const int R0 = ttr[blockIdx.x] * TS, C0 = ttc[blockIdx.x] * TS;
  const double* L21 = si.Ls + si.nc;   // rows of L21 start at panel row nc
  __shared__ double As[NB][TS + 1];
  __shared__ double Bs[NB][TS + 1];
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  double acc[4][4] = {};
  for (int kk = 0; kk < si.nc; kk += NB) {
    int kc = min(NB, si.nc - kk);
    __syncthreads();
    stage(As, L21, si.f, R0, si.nr, kk, kc);
    stage(Bs, L21, si.f, C0, si.nr, kk, kc);
    __syncthreads();
    tile_mac(As, Bs, kc, ty, tx, acc);
  }
  double* Us = c.U + c.u_off[s];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int r = R0 + ty + 16 * i, col = C0 + tx + 16 * j;
      if (r < si.nr && col < si.nr && r >= col) Us[static_cast<int64_t>(col) * si.nr + r] = -acc[i][j];
    }
}

// ----------------------------------------------------------------- solves
__device__ double block_reduce(double v, double* sh)
{
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Forward: v = [x(cols); 0] + children's update vectors; y = L11^{-1} v(cols);
// u = v(rows) - L21 y.  Working vector in global scratch `w` (f entries).
__global__ void __launch_bounds__(256) k_trsv_fwd(CholDev c, const int* __restrict__ sns,
                                                  double* __restrict__ x,
                                                  double* __restrict__ ubuf,
                                                  const int64_t* __restrict__ uoff,
                                                  double* __restrict__ wbuf, int max_front)
{
  const int s = sns[blockIdx.x];
  const SnInfo si = sninfo(c, s);
  double* v = wbuf + static_cast<int64_t>(blockIdx.x) * max_front;
  for (int i = threadIdx.x; i < si.f; i += blockDim.x) v[i] = (i < si.nc) ? x[si.a + i] : 0.0;
  __syncthreads();
  for (int t = c.child_ptr[s]; t < c.child_ptr[s + 1]; t++) {
    const int ch = c.child_list[t];
    const int nrc = c.rows_ptr[ch + 1] - c.rows_ptr[ch];
    const int* relc = c.rel + c.rows_ptr[ch];
    const double* uc = ubuf + uoff[ch];
    for (int k = threadIdx.x; k < nrc; k += blockDim.x) v[relc[k]] += uc[k];
    __syncthreads();
  }
  for (int j = 0; j < si.nc; j++) {
    if (threadIdx.x == 0) v[j] = v[j] / si.Ls[static_cast<int64_t>(j) * si.f + j];
    __syncthreads();
    const double yj = v[j];
    const double* col = si.Ls + static_cast<int64_t>(j) * si.f;
    for (int r = j + 1 + threadIdx.x; r < si.f; r += blockDim.x) v[r] -= col[r] * yj;
    __syncthreads();
  }
  for (int i = threadIdx.x; i < si.nc; i += blockDim.x) x[si.a + i] = v[i];
  for (int i = threadIdx.x; i < si.nr; i += blockDim.x) ubuf[uoff[s] + i] = v[si.nc + i];
}

// Backward: x(cols) = L11^{-T} (y(cols) - L21^T x(rows)).
__global__ void __launch_bounds__(256) k_trsv_bwd(CholDev c, const int* __restrict__ sns,
                                                  double* __restrict__ x,
                                                  double* __restrict__ wbuf, int max_front)
{
  const int s = sns[blockIdx.x];
  const SnInfo si = sninfo(c, s);
  __shared__ double sh[256];
  double* v = wbuf + static_cast<int64_t>(blockIdx.x) * max_front;
  const int* rows = c.rows + c.rows_ptr[s];
  // xr = x(rows) staged after the nc entries
  for (int i = threadIdx.x; i < si.nr; i += blockDim.x) v[si.nc + i] = x[rows[i]];
  for (int i = threadIdx.x; i < si.nc; i += blockDim.x) v[i] = x[si.a + i];
  __syncthreads();
  for (int j = 0; j < si.nc; j++) {
    const double* col = si.Ls + static_cast<int64_t>(j) * si.f + si.nc;
    double t = 0.0;
    for (int k = threadIdx.x; k < si.nr; k += blockDim.x) t += col[k] * v[si.nc + k];
    t = block_reduce(t, sh);
    if (threadIdx.x == 0) v[j] -= t;
  }
  __syncthreads();
  for (int j = si.nc - 1; j >= 0; j--) {
    const double* col = si.Ls + static_cast<int64_t>(j) * si.f;
    double t = 0.0;
    for (int i = j + 1 + threadIdx.x; i < si.nc; i += blockDim.x) t += col[i] * v[i];
    t = block_reduce(t, sh);
    if (threadIdx.x == 0) v[j] = (v[j] - t) / col[j];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < si.nc; i += blockDim.x) x[si.a + i] = v[i];
}

}  // namespace

// =================================================================== launchers
void extend_add(const CholDev& c, int ntasks, const int* task_p, const int* task_col,
                const int* cptr, const int* cc, const int* ccol, hipStream_t st)
{
  if (ntasks <= 0) return;
  k_extend_add<<<ntasks, 256, 0, st>>>(c, task_p, task_col, cptr, cc, ccol);
  HIPCHK(hipGetLastError());
}

void potrf_block(const CholDev& c, int nsn, const int* sns, int step, hipStream_t st)
{
  if (nsn <= 0) return;
  k_potrf<<<nsn, 256, 0, st>>>(c, sns, step);
  HIPCHK(hipGetLastError());
}

void trsm_block(const CholDev& c, int ntasks, const int* tsn, const int* ttile, int step,
                hipStream_t st)
{
  if (ntasks <= 0) return;
  k_trsm<<<ntasks, 64, 0, st>>>(c, tsn, ttile, step);
  HIPCHK(hipGetLastError());
}

void panel_update(const CholDev& c, int ntasks, const int* tsn, const int* ttr, const int* ttc,
                  int step, hipStream_t st)
{
  if (ntasks <= 0) return;
  k_panel_update<<<ntasks, 256, 0, st>>>(c, tsn, ttr, ttc, step);
  HIPCHK(hipGetLastError());
}

void syrk_update(const CholDev& c, int ntasks, const int* tsn, const int* ttr, const int* ttc,
                 hipStream_t st)
{
  if (ntasks <= 0) return;
  k_syrk<<<ntasks, 256, 0, st>>>(c, tsn, ttr, ttc);
  HIPCHK(hipGetLastError());
}

void trsv_forward(const CholDev& c, int nsn, const int* sns, double* x, double* ubuf,
                  const int64_t* uoff, double* wbuf, int max_front, hipStream_t st)
{
  if (nsn <= 0) return;
  k_trsv_fwd<<<nsn, 256, 0, st>>>(c, sns, x, ubuf, uoff, wbuf, max_front);
  HIPCHK(hipGetLastError());
}

void trsv_backward(const CholDev& c, int nsn, const int* sns, double* x, double* wbuf,
                   int max_front, hipStream_t st)
{
  if (nsn <= 0) return;
  k_trsv_bwd<<<nsn, 256, 0, st>>>(c, sns, x, wbuf, max_front);
  HIPCHK(hipGetLastError());
}

}  // namespace dev
}  // namespace hykkt
