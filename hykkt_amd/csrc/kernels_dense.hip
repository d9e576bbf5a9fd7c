// kernels_dense.hip -- dense FP64 kernels of the supernodal Cholesky on
// gfx950: 128x128-tile NT-GEMM on the matrix cores (v_mfma_f64_16x16x4f64),
// and the factorization + inversion of 128x128 diagonal blocks.
//
// MFMA f64 16x16x4 operand layout on gfx950 (lane l, 64 lanes):
//   A operand: A[i = l%16][k = l/16];  B operand: B[k = l/16][j = l%16];
//   result element v (0..3): D[i = 4v + l/16][j = l%16].
// The GEMM computes C(r, c) = beta*C(r, c) + alpha * sum_k A(r, k) * B(c, k)
// with A, B, C column-major.  It feeds the MFMA with B's rows as the A
// operand and A's rows as the B operand, so each result register holds 16
// consecutive rows of one C column (coalesced column-major stores).
#include <hip/hip_runtime.h>

#include <climits>

#include "kernels.hpp"

namespace hykkt {
namespace dev {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BT = 128;        // tile edge (rows and columns of C)
constexpr int KC = 16;         // K chunk staged in LDS
constexpr int LDA_S = BT + 2;  // LDS row length (doubles)

struct GemmArgs {
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  double* C;
  int64_t ldc;
  int M, N, K;       // C is M x N (tile-local bounds), K inner
  double alpha, beta;
};

// One 128x128 tile of C = beta*C + alpha*A*B^T.  256 threads: 4 waves in a
// 2x2 arrangement, each owning a 64x64 quadrant = 4x4 MFMA blocks.
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, double* smem)
{
  double* As = smem;                       // [KC][LDA_S]
  double* Bs = smem + KC * LDA_S;          // [KC][LDA_S]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  // staging assignment: thread t loads k = t/16, rows (t%16)*8 .. +8
  const int sk = t >> 4, sr = (t & 15) * 8;
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    const int kk = k0 + sk;
    double ra[8], rb[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int r = sr + q;
      ra[q] = (kk < g.K && r < g.M) ? g.A[kk * g.lda + r] : 0.0;
      rb[q] = (kk < g.K && r < g.N) ? g.B[kk * g.ldb + r] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) {
      As[sk * LDA_S + sr + q] = ra[q];
      Bs[sk * LDA_S + sr + q] = rb[q];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      double fa[4], fb[4];
#pragma unroll
      for (int i = 0; i < 4; i++) fa[i] = As[kr * LDA_S + wr + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; j++) fb[j] = Bs[kr * LDA_S + wc + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i], acc[i][j], 0, 0, 0);
    }
  }
  // acc[i][j][v]: C row = wr + 16 i + lane%16, C col = wc + 16 j + 4 v + lane/16
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = wr + i * 16 + (lane & 15);
    if (r >= g.M) continue;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = wc + j * 16 + 4 * v + (lane >> 4);
        if (c < g.N) {
          double* p = g.C + c * g.ldc + r;
          double val = g.alpha * acc[i][j][v];
          *p = (g.beta == 0.0) ? val : g.beta * *p + val;
        }
      }
  }
}

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

// mode 0: trailing panel update after block `step`   C = C - A A^T (K = block)
// mode 1: update matrix U = -L21 L21^T                 (K = nc, U overwritten)
// mode 2: TRSM L21(block) = A21 * Dinv^T               (K = block, in place)
__global__ void __launch_bounds__(256) k_dense_nt(CholDev c, const int* __restrict__ tsn,
                                                  const int* __restrict__ ttr,
                                                  const int* __restrict__ ttc, int mode, int step)
{
  extern __shared__ double smem[];
  const SnInfo si = sninfo(c, tsn[blockIdx.x]);
  const int tr = ttr[blockIdx.x], tc = ttc[blockIdx.x];
  GemmArgs g;
  if (mode == 1) {
    const double* L21 = si.Ls + si.nc;
    const int R0 = tr * BT, C0 = tc * BT;
    g.A = L21 + R0;
    g.B = L21 + C0;
    g.lda = g.ldb = si.f;
    g.C = c.U + c.u_off[tsn[blockIdx.x]] + static_cast<int64_t>(C0) * si.nr + R0;
    g.ldc = si.nr;
    g.M = min(BT, si.nr - R0);
    g.N = min(BT, si.nr - C0);
    g.K = si.nc;
    g.alpha = -1.0;
    g.beta = 0.0;
  } else {
    const int k0 = step * BT;
    const int b = min(BT, si.nc - k0);
    const int cs = k0 + b;
    double* panel = si.Ls + static_cast<int64_t>(k0) * si.f;   // columns k0..k0+b
    if (mode == 0) {
      const int R0 = cs + tr * BT, C0 = cs + tc * BT;
      g.A = panel + R0;
      g.B = panel + C0;
      g.lda = g.ldb = si.f;
      g.C = si.Ls + static_cast<int64_t>(C0) * si.f + R0;
      g.ldc = si.f;
      g.M = min(BT, si.f - R0);
      g.N = min(BT, si.nc - C0);
      g.K = b;
      g.alpha = -1.0;
      g.beta = 1.0;
    } else {
      const int R0 = cs + tr * BT;
      g.A = panel + R0;
      g.lda = si.f;
      g.B = c.dinv + c.dinv_off[tsn[blockIdx.x]] + static_cast<int64_t>(step) * BT * BT;
      g.ldb = BT;
      g.C = panel + R0;
      g.ldc = si.f;
      g.M = min(BT, si.f - R0);
      g.N = b;
      g.K = b;
      g.alpha = 1.0;
      g.beta = 0.0;
    }
  }
  if (g.M <= 0 || g.N <= 0) return;
  gemm_tile(g, smem);
}

// Factor the diagonal block of step `step` (b <= 128) in LDS, write L_kk back
// and its inverse (lower triangular, zero-padded to 128 x 128, column-major)
// to dinv.  LDS holds one 128 x 129 array P: L(r, c) (r >= c) at P[r][c]; the
// inverse X(i, j) (i > j) at P[j][i] and X(j, j) at P[j][128].
__global__ void __launch_bounds__(256) k_potrf_inv(CholDev c, const int* __restrict__ sns,
                                                   int step)
{
  extern __shared__ double smem[];
  const int s = sns[blockIdx.x];
  const SnInfo si = sninfo(c, s);
  const int k0 = step * BT;
  if (k0 >= si.nc) return;
  const int b = min(BT, si.nc - k0);
  double (*P)[BT + 1] = reinterpret_cast<double (*)[BT + 1]>(smem);
  __shared__ int bad;
  double* D = si.Ls + static_cast<int64_t>(k0) * si.f + k0;
  for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
    int col = idx / b, r = idx % b;
    if (r >= col) P[r][col] = D[r + static_cast<int64_t>(col) * si.f];
  }
  if (threadIdx.x == 0) bad = INT_MAX;
  __syncthreads();
  // right-looking unblocked Cholesky of the b x b block
  for (int k = 0; k < b; k++) {
    if (threadIdx.x == 0) {
      double d = P[k][k];
      if (!(d > 0.0) || !(sqrt(d) > c.pivot_tol)) {
        if (bad == INT_MAX) bad = si.a + k0 + k;
        d = 1.0;   // keep the arithmetic finite; the failure is reported
      }
      P[k][k] = sqrt(d);
    }
    __syncthreads();
    const double dk = P[k][k];
    for (int r = k + 1 + threadIdx.x; r < b; r += blockDim.x) P[r][k] /= dk;
    __syncthreads();
    const int m = b - k - 1;
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
      int col = k + 1 + idx / m, r = k + 1 + idx % m;
      if (col <= r) P[r][col] -= P[r][k] * P[col][k];
    }
    __syncthreads();
  }
  // X = L^{-1} row by row: X(i, j) = -(sum_{k<i} L(i,k) X(k,j)) / L(i,i), j < i,
  // thread j owns column j; L(i, k) is read by all threads at once (broadcast).
  const int j = threadIdx.x;
  if (j < b) P[j][BT] = 1.0 / P[j][j];
  __syncthreads();
  for (int i = 1; i < b; i++) {
    if (j < i) {
      double t = P[i][j] * P[j][BT];
      for (int k = j + 1; k < i; k++) t += P[i][k] * P[j][k];
      P[j][i] = -t / P[i][i];
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
    int col = idx / b, r = idx % b;
    if (r >= col) D[r + static_cast<int64_t>(col) * si.f] = P[r][col];
  }
  double* W = c.dinv + c.dinv_off[s] + static_cast<int64_t>(step) * BT * BT;
  for (int idx = threadIdx.x; idx < BT * BT; idx += blockDim.x) {
    int col = idx / BT, r = idx % BT;
    double x = 0.0;
    if (r < b && col < b && r >= col) x = (r == col) ? P[col][BT] : P[col][r];
    W[idx] = x;
  }
  if (threadIdx.x == 0 && bad != INT_MAX) atomicMin(c.pivot, bad);
}

}  // namespace

size_t dense_nt_smem() { return 2 * KC * LDA_S * sizeof(double); }
size_t potrf_inv_smem() { return BT * (BT + 1) * sizeof(double); }

void dense_nt(const CholDev& c, int ntasks, const int* tsn, const int* ttr, const int* ttc,
              int mode, int step, hipStream_t st)
{
  if (ntasks <= 0) return;
  k_dense_nt<<<ntasks, 256, dense_nt_smem(), st>>>(c, tsn, ttr, ttc, mode, step);
  HIPCHK(hipGetLastError());
}

void potrf_inv(const CholDev& c, int nsn, const int* sns, int step, hipStream_t st)
{
  if (nsn <= 0) return;
  static bool attr = false;
  if (!attr) {
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_potrf_inv),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(potrf_inv_smem())));
    attr = true;
  }
  k_potrf_inv<<<nsn, 256, potrf_inv_smem(), st>>>(c, sns, step);
  HIPCHK(hipGetLastError());
}

}  // namespace dev
}  // namespace hykkt
