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

// One 128x128 tile of C = beta*C + alpha*A*B^T with beta in {0, 1} and
// alpha in {1, -1}.  256 threads: 4 waves in a 2x2 arrangement, each owning a
// 64x64 quadrant = 4x4 MFMA blocks.  The accumulators start from beta*C (all
// loads in flight at once) and are stored once at the end; alpha is applied
// exactly to the staged A values.  K is staged through LDS in chunks of KC,
// the next chunk's global loads overlapping the current chunk's MFMAs.
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, double* smem)
{
  double* As = smem;                       // [KC][LDA_S]
  double* Bs = smem + KC * LDA_S;          // [KC][LDA_S]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  d4 acc[4][4];
  // acc[i][j][v]: C row = wr + 16 i + lane%16, C col = wc + 16 j + 4 v + lane/16
  if (g.beta != 0.0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = wr + i * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = wc + j * 16 + 4 * v + (lane >> 4);
          acc[i][j][v] = (r < g.M && c < g.N) ? g.C[c * g.ldc + r] : 0.0;
        }
    }
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  }

  // staging assignment: thread t loads k = t/16, rows (t%16)*8 .. +8
  const int sk = t >> 4, sr = (t & 15) * 8;
  double ra[8], rb[8];
  auto gload = [&](int k0) {
    const int kk = k0 + sk;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int r = sr + q;
      ra[q] = (kk < g.K && r < g.M) ? g.A[kk * g.lda + r] : 0.0;
      rb[q] = (kk < g.K && r < g.N) ? g.B[kk * g.ldb + r] : 0.0;
    }
  };
  gload(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    __syncthreads();   // the previous chunk's fragments have been read
#pragma unroll
    for (int q = 0; q < 8; q++) {
      As[sk * LDA_S + sr + q] = g.alpha * ra[q];   // exact: alpha is +-1
      Bs[sk * LDA_S + sr + q] = rb[q];
    }
    __syncthreads();
    if (k0 + KC < g.K) gload(k0 + KC);
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
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = wr + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = wc + j * 16 + 4 * v + (lane >> 4);
        if (r < g.M && c < g.N) g.C[c * g.ldc + r] = acc[i][j][v];
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
// inverse X(i, j) (i > j) at P[j][i] and X(j, j) at P[j][128]; plus a 3 x 32 x 33
// scratch for the off-diagonal blocks of the inverse.  The block is padded to
// 128 with the identity.
//
// Both steps are blocked by 32.  Cholesky: wave 0 factors the 32x32 diagonal
// block in registers (lane r holds row r; columns move between lanes by
// shuffles), then every thread solves one row below against it, then the
// trailing lower triangle is updated.  Inverse: wave w inverts diagonal block
// w in registers (lane j computes column j), then the block diagonals below
// follow from X(i,j) = -X(i,i) * sum_{k=j}^{i-1} L(i,k) X(k,j).
constexpr int SUB = 32;

// Lane `lane`'s value of v, as a wave-uniform (scalar register) value.
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(u), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(u >> 32), lane);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// BK: padded block edge (32, 64 or 128) chosen per launch from the largest
// diagonal block in it, so small supernodes need little LDS.
template <int BK>
__global__ void __launch_bounds__(256) k_potrf_inv(CholDev c, const int* __restrict__ sns,
                                                   int step)
{
  extern __shared__ double smem[];
  const int s = sns[blockIdx.x];
  const SnInfo si = sninfo(c, s);
  const int k0 = step * BT;
  if (k0 >= si.nc) return;
  const int b = min(BK, si.nc - k0);
  double (*P)[BK + 1] = reinterpret_cast<double (*)[BK + 1]>(smem);
  double (*T)[SUB][SUB + 1] = reinterpret_cast<double (*)[SUB][SUB + 1]>(smem + BK * (BK + 1));
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double* D = si.Ls + static_cast<int64_t>(k0) * si.f + k0;
  // load the block up to 32 elements per thread at a time (all in flight
  // before the LDS stores): element e = col*BK + r, lower part only
  constexpr int NE = BK * BK, PER = NE / 256 < 32 ? NE / 256 : 32, STEP = 256 * PER;
  for (int e0 = tid; e0 < NE; e0 += STEP) {
    double v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = e0 + u * 256, r = e % BK, col = e / BK;
      v[u] = (e < NE && r >= col && r < b) ? D[r + static_cast<int64_t>(col) * si.f]
                                           : (r == col ? 1.0 : 0.0);
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = e0 + u * 256, r = e % BK, col = e / BK;
      if (e < NE && r >= col) P[r][col] = v[u];
    }
  }
  if (tid == 0) bad = INT_MAX;
  __syncthreads();

  // ---- blocked Cholesky ----
  for (int kb = 0; kb < BK; kb += SUB) {
    if (wave == 0) {
      // lane r holds row r; entries above the diagonal are updated too (they
      // are never stored), which keeps the loop free of lane-dependent branches.
      // Column k of L reaches every lane through v_readlane (scalar registers),
      // not through the LDS crossbar.
      const int r = lane & (SUB - 1);
      double v[SUB];
#pragma unroll
      for (int q = 0; q < SUB; q++) v[q] = (q <= r) ? P[kb + r][kb + q] : 0.0;
      int badk = -1;
#pragma unroll
      for (int k = 0; k < SUB; k++) {
        const double dkk = readlane_f64(v[k], k);
        const bool ok = (dkk > 0.0) && (sqrt(dkk) > c.pivot_tol);
        badk = (badk < 0 && !ok && kb + k < b) ? k : badk;
        const double d = sqrt(ok ? dkk : 1.0);   // failure reported, arithmetic kept finite
        v[k] = (r == k) ? d : v[k] / d;
        const double vk = v[k];
#pragma unroll
        for (int q = k + 1; q < SUB; q++) v[q] -= vk * readlane_f64(vk, q);
      }
      if (lane == 0 && badk >= 0 && bad == INT_MAX) bad = si.a + k0 + kb + badk;
      if (lane < SUB) {
#pragma unroll
        for (int q = 0; q < SUB; q++)
          if (q <= r) P[kb + r][kb + q] = v[q];
      }
    }
    __syncthreads();
    const int rbeg = kb + SUB;
    if (rbeg < BK) {
      // rows below: x = a * L_kb^{-T}, one row per thread, row held in registers
      for (int r = rbeg + tid; r < BK; r += blockDim.x) {
        double x[SUB];
#pragma unroll
        for (int q = 0; q < SUB; q++) x[q] = P[r][kb + q];
#pragma unroll
        for (int k = 0; k < SUB; k++) {
          double t0 = x[k], t1 = 0.0;
#pragma unroll
          for (int q = 0; q < k; q++) {
            if (q & 1) t1 -= x[q] * P[kb + k][kb + q];
            else t0 -= x[q] * P[kb + k][kb + q];
          }
          x[k] = (t0 + t1) / P[kb + k][kb + k];
        }
#pragma unroll
        for (int q = 0; q < SUB; q++) P[r][kb + q] = x[q];
      }
      __syncthreads();
      // trailing update P[r][cc] -= sum_k P[r][k] P[cc][k], r >= cc >= rbeg,
      // in 4x4 register tiles (lower tiles only)
      const int mt = (BK - rbeg) / 4;
      for (int e = tid; e < mt * mt; e += blockDim.x) {
        const int ti = e / mt, tj = e % mt;
        if (tj > ti) continue;
        const int r0 = rbeg + 4 * ti, c0 = rbeg + 4 * tj;
        double acc[4][4] = {};
#pragma unroll 4
        for (int k = kb; k < kb + SUB; k++) {
          double a[4], bb[4];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            a[q] = P[r0 + q][k];
            bb[q] = P[c0 + q][k];
          }
#pragma unroll
          for (int q = 0; q < 4; q++)
#pragma unroll
            for (int w = 0; w < 4; w++) acc[q][w] += a[q] * bb[w];
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int w = 0; w < 4; w++)
            if (c0 + w <= r0 + q) P[r0 + q][c0 + w] -= acc[q][w];
      }
      __syncthreads();
    }
  }

  // ---- inverse of L ----
  {
    // wave w inverts diagonal block w; lane j < 32 computes column j
    const int base = wave * SUB, j = lane;
    if (j < SUB && base < BK) {
      double x[SUB];
#pragma unroll
      for (int i = 0; i < SUB; i++) {
        double t0 = (i == j) ? 1.0 : 0.0, t1 = 0.0;
#pragma unroll
        for (int k = 0; k < i; k++) {
          if (k & 1) t1 -= P[base + i][base + k] * x[k];
          else t0 -= P[base + i][base + k] * x[k];
        }
        x[i] = (i >= j) ? (t0 + t1) / P[base + i][base + i] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < SUB; i++) {
        if (i == j) P[base + j][BK] = x[i];
        else if (i > j) P[base + j][base + i] = x[i];
      }
    }
  }
  __syncthreads();
  auto X = [&](int i, int j) -> double { return (i == j) ? P[j][BK] : P[j][i]; };
  // block diagonals: T_q = sum_{k=j..ib-1} L(i,k) X(k,j), then X(i,j) = -X_ii T_q,
  // 4x4 register tiles, (q, tile) spread over the threads
  for (int d = 1; d < BK / SUB; d++) {
    const int nblk = BK / SUB - d;
    constexpr int TPB4 = (SUB / 4) * (SUB / 4);   // 4x4 tiles per 32x32 block
    for (int e = tid; e < nblk * TPB4; e += blockDim.x) {
      const int q = e / TPB4, ti = (e % TPB4) / (SUB / 4), tj = e % (SUB / 4);
      const int ib = (q + d) * SUB, i0 = ib + 4 * ti, j0 = q * SUB + 4 * tj;
      double acc[4][4] = {};
      // X(k, j) = 0 for k < j: start at the tile's first column; the first
      // four k straddle the diagonal of X, beyond them X(k, j) = P[j][k]
#pragma unroll
      for (int k = j0; k < j0 + 4; k++) {
        double a[4], xv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          a[u] = P[i0 + u][k];
          const int j = j0 + u;
          xv[u] = (k > j) ? P[j][k] : ((k == j) ? P[j][BK] : 0.0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int w = 0; w < 4; w++) acc[u][w] += a[u] * xv[w];
      }
      for (int k = j0 + 4; k < ib; k++) {
        double a[4], xv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          a[u] = P[i0 + u][k];
          xv[u] = P[j0 + u][k];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int w = 0; w < 4; w++) acc[u][w] += a[u] * xv[w];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int w = 0; w < 4; w++) T[q][4 * ti + u][4 * tj + w] = acc[u][w];
    }
    __syncthreads();
    for (int e = tid; e < nblk * TPB4; e += blockDim.x) {
      const int q = e / TPB4, ti = (e % TPB4) / (SUB / 4), tj = e % (SUB / 4);
      const int ib = (q + d) * SUB, i0 = ib + 4 * ti, j0 = q * SUB + 4 * tj;
      double acc[4][4] = {};
      // X(i, ib + k) for k below the tile's rows is P[ib + k][i]; the last four
      // k straddle the diagonal of X
      for (int k = 0; k < 4 * ti; k++) {
        double xi[4], tv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          xi[u] = P[ib + k][i0 + u];
          tv[u] = T[q][k][4 * tj + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int w = 0; w < 4; w++) acc[u][w] += xi[u] * tv[w];
      }
#pragma unroll
      for (int k = 4 * ti; k < 4 * ti + 4; k++) {
        double xi[4], tv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int i = i0 + u, kk = ib + k;
          xi[u] = (i > kk) ? P[kk][i] : ((i == kk) ? P[kk][BK] : 0.0);
          tv[u] = T[q][k][4 * tj + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int w = 0; w < 4; w++) acc[u][w] += xi[u] * tv[w];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int w = 0; w < 4; w++) P[j0 + w][i0 + u] = -acc[u][w];
    }
    __syncthreads();
  }

  for (int col = tid / 64; col < b; col += 4)
    for (int r = (tid & 63) + col; r < b; r += 64) D[r + static_cast<int64_t>(col) * si.f] = P[r][col];
  // inverse into the 128 x 128 dinv slot (column-major); entries outside the
  // b x b lower triangle stay at the zeros written when the slot was created
  double* W = c.dinv + c.dinv_off[s] + static_cast<int64_t>(step) * BT * BT;
  for (int col = tid / 64; col < b; col += 4)
    for (int r = (tid & 63) + col; r < b; r += 64) W[static_cast<int64_t>(col) * BT + r] = X(r, col);
  if (tid == 0 && bad != INT_MAX) atomicMin(c.pivot, bad);
}

}  // namespace

size_t dense_nt_smem() { return 2 * KC * LDA_S * sizeof(double); }
void dense_nt(const CholDev& c, int ntasks, const int* tsn, const int* ttr, const int* ttc,
              int mode, int step, hipStream_t st)
{
  if (ntasks <= 0) return;
  k_dense_nt<<<ntasks, 256, dense_nt_smem(), st>>>(c, tsn, ttr, ttc, mode, step);
  HIPCHK(hipGetLastError());
}

template <int BK>
static size_t potrf_smem()
{
  return (BK * (BK + 1) + (BK / SUB - 1) * SUB * (SUB + 1)) * sizeof(double);
}

template <int BK>
static void launch_potrf(const CholDev& c, int nsn, const int* sns, int step, hipStream_t st)
{
  static bool attr = false;
  if (!attr) {
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_potrf_inv<BK>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(potrf_smem<BK>())));
    attr = true;
  }
  k_potrf_inv<BK><<<nsn, 256, potrf_smem<BK>(), st>>>(c, sns, step);
  HIPCHK(hipGetLastError());
}

void potrf_inv(const CholDev& c, int nsn, const int* sns, int step, int max_b, hipStream_t st)
{
  if (nsn <= 0) return;
  if (max_b <= 32) launch_potrf<32>(c, nsn, sns, step, st);
  else if (max_b <= 64) launch_potrf<64>(c, nsn, sns, step, st);
  else launch_potrf<128>(c, nsn, sns, step, st);
}

}  // namespace dev
}  // namespace hykkt
