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
  // pk > 0: C is the pk x pk lower triangle packed by columns and the tile
  // sits at (pr0, pc0) in it; entries above the diagonal are not stored
  int64_t pk = 0, pr0 = 0, pc0 = 0;
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
  if (g.pk > 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = wr + i * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = wc + j * 16 + 4 * v + (lane >> 4);
          const int64_t gr = g.pr0 + r, gc = g.pc0 + c;
          if (r < g.M && c < g.N && gr >= gc) g.C[packed_lower(gr, gc, g.pk)] = acc[i][j][v];
        }
    }
    return;
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
__global__ void __launch_bounds__(256, 2) k_dense_nt(CholDev c, const int* __restrict__ tsn,
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
    g.C = c.U + c.u_off[tsn[blockIdx.x]];   // packed lower triangle of U
    g.ldc = si.nr;
    g.pk = si.nr;
    g.pr0 = R0;
    g.pc0 = C0;
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
// to dinv.  LDS holds one BK x (BK+1) array P: L(r, c) (r >= c) at P[r][c];
// the inverse X(i, j) (i > j) at P[j][i] and X(j, j) at P[j][BK]; the
// reciprocals of L's diagonal; and 32 x 33 scratch blocks for the inverse's
// off-diagonal products (one per block pair of a block diagonal).  The block is padded to BK with the
// identity.
//
// Both steps are blocked by 32.  Cholesky, per 32-column panel: wave 0
// factors the 32x32 diagonal block in registers (lane r holds row r; column
// k reaches all lanes through v_readlane), every thread solves one row below
// against it (multiplying by reciprocals, no divisions on the chain), and the
// trailing lower part is updated with FP64 MFMA 16x16x4 tiles.  Inverse: wave
// w inverts diagonal block w (lane j computes column j), then the block
// diagonals below follow from X(i,j) = -X(i,i) * sum_{k=j}^{i-1} L(i,k) X(k,j),
// both products on the matrix cores.
//
// MFMA f64 16x16x4 (lane l of 64): A operand A(l%16, l/16), B operand
// B(l/16, l%16), result element v = D(4v + l/16, l%16) of D = A B.
constexpr int SUB = 32;

// Lane `lane`'s value of v, as a wave-uniform (scalar register) value.
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(u), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(u >> 32), lane);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// One 16x16 tile of D = A B^T-style products accumulated over K (a multiple
// of 4): a_at(i, k) and b_at(k, j) give the operands.
template <class FA, class FB>
__device__ __forceinline__ d4 mfma16(int K, int lane, FA a_at, FB b_at)
{
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  const int i = lane & 15, kq = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a_at(i, k0 + kq), b_at(k0 + kq, i), acc, 0, 0, 0);
  return acc;
}

// A 32x32 block of D = A B as four 16x16 tiles with independent accumulators
// (interleaved, so the matrix core pipelines them): acc[2*ti + tj] holds
// tile (ti, tj), element v at D(16 ti + 4v + lane/16, 16 tj + lane%16).
template <class FA, class FB>
__device__ __forceinline__ void mfma32(int K, int lane, FA a_at, FB b_at, d4 (&acc)[4])
{
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
  const int i = lane & 15, kq = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += 4) {
    const double a0 = a_at(i, k0 + kq), a1 = a_at(16 + i, k0 + kq);
    const double b0 = b_at(k0 + kq, i), b1 = b_at(k0 + kq, 16 + i);
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[3], 0, 0, 0);
  }
}

// BK: padded block edge (32, 64 or 128) chosen per launch from the largest
// diagonal block in it, so small supernodes need little LDS.
template <int BK>
__global__ void __launch_bounds__(256) k_potrf_inv(CholDev c, const int* __restrict__ sns,
                                                   int step)
{
  constexpr int NB = BK / SUB;   // 32-wide panels
  extern __shared__ double smem[];
  const int s = sns[blockIdx.x];
  const SnInfo si = sninfo(c, s);
  const int k0 = step * BT;
  if (k0 >= si.nc) return;
  const int b = min(BK, si.nc - k0);
  double (*P)[BK + 1] = reinterpret_cast<double (*)[BK + 1]>(smem);
  double* rd = smem + BK * (BK + 1);                                   // 1 / L(r, r)
  double (*T)[SUB][SUB + 1] = reinterpret_cast<double (*)[SUB][SUB + 1]>(rd + BK);
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double* D = si.Ls + static_cast<int64_t>(k0) * si.f + k0;
  // load the block, up to 32 elements per thread in flight: element
  // e = col*BK + r, lower part only, padded with the identity
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

  // X(i, j) for i >= j from its storage; 0 above the diagonal
  auto X = [&](int i, int j) -> double { return (i > j) ? P[j][i] : ((i == j) ? P[j][BK] : 0.0); };
  const double tol2 = c.pivot_tol * c.pivot_tol;

  // ---- blocked Cholesky; the diagonal blocks of the inverse on the way ----
#pragma unroll 1
  for (int p = 0; p < NB; p++) {
    const int kb = p * SUB;
    if (wave == 0) {
      // lane r holds row r of the diagonal block; entries above the diagonal
      // are carried along (never stored), so no lane-dependent branches.
      // Step k: pivot from lane k, 1/sqrt from v_rsq_f64 refined by two
      // Newton steps, column k to every lane through v_readlane.
      const int r = lane & (SUB - 1);
      double v[SUB];
#pragma unroll
      for (int q = 0; q < SUB; q++) v[q] = P[kb + r][kb + q];
      int badk = -1;
      double myinv = 1.0;   // 1 / L(r, r), taken at step k == r (no dynamic register index)
#pragma unroll
      for (int k = 0; k < SUB; k++) {
        const double dkk = readlane_f64(v[k], k);
        const bool ok = (dkk > 0.0) && (dkk > tol2);   // sqrt(dkk) > pivot_tol
        badk = (badk < 0 && !ok && kb + k < b) ? k : badk;
        const double xk = ok ? dkk : 1.0;   // failure reported, arithmetic kept finite
        double y = __builtin_amdgcn_rsq(xk);
        const double hx = 0.5 * xk;
        y = y * fma(-hx, y * y, 1.5);
        y = y * fma(-hx, y * y, 1.5);
        v[k] = (r == k) ? xk * y : v[k] * y;
        myinv = (r == k) ? y : myinv;
        const double vk = v[k];
#pragma unroll
        for (int q = k + 1; q < SUB; q++) v[q] -= vk * readlane_f64(vk, q);
      }
      if (lane == 0 && badk >= 0) bad = si.a + k0 + kb + badk;
      if (lane < SUB) {
#pragma unroll
        for (int q = 0; q < SUB; q++)
          if (q <= r) P[kb + r][kb + q] = v[q];
        rd[kb + r] = myinv;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // inverse of this diagonal block: lane j < 32 computes column j;
      // X(i, j) (i > j) goes to P[j][i], X(j, j) to P[j][BK]
      if (lane < SUB) {
        const int j = lane;
        double x[SUB];
#pragma unroll
        for (int i = 0; i < SUB; i++) {
          double t0 = (i == j) ? 1.0 : 0.0, t1 = 0.0;
#pragma unroll
          for (int k = 0; k < i; k++) {
            if (k & 1) t1 -= P[kb + i][kb + k] * x[k];
            else t0 -= P[kb + i][kb + k] * x[k];
          }
          x[i] = (i >= j) ? (t0 + t1) * rd[kb + i] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < SUB; i++) {
          if (i == j) P[kb + j][BK] = x[i];
          else if (i > j) P[kb + j][kb + i] = x[i];
        }
      }
    }
    __syncthreads();
    const int rbeg = kb + SUB;
    if (rbeg < BK) {
      // rows below: L(r, panel) = A(r, panel) X_pp^T on the matrix cores,
      // 16-row x 16-column tiles; all reads before the in-place writes
      const int nrt = (BK - rbeg) / 16, ntl = 2 * nrt;
      d4 acc[3];
#pragma unroll
      for (int u = 0; u < 3; u++) {
        const int t = wave + 4 * u;
        if (t < ntl) {
          const int r0 = rbeg + 16 * (t >> 1), c0 = 16 * (t & 1);
          acc[u] = mfma16(SUB, lane, [&](int i, int k) { return P[r0 + i][kb + k]; },
                          [&](int k, int j) { return X(kb + c0 + j, kb + k); });
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 3; u++) {
        const int t = wave + 4 * u;
        if (t < ntl) {
          const int r0 = rbeg + 16 * (t >> 1), c0 = 16 * (t & 1);
#pragma unroll
          for (int v = 0; v < 4; v++) P[r0 + 4 * v + (lane >> 4)][kb + c0 + (lane & 15)] = acc[u][v];
        }
      }
      __syncthreads();
      // trailing update P[r][cc] -= sum_k P[r][k] P[cc][k] over the panel,
      // r >= cc >= rbeg, in 16x16 MFMA tiles (lower tiles only; the upper
      // half of diagonal tiles is scratch)
      const int mt = (BK - rbeg) / 16;
      for (int e = wave; e < mt * (mt + 1) / 2; e += 4) {
        int ti = 0;
        while ((ti + 1) * (ti + 2) / 2 <= e) ti++;
        const int tj = e - ti * (ti + 1) / 2;
        const int r0 = rbeg + 16 * ti, c0 = rbeg + 16 * tj;
        const d4 a = mfma16(SUB, lane, [&](int i, int k) { return P[r0 + i][kb + k]; },
                            [&](int k, int j) { return P[c0 + j][kb + k]; });
#pragma unroll
        for (int v = 0; v < 4; v++) P[r0 + 4 * v + (lane >> 4)][c0 + (lane & 15)] -= a[v];
      }
      __syncthreads();
    }
  }

  // block diagonals d = 1 .. NB-1: pair q (rows ib = (q+d)*32, cols jb = q*32) on wave q
#pragma unroll 1
  for (int d = 1; d < NB; d++) {
    const int q = wave;
    const bool act = q < NB - d;
    const int ib = (q + d) * SUB, jb = q * SUB;
    d4 acc[4];
    if (act) {
      // T = sum_{k = jb}^{ib-1} L(ib.., k) X(k, jb..)   (32 x 32, K = 32 d)
      mfma32(SUB * d, lane, [&](int i, int k) { return P[ib + i][jb + k]; },
             [&](int k, int j) { return X(jb + k, jb + j); }, acc);
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int v = 0; v < 4; v++)
          T[q][16 * (t >> 1) + 4 * v + (lane >> 4)][16 * (t & 1) + (lane & 15)] = acc[t][v];
    }
    __syncthreads();
    if (act) {
      // X(ib.., jb..) = -X(ib.., ib..) T
      mfma32(SUB, lane, [&](int i, int k) { return X(ib + i, ib + k); },
             [&](int k, int j) { return T[q][k][j]; }, acc);
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int v = 0; v < 4; v++)
          P[jb + 16 * (t & 1) + (lane & 15)][ib + 16 * (t >> 1) + 4 * v + (lane >> 4)] = -acc[t][v];
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
  // P, the reciprocals, and one 32 x 33 product block per off-diagonal pair
  // of the first block diagonal (at most NB-1 pairs at a time)
  return (BK * (BK + 1) + BK + (BK / SUB - 1) * SUB * (SUB + 1)) * sizeof(double);
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
