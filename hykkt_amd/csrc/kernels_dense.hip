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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "batch_dev.hpp"
#include "kernels.hpp"

namespace hykkt {
namespace dev {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BT = 128;        // tile edge (rows and columns of C)
constexpr int KC = 16;         // K chunk staged in LDS
constexpr int LDA_S = BT + 2;  // LDS row length (doubles)

// Global loads and stores of factor data.  COH: device-scope relaxed atomics
// (coherent across the XCDs' L2 caches), for data another workgroup of the
// same launch produced or will consume.
template <bool COH>
__device__ __forceinline__ double ldg(const double* p)
{
  if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}
template <bool COH>
__device__ __forceinline__ void stg(double* p, double v)
{
  if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

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

// One TM x TN tile of C = beta*C + alpha*A*B^T with beta in {0, 1} and
// alpha in {1, -1}.  256 threads: 4 waves in a WM x WN arrangement (2x2 for
// 128x128 tiles, 1x4 for 32x128), each owning a QM x QN block of IM x JN
// MFMA tiles.  The accumulators start from beta*C (all loads in flight at
// once) and are stored once at the end; alpha is applied exactly to the
// staged A values.  K is staged through LDS in chunks of KC, the next chunk's
// global loads overlapping the current chunk's MFMAs.
template <int TM, int TN, bool COH = false>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, double* smem)
{
  constexpr int WN = TM == TN ? 2 : 4, WM = 4 / WN;
  constexpr int QM = TM / WM, QN = TN / WN, IM = QM / 16, JN = QN / 16;
  constexpr int LA = TM + 2, LB = TN + 2;          // LDS row lengths (doubles)
  constexpr int EA = TM / 16, EB = TN / 16;        // staged elements per thread
  double* As = smem;                               // [KC][LA]
  double* Bs = smem + KC * LA;                     // [KC][LB]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = (wave / WN) * QM, wc = (wave % WN) * QN;
  d4 acc[IM][JN];
  // acc[i][j][v]: C row = wr + 16 i + lane%16, C col = wc + 16 j + 4 v + lane/16
  if (g.beta != 0.0) {
#pragma unroll
    for (int i = 0; i < IM; i++) {
      const int r = wr + i * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = wc + j * 16 + 4 * v + (lane >> 4);
          acc[i][j][v] = (r < g.M && c < g.N) ? ldg<COH>(g.C + c * g.ldc + r) : 0.0;
        }
    }
  } else {
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < JN; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  }

  // staging assignment: thread t loads k = t/16, rows t%16 + 16 q (q < E): the
  // 16 lanes of a k write 16 consecutive LDS words (no bank conflicts)
  const int sk = t >> 4, sa = t & 15, sb = t & 15;
  double ra[EA], rb[EB];
  auto gload = [&](int k0) {
    const int kk = k0 + sk;
#pragma unroll
    for (int q = 0; q < EA; q++) {
      const int r = sa + 16 * q;
      ra[q] = (kk < g.K && r < g.M) ? ldg<COH>(g.A + kk * g.lda + r) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < EB; q++) {
      const int r = sb + 16 * q;
      rb[q] = (kk < g.K && r < g.N) ? ldg<COH>(g.B + kk * g.ldb + r) : 0.0;
    }
  };
  gload(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    __syncthreads();   // the previous chunk's fragments have been read
#pragma unroll
    for (int q = 0; q < EA; q++) As[sk * LA + sa + 16 * q] = g.alpha * ra[q];   // exact: alpha is +-1
#pragma unroll
    for (int q = 0; q < EB; q++) Bs[sk * LB + sb + 16 * q] = rb[q];
    __syncthreads();
    if (k0 + KC < g.K) gload(k0 + KC);
#pragma unroll
    for (int ks = 0; ks < KC / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      double fa[IM], fb[JN];
#pragma unroll
      for (int i = 0; i < IM; i++) fa[i] = As[kr * LA + wr + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < JN; j++) fb[j] = Bs[kr * LB + wc + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < IM; i++)
#pragma unroll
        for (int j = 0; j < JN; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i], acc[i][j], 0, 0, 0);
    }
  }
  if (g.pk > 0) {
#pragma unroll
    for (int i = 0; i < IM; i++) {
      const int r = wr + i * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = wc + j * 16 + 4 * v + (lane >> 4);
          const int64_t gr = g.pr0 + r, gc = g.pc0 + c;
          if (r < g.M && c < g.N && gr >= gc) stg<COH>(g.C + packed_lower(gr, gc, g.pk), acc[i][j][v]);
        }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < IM; i++) {
    const int r = wr + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < JN; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = wc + j * 16 + 4 * v + (lane >> 4);
        if (r < g.M && c < g.N) stg<COH>(g.C + c * g.ldc + r, acc[i][j][v]);
      }
  }
}

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

// mode 0: trailing panel update after block `step`   C = C - A A^T (K = block)
// mode 1: update matrix U = -L21 L21^T                 (K = nc, U overwritten)
// mode 2: TRSM L21(block) = A21 * Dinv^T               (K = block, in place)
// mode m >= 3: as mode 0 after blocks `step` .. `step`+m-2 (K = m-1 blocks)
// TM = 128: 128x128 tiles (all modes).  TM = 32: 32-row x 128-column tiles
// (modes 0 and 2), four times as many workgroups for the short launches on
// the critical path (the TRSM and the next-column update of a step with few
// tiles).  System blockIdx.y.
template <int TM>
__global__ void __launch_bounds__(256, 2) k_dense_nt(CholDev c, const int* __restrict__ tsn,
                                                  const int* __restrict__ ttr,
                                                  const int* __restrict__ ttc, int mode, int step,
                                                  Batch bt)
{
  // the 32-row launches are the factorization's critical chain: their waves
  // take issue priority over the trailing updates they share SIMDs with
  if constexpr (TM == 32) __builtin_amdgcn_s_setprio(3);
  extern __shared__ double smem[];
  const int sy = batch_sys(bt);
  if (sy < 0) return;
  const SnInfo si = sninfo(c, batch_at(c.L, sy, bt), tsn[blockIdx.x]);
  const int tr = ttr[blockIdx.x], tc = ttc[blockIdx.x];
  GemmArgs g;
  if (mode == 1) {
    const double* L21 = si.Ls + si.nc;
    const int R0 = tr * BT, C0 = tc * BT;
    g.A = L21 + R0;
    g.B = L21 + C0;
    g.lda = g.ldb = si.f;
    g.C = batch_at(c.U, sy, bt) + c.u_off[tsn[blockIdx.x]];   // packed lower triangle of U
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
    // modes 3 and up: the update by blocks step .. step + mode - 2 together
    const int b = min(mode >= 3 ? (mode - 1) * BT : BT, si.nc - k0);
    const int cs = k0 + b;
    double* panel = si.Ls + static_cast<int64_t>(k0) * si.f;   // columns k0..k0+b
    if (mode != 2) {
      const int R0 = cs + tr * TM, C0 = cs + tc * BT;
      g.A = panel + R0;
      g.B = panel + C0;
      g.lda = g.ldb = si.f;
      g.C = si.Ls + static_cast<int64_t>(C0) * si.f + R0;
      g.ldc = si.f;
      g.M = min(TM, si.f - R0);
      g.N = min(BT, si.nc - C0);
      g.K = b;
      g.alpha = -1.0;
      g.beta = 1.0;
    } else {
      const int R0 = cs + tr * TM;
      g.A = panel + R0;
      g.lda = si.f;
      g.B = batch_at(c.dinv, sy, bt) + c.dinv_off[tsn[blockIdx.x]] +
            static_cast<int64_t>(step) * BT * BT;
      g.ldb = BT;
      g.C = panel + R0;
      g.ldc = si.f;
      g.M = min(TM, si.f - R0);
      g.N = b;
      g.K = b;
      g.alpha = 1.0;
      g.beta = 0.0;
    }
  }
  if (g.M <= 0 || g.N <= 0) return;
  gemm_tile<TM, BT>(g, smem);
}

// Factor the diagonal block of step `step` (b <= 128) in LDS, write L_kk back
// and its inverse (lower triangular, zero-padded to 128 x 128, column-major)
// to dinv.  LDS holds one BK x (BK+1) array P, padded to BK with the
// identity: L(r, c) (r >= c) at P[r][c], the inverse X(i, j) (i > j) at
// P[j][i] and X(j, j) at P[j][BK]; and the reciprocals of L's diagonal.
//
// Cholesky, blocked by 16-column panels: wave 0 factors the whole panel in
// registers (two rows per lane; the pivot row's entries reach every lane
// through v_readlane; 1/sqrt of the pivot is v_rsq_f64 refined by two Newton
// steps, so the chain carries no division or square root); the trailing
// lower part takes the rank-16 update as FP64 MFMA 16x16x4 tiles spread over
// the four waves.  Inverse: the 16x16 diagonal
// blocks are inverted together (16 lanes per block, lane j computes column
// j), then blocks of 16, 32 and 64 are doubled with
//   inv([L11 0; L21 L22]) = [X11 0; -X22 (L21 X11)  X22],
// both products on the matrix cores; L21 X11 is parked in the slot that
// -X22 (L21 X11) then overwrites.
//
// MFMA f64 16x16x4 (lane l of 64): A operand A(l%16, l/16), B operand
// B(l/16, l%16), result element v = D(4v + l/16, l%16) of D = A B.
constexpr int PW = 16;   // panel width

// Lane `lane`'s value of v, as a wave-uniform (scalar register) value.
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(u), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(u >> 32), lane);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// One 16x16 tile of D = A B accumulated over K (a multiple of 4): a_at(i, k)
// and b_at(k, j) give the operands.
template <int K, class FA, class FB>
__device__ __forceinline__ d4 mfma16(int lane, FA a_at, FB b_at)
{
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  const int i = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a_at(i, k0 + kq), b_at(k0 + kq, i), acc, 0, 0, 0);
  return acc;
}

// X(i, j) of the inverse from its storage in P (0 above the diagonal).
template <int BK>
__device__ __forceinline__ double xinv(const double (*P)[BK + 1], int i, int j)
{
  return (i > j) ? P[j][i] : ((i == j) ? P[j][BK] : 0.0);
}

// One doubling level: for every pair of inverted diagonal blocks of size H,
// X21 = -X22 (L21 X11), in 16 x 16 tiles (HB = H/16 per side).  The pair's
// HB waves take HB tiles each: in T = L21 X11 wave wp takes tile columns
// cb = u (u < HB) in rows (wp + u) % HB; in -X22 T, tile rows rb = u in
// columns (wp + u) % HB.  As X11 and X22 are lower triangular, a tile takes
// the k-blocks kk >= cb (resp. kk <= rb) only, so every wave does the same
// HB (HB + 1) / 2 block products.  All of a wave's operands are read from
// LDS before its MFMA chains run, interleaved.
template <int BK, int H>
__device__ __forceinline__ void inverse_level(double (*P)[BK + 1], int wave, int lane)
{
  constexpr int HB = H / 16, NPR = BK / (2 * H), WPP = 4 / NPR;   // waves per pair
  const int pair = wave / WPP, wp = wave % WPP;
  const bool act = wp < HB;
  const int b1 = 2 * H * pair;
  const int i = lane & 15, kq = lane >> 4;
  d4 acc[HB];
  double fa[HB][HB][4], fb[HB][HB][4];
  // T = L21 X11, stored where X21 goes (T(r, c) at P[c][r])
  if (act) {
#pragma unroll
    for (int u = 0; u < HB; u++) {
      const int r0 = b1 + H + 16 * ((wp + u) % HB), c0 = b1 + 16 * u;
#pragma unroll
      for (int kk = u; kk < HB; kk++)
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const int k = b1 + 16 * kk + 4 * m + kq;
          fa[u][kk][m] = P[r0 + i][k];
          fb[u][kk][m] = (kk == u) ? xinv<BK>(P, k, c0 + i) : P[c0 + i][k];
        }
    }
#pragma unroll
    for (int u = 0; u < HB; u++) acc[u] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < HB; kk++)
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int u = 0; u <= kk; u++)
          acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[u][kk][m], fb[u][kk][m], acc[u], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < HB; u++) {
      const int r0 = b1 + H + 16 * ((wp + u) % HB), c0 = b1 + 16 * u;
#pragma unroll
      for (int v = 0; v < 4; v++) P[c0 + i][r0 + 4 * v + kq] = acc[u][v];
    }
  }
  __syncthreads();
  // X21 = -X22 T; every read of T before the in-place writes
  if (act) {
#pragma unroll
    for (int u = 0; u < HB; u++) {
      const int r0 = b1 + H + 16 * u, c0 = b1 + 16 * ((wp + u) % HB);
#pragma unroll
      for (int kk = 0; kk <= u; kk++)
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const int k = b1 + H + 16 * kk + 4 * m + kq;
          fa[u][kk][m] = (kk == u) ? xinv<BK>(P, r0 + i, k) : P[k][r0 + i];
          fb[u][kk][m] = P[c0 + i][k];
        }
    }
#pragma unroll
    for (int u = 0; u < HB; u++) acc[u] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < HB; kk++)
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int u = kk; u < HB; u++)
          acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[u][kk][m], fb[u][kk][m], acc[u], 0, 0, 0);
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int u = 0; u < HB; u++) {
      const int r0 = b1 + H + 16 * u, c0 = b1 + 16 * ((wp + u) % HB);
#pragma unroll
      for (int v = 0; v < 4; v++) P[c0 + i][r0 + 4 * v + kq] = -acc[u][v];
    }
  }
  __syncthreads();
}

// Entry e < BK (BK + 1) / 2 of the lower triangle of a BK x BK block (BK
// even), columns folded in pairs c, BK-1-c (together BK + 1 entries), so
// consecutive e run down one column.
template <int BK>
__device__ __forceinline__ void lower_entry(int e, int& r, int& col)
{
  const int pr = e / (BK + 1), o = e - pr * (BK + 1);
  const bool first = o < BK - pr;
  col = first ? pr : BK - 1 - pr;
  r = first ? pr + o : col + o - (BK - pr);
}

// The doubling levels and the store of the whole inverse, from P holding L
// in its lower part and the inverses of the 16x16 diagonal blocks (see
// potrf_block), into the 128 x 128 dinv slot W (column-major; entries
// outside the b x b lower triangle stay at the zeros written when the slot
// was created).  All threads of the workgroup.
template <int BK>
__device__ __forceinline__ void inverse_full(double (*P)[BK + 1], double* W, int b, int tid,
                                             int wave, int lane)
{
  constexpr int NL = BK * (BK + 1) / 2, PER = (NL + 255) / 256;
  if constexpr (BK >= 32) inverse_level<BK, 16>(P, wave, lane);
  if constexpr (BK >= 64) inverse_level<BK, 32>(P, wave, lane);
  if constexpr (BK >= 128) inverse_level<BK, 64>(P, wave, lane);
  // over the folded lower triangle, LDS reads first so they overlap
  constexpr int SB2 = (PER + 1) / 2;   // entries per batch
#pragma unroll
  for (int u0 = 0; u0 < PER; u0 += SB2) {
    double xv[SB2];
#pragma unroll
    for (int u = 0; u < SB2; u++) {
      int r, col;
      lower_entry<BK>(min(tid + (u0 + u) * 256, NL - 1), r, col);
      xv[u] = xinv<BK>(P, r, col);
    }
#pragma unroll
    for (int u = 0; u < SB2; u++) {
      int r, col;
      const int e = tid + (u0 + u) * 256;
      lower_entry<BK>(min(e, NL - 1), r, col);
      if (u0 + u < PER && e < NL && r < b) W[static_cast<int64_t>(col) * BT + r] = xv[u];
    }
  }
}

// P[r][c] -= sum_k P[r][k] P[c][k] over the np 16-column panels from column
// kb0 on, for the 16 x 16 tiles e < nt whose top-left corners tile(e, r0, c0)
// gives, tile e to wave w0 + e % nw (the calling waves are w0 .. w0+nw-1),
// two tiles at a time: each panel's operand reads first, the two MFMA
// chains interleaved.  Whole tiles: the upper half of a diagonal tile is
// scratch until the inverse is written.
template <int BK, class TILE>
__device__ __forceinline__ void panel_tiles(double (*P)[BK + 1], int kb0, int np, int nt, int w0,
                                            int nw, int wave, int lane, TILE tile)
{
  const int i = lane & 15, kq = lane >> 4;
  for (int e = wave - w0; e < nt; e += 2 * nw) {
    int r0[2], c0[2];
    const bool two = e + nw < nt;
    tile(e, r0[0], c0[0]);
    tile(two ? e + nw : e, r0[1], c0[1]);
    d4 acc[2];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc[h][v] = P[r0[h] + 4 * v + kq][c0[h] + i];
    for (int q = 0; q < np; q++) {
      const int kb = kb0 + PW * q;
      double fa[2][4], fb[2][4];
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
          fa[h][k] = -P[r0[h] + i][kb + 4 * k + kq];
          fb[h][k] = P[c0[h] + i][kb + 4 * k + kq];
        }
      // C(r, c) -= sum_k A(r, k) B(c, k): result element v is row 4v + lane/16
      // of the tile, column lane % 16
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int h = 0; h < 2; h++)
          acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[h][k], fb[h][k], acc[h], 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) P[r0[0] + 4 * v + kq][c0[0] + i] = acc[0][v];
    if (two) {
#pragma unroll
      for (int v = 0; v < 4; v++) P[r0[1] + 4 * v + kq][c0[1] + i] = acc[1][v];
    }
  }
}

// Diagonal block `step` of supernode s, by one workgroup of 256 threads with
// smem >= potrf_smem<BK>().  BK: padded block edge (32, 64 or 128) >= b.
// FULL: the whole inverse; otherwise only the inverses of its 16x16 diagonal
// blocks (the blocked TRSM's), the rest formed after the level (k_inv_blocks).
template <int BK, bool FULL>
__device__ __forceinline__ void potrf_block(const CholDev& c, double* L, double* dinv, int* pivot,
                                            int s, int step, double* smem)
{
  constexpr int NPN = BK / PW;   // panels
  const SnInfo si = sninfo(c, L, s);
  const int k0 = step * BT;
  if (k0 >= si.nc) return;
  const int b = min(BK, si.nc - k0);
  double (*P)[BK + 1] = reinterpret_cast<double (*)[BK + 1]>(smem);
  double* rd = smem + BK * (BK + 1);   // 1 / L(r, r)
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // scalar tile bookkeeping
  double* D = si.Ls + static_cast<int64_t>(k0) * si.f + k0;
  // load the lower triangle (padded with the identity past b), all of a
  // thread's entries in flight at once.  Every lane loads (entries past b or
  // past the triangle read D[0] and are replaced), so the loads are not split
  // by branches that would wait on each other.
  constexpr int NL = BK * (BK + 1) / 2, PER = (NL + 255) / 256;
  {
    double v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      int r, col;
      lower_entry<BK>(min(tid + u * 256, NL - 1), r, col);
      const bool in = r < b;
      const double t = ldg<false>(D + (in ? r + static_cast<int64_t>(col) * si.f : 0));
      v[u] = in ? t : (r == col ? 1.0 : 0.0);
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      int r, col;
      lower_entry<BK>(min(tid + u * 256, NL - 1), r, col);
      if (tid + u * 256 < NL) P[r][col] = v[u];
    }
  }
  if (tid == 0) bad = INT_MAX;
  __syncthreads();

  // ---- blocked Cholesky, one 16-column panel at a time ----
  // Each factoring wave takes the panel's 16 x 16 diagonal block in lanes
  // 0..15 (lane l holds row kb + l; every such wave the same) and 48 of the
  // rows below in lanes 16..63, so the whole panel is factored in one pass:
  // at step k the pivot row's entries reach every lane of the wave through
  // v_readlane (read before the pivot's 1/sqrt is known, scaled after),
  // column k is scaled and the later columns take its rank-1 update.  Entries
  // above the diagonal of the 16 x 16 block are carried along, never stored,
  // so no lane-dependent branches.  Only as many waves factor as the panel's
  // rows need (three for the first panel, then two, then one); the others
  // meanwhile bring the next tile column up to date with all earlier panels
  // (left-looking, so this work stays small and even), and after the panel
  // all waves apply the panel itself to that column.
  {
    const bool dg = lane < PW;
    const int prow_off = dg ? lane : PW + 48 * wave + (lane - PW);   // row - kb
#pragma unroll 1
    for (int p = 0; p < NPN; p++) {
      const int kb = p * PW;
      const int nf = max(1, (BK - kb - PW + 47) / 48);   // factoring waves
      if (wave < nf) {
        const int r = kb + prow_off;
        const bool mine = r < BK && (dg ? wave == 0 : true);   // this lane stores row r
        const int rr = min(r, BK - 1);
        double v[PW];
#pragma unroll
        for (int q = 0; q < PW; q++) v[q] = P[rr][kb + q];
        double myinv = 1.0;   // 1 / L(kb + lane, kb + lane), taken at step k == lane
#pragma unroll
        for (int k = 0; k < PW; k++) {
          const double dkk = readlane_f64(v[k], k);
          double lu[PW];   // unscaled entries (kb+q, kb+k) of rows q > k
#pragma unroll
          for (int q = k + 1; q < PW; q++) lu[q] = readlane_f64(v[k], q);
          // 1/sqrt(dkk), v_rsq_f64 and two Newton steps; the pivot test
          // comes after the panel, off the column-to-column chain
          double y = __builtin_amdgcn_rsq(dkk);
          const double hx = 0.5 * dkk;
          y = y * fma(-hx, y * y, 1.5);
          y = y * fma(-hx, y * y, 1.5);
          // L(r, kb+k) = v_k y; the update L(r, kb+k) L(kb+q, kb+k) is v_k y^2 lu_q
          const double wk = v[k] * (y * y);
          v[k] = (lane == k) ? dkk * y : v[k] * y;
          myinv = (lane == k) ? y : myinv;
          // formed here, so the step's masks end with the step (the compiler
          // would otherwise defer these selects to the stores, holding every
          // step's masks in scalar registers)
          asm volatile("" : "+v"(v[k]), "+v"(myinv));
#pragma unroll
          for (int q = k + 1; q < PW; q++) v[q] -= wk * lu[q];
        }
        // whole rows, also the diagonal block's carried entries above the
        // diagonal: that half of a diagonal tile is scratch until the inverse
        // is written there (one store mask, no per-entry branches)
        if (mine) {
#pragma unroll
          for (int q = 0; q < PW; q++) P[r][kb + q] = v[q];
        }
        if (wave == 0) {
          // the pivot test, sqrt(pivot) > pivot_tol, on L(kb+l, kb+l) as lane l
          // just stored it: a failed pivot leaves a small value or NaN there
          // (NaN then runs through the rest of this system's factor; the
          // system is reported and kept out of the solve).  The first
          // failing column of the first failing panel is kept.
          const double lkk = P[kb + (lane & 15)][kb + (lane & 15)];
          const uint64_t f = __builtin_amdgcn_ballot_w64(dg && kb + lane < b && !(lkk > c.pivot_tol));
          if (f != 0 && tid == 0 && bad == INT_MAX) bad = si.a + k0 + kb + __builtin_ctzll(f);
          if (dg) rd[kb + lane] = myinv;
        }
      } else if (p > 0 && kb + PW < BK) {
        // the next tile column (its tiles from the diagonal down) by the
        // panels before this one
        const int cb = kb + PW;
        panel_tiles<BK>(P, 0, p, (BK - cb) / PW, nf, 4 - nf, wave, lane,
                        [&](int e, int& r0, int& c0) {
                          r0 = cb + PW * e;
                          c0 = cb;
                        });
      }
      __syncthreads();
      if (kb + PW < BK) {
        // and by this panel
        const int cb = kb + PW;
        panel_tiles<BK>(P, kb, 1, (BK - cb) / PW, 0, 4, wave, lane, [&](int e, int& r0, int& c0) {
          r0 = cb + PW * e;
          c0 = cb;
        });
        __syncthreads();
      }
    }
  }

  // L_kk back to the factor (the inverse below keeps P's lower part), the
  // stores draining while the inverse is formed
#pragma unroll
  for (int u = 0; u < PER; u++) {
    int r, col;
    const int e = tid + u * 256;
    lower_entry<BK>(min(e, NL - 1), r, col);
    const double lv = P[r][col];
    if (e < NL && r < b) stg<false>(D + r + static_cast<int64_t>(col) * si.f, lv);
  }

  // ---- inverse: the 16x16 diagonal blocks (16 lanes each), then doubling ----
  if (tid < BK) {
    const int bb = (tid / PW) * PW, j = tid % PW;
    double x[PW];
#pragma unroll
    for (int i = 0; i < PW; i++) {
      double t0 = (i == j) ? 1.0 : 0.0, t1 = 0.0;
#pragma unroll
      for (int k = 0; k < i; k++) {
        if (k & 1) t1 -= P[bb + i][bb + k] * x[k];
        else t0 -= P[bb + i][bb + k] * x[k];
      }
      x[i] = (i >= j) ? (t0 + t1) * rd[bb + i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < PW; i++) {
      if (i == j) P[bb + j][BK] = x[i];
      else if (i > j) P[bb + j][bb + i] = x[i];
    }
  }
  __syncthreads();
  double* W = dinv + c.dinv_off[s] + static_cast<int64_t>(step) * BT * BT;
  if constexpr (FULL) {
    inverse_full<BK>(P, W, b, tid, wave, lane);
  } else {
    // the 16x16 diagonal blocks of the inverse, at their places in the slot
    constexpr int NE = (BK / PW) * (PW * (PW + 1) / 2);
    for (int e = tid; e < NE; e += 256) {
      const int bb = (e / (PW * (PW + 1) / 2)) * PW;
      int r, col;
      lower_entry<PW>(e % (PW * (PW + 1) / 2), r, col);
      if (bb + r < b) W[static_cast<int64_t>(bb + col) * BT + bb + r] = xinv<BK>(P, bb + r, bb + col);
    }
  }
  if (tid == 0 && bad != INT_MAX) atomicMin(pivot, bad);
}

template <int BK>
__global__ void __launch_bounds__(256) k_potrf_inv(CholDev c, const int* __restrict__ sns,
                                                   int step, Batch bt)
{
  __builtin_amdgcn_s_setprio(3);   // the critical chain: issue priority (see k_dense_nt)
  extern __shared__ double smem[];
  const int sy = batch_sys(bt);
  if (sy < 0) return;
  const int s = sns[blockIdx.x];
  if (step * BT >= c.sn_first[s + 1] - c.sn_first[s]) return;
  potrf_block<BK, (BK < 128)>(c, batch_at(c.L, sy, bt), batch_at(c.dinv, sy, bt), c.pivot + sy,
                              s, step, smem);
}

// The whole inverse of 128-wide diagonal blocks whose factorization left only
// the 16x16 diagonal blocks' inverses (potrf_block<128, false>): L_kk from
// the panel, those inverses from the dinv slot, then the doubling levels.
// Task (s, step); system blockIdx.y.  Same arithmetic as the in-kernel
// inverse, so the same bits.
__global__ void __launch_bounds__(256) k_inv_blocks(CholDev c, const int* __restrict__ tsn,
                                                    const int* __restrict__ tstep, Batch bt)
{
  constexpr int BK = BT;
  constexpr int NL = BK * (BK + 1) / 2, PER = (NL + 255) / 256;
  constexpr int T16 = PW * (PW + 1) / 2;
  extern __shared__ double smem[];
  const int sy = batch_sys(bt);
  if (sy < 0) return;
  const int s = tsn[blockIdx.x], step = tstep[blockIdx.x];
  const SnInfo si = sninfo(c, batch_at(c.L, sy, bt), s);
  const int k0 = step * BT;
  if (k0 >= si.nc) return;
  const int b = min(BK, si.nc - k0);
  double (*P)[BK + 1] = reinterpret_cast<double (*)[BK + 1]>(smem);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const double* D = si.Ls + static_cast<int64_t>(k0) * si.f + k0;
  double* W = batch_at(c.dinv, sy, bt) + c.dinv_off[s] + static_cast<int64_t>(step) * BT * BT;
  {
    double v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      int r, col;
      lower_entry<BK>(min(tid + u * 256, NL - 1), r, col);
      const bool in = r < b;
      const double t = D[in ? r + static_cast<int64_t>(col) * si.f : 0];
      v[u] = in ? t : (r == col ? 1.0 : 0.0);
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      int r, col;
      lower_entry<BK>(min(tid + u * 256, NL - 1), r, col);
      if (tid + u * 256 < NL) P[r][col] = v[u];
    }
  }
  for (int e = tid; e < (BK / PW) * T16; e += 256) {
    const int bb = (e / T16) * PW;
    int r, col;
    lower_entry<PW>(e % T16, r, col);
    const double x = (bb + r < b) ? W[static_cast<int64_t>(bb + col) * BT + bb + r]
                                  : (r == col ? 1.0 : 0.0);
    if (r == col) P[bb + col][BK] = x;
    else P[bb + col][bb + r] = x;
  }
  __syncthreads();
  inverse_full<BK>(P, W, b, tid, wave, lane);
}

// L21 = A21 L_kk^{-T} for the rows below 128-wide diagonal block `step`, by
// blocked forward substitution over its eight 16-column panels (no inverse of
// the whole block needed): X_p^T = inv16(L_pp) (A_p^T - sum_{q<p} L_pq X_q^T).
// A wave owns 16-row groups; a group's X_q^T live in MFMA accumulators, whose
// lane layout (element v = D[4v + l/16][l%16]) is the B operand's of k-step
// v, so they feed the later panels' products without going through LDS.  The
// 28 blocks L_pq (p > q) and the 8 inverses inv16(L_pp) are staged in LDS
// first (each 16 x 16, column-major, in batches of loads), each panel's A
// values are loaded while the one before is computed.  Few registers: a
// 32-row tile fits beside two workgroups of the trailing update on one CU,
// so the critical path's TRSM does not wait for them to drain.  TM rows per tile (32 or 128: 2 or 8
// groups); every row takes the same operations in the same order whatever
// the tile, so both shapes give the same bits.  In place.  System blockIdx.y.
constexpr int TRSM_NB = (BT / PW) * (BT / PW - 1) / 2;   // blocks L_pq, p > q
constexpr size_t TRSM_SMEM = (TRSM_NB + BT / PW) * PW * PW * sizeof(double);

template <int TM>
__global__ void __launch_bounds__(256) k_trsm(CholDev c, const int* __restrict__ tsn,
                                              const int* __restrict__ ttr, int step, Batch bt)
{
  __builtin_amdgcn_s_setprio(3);   // the critical chain: issue priority (see k_dense_nt)
  constexpr int NG = TM / 16, NGW = (NG + 3) / 4;   // groups, groups per wave
  constexpr int NPN = BT / PW;                       // panels
  extern __shared__ double smem[];
  double* Ls = smem;                                 // block (p, q) at (p(p-1)/2 + q) * 256
  double* Is = smem + TRSM_NB * PW * PW;             // inv16(L_pp) at p * 256
  const int sy = batch_sys(bt);
  if (sy < 0) return;
  const int s = tsn[blockIdx.x], tr = ttr[blockIdx.x];
  const SnInfo si = sninfo(c, batch_at(c.L, sy, bt), s);
  const int k0 = step * BT;
  const int b = min(BT, si.nc - k0);
  const int R0 = k0 + b + tr * TM;
  const int M = min(TM, si.f - R0);
  if (b <= 0 || M <= 0) return;   // uniform over the workgroup
  const int64_t f = si.f;
  const double* Lkk = si.Ls + static_cast<int64_t>(k0) * f + k0;   // L(k0 + r, k0 + c) at c*f + r
  double* A = si.Ls + static_cast<int64_t>(k0) * f + R0;            // A(row, col) at col*f + row
  const double* W = batch_at(c.dinv, sy, bt) + c.dinv_off[s] + static_cast<int64_t>(step) * BT * BT;
  const int np = (b + PW - 1) / PW;
  // stage: thread t holds entry (row t % 16, column t / 16) of every block;
  // rows at or past b (the panel rows below a partial block) read as 0.  All
  // loads are issued before the first LDS store.
  {
    const int rr = threadIdx.x % PW, cc = threadIdx.x / PW;
    constexpr int NBAT = 9;   // loads in flight per batch
#pragma unroll
    for (int t0 = 0; t0 < TRSM_NB + NPN; t0 += NBAT) {
    double v[NBAT];
#pragma unroll
    for (int t = t0; t < t0 + NBAT; t++) {
      if (t < TRSM_NB) {
        int p = 1;
        while ((p + 1) * p / 2 <= t) p++;   // compile-time after unrolling
        const int q = t - p * (p - 1) / 2;
        const int row = PW * p + rr;
        const bool in = p < np && row < b;
        const double w = Lkk[in ? static_cast<int64_t>(PW * q + cc) * f + row : 0];
        v[t - t0] = in ? w : 0.0;
      } else {
        const int p = t - TRSM_NB;
        v[t - t0] = W[static_cast<int64_t>(PW * p + cc) * BT + PW * p + rr];
      }
    }
#pragma unroll
    for (int t = t0; t < t0 + NBAT; t++) smem[t * PW * PW + threadIdx.x] = v[t - t0];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i16 = lane & 15, k4 = lane >> 4;
  int grow[NGW];   // first row of each of this wave's groups (-1: none)
#pragma unroll
  for (int gi = 0; gi < NGW; gi++) {
    const int g = wave + 4 * gi;
    grow[gi] = (g < NG && g * 16 < M) ? g * 16 : -1;
  }
  if (grow[0] < 0) return;   // no barrier follows
  // A_p^T of each group: element v = A(row i16, column 16p + 4v + k4)
  auto load_a = [&](int p, d4 (&a)[NGW]) {
#pragma unroll
    for (int gi = 0; gi < NGW; gi++) {
      const int row = grow[gi] + i16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int col = PW * p + 4 * v + k4;
        a[gi][v] = (grow[gi] >= 0 && row < M && col < b) ? A[col * f + row] : 0.0;
      }
    }
  };
  d4 x[NGW][NPN];
  d4 an[NGW];
  load_a(0, an);
#pragma unroll
  for (int p = 0; p < NPN; p++) {
    if (p >= np) break;
    d4 acc[NGW];
#pragma unroll
    for (int gi = 0; gi < NGW; gi++) acc[gi] = an[gi];
    if (p + 1 < np) load_a(p + 1, an);
    // - sum_{q<p} L_pq X_q^T: A operand -L(16p + i16, 16q + 4 s4 + k4)
#pragma unroll
    for (int q = 0; q < NPN; q++) {
      if (q >= p) break;
      const double* blk = Ls + (p * (p - 1) / 2 + q) * PW * PW;
#pragma unroll
      for (int s4 = 0; s4 < 4; s4++) {
        const double a = -blk[(4 * s4 + k4) * PW + i16];
#pragma unroll
        for (int gi = 0; gi < NGW; gi++)
          acc[gi] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x[gi][q][s4], acc[gi], 0, 0, 0);
      }
    }
    // X_p^T = inv16(L_pp) acc: A operand inv16(i16, 4 s4 + k4)
    d4 r[NGW];
#pragma unroll
    for (int gi = 0; gi < NGW; gi++) r[gi] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
      const double a = Is[p * PW * PW + (4 * s4 + k4) * PW + i16];
#pragma unroll
      for (int gi = 0; gi < NGW; gi++)
        r[gi] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, acc[gi][s4], r[gi], 0, 0, 0);
    }
#pragma unroll
    for (int gi = 0; gi < NGW; gi++) {
      if (p + 1 < NPN) x[gi][p] = r[gi];   // the last panel's X is not read again
      const int row = grow[gi] + i16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int col = PW * p + 4 * v + k4;
        if (grow[gi] >= 0 && row < M && col < b) A[col * f + row] = r[gi][v];
      }
    }
  }
}

}  // namespace

size_t dense_nt_smem() { return 2 * KC * LDA_S * sizeof(double); }
void dense_nt(const CholDev& c, int ntasks, const int* tsn, const int* ttr, const int* ttc,
              int mode, int step, const Batch& b, hipStream_t st, bool short_rows, hipEvent_t done)
{
  if (ntasks <= 0 || b.count <= 0) return;
  const dim3 grid(ntasks, b.count);
  // `done` rides on the launch itself (its completion signal), so no marker
  // of its own sits between this kernel and the next one of the stream
  if (short_rows)
    hipExtLaunchKernelGGL(k_dense_nt<32>, grid, dim3(256), dense_nt_smem(), st, nullptr, done, 0,
                          c, tsn, ttr, ttc, mode, step, b);
  else
    hipExtLaunchKernelGGL(k_dense_nt<128>, grid, dim3(256), dense_nt_smem(), st, nullptr, done, 0,
                          c, tsn, ttr, ttc, mode, step, b);
  HIPCHK(hipGetLastError());
}

void inv_blocks(const CholDev& c, int ntasks, const int* tsn, const int* tstep, const Batch& b,
                hipStream_t st)
{
  if (ntasks <= 0 || b.count <= 0) return;
  constexpr size_t SM = (BT * (BT + 1) + BT) * sizeof(double);
  static const bool attr = [] {
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_inv_blocks),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(SM)));
    return true;
  }();
  (void)attr;
  k_inv_blocks<<<dim3(ntasks, b.count), 256, SM, st>>>(c, tsn, tstep, b);
  HIPCHK(hipGetLastError());
}

void trsm(const CholDev& c, int ntasks, const int* tsn, const int* ttr, int step, const Batch& b,
          hipStream_t st, bool short_rows)
{
  if (ntasks <= 0 || b.count <= 0) return;
  static const bool attr = [] {
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_trsm<32>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(TRSM_SMEM)));
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_trsm<128>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(TRSM_SMEM)));
    return true;
  }();
  (void)attr;
  const dim3 grid(ntasks, b.count);
  if (short_rows) k_trsm<32><<<grid, 256, TRSM_SMEM, st>>>(c, tsn, ttr, step, b);
  else k_trsm<128><<<grid, 256, TRSM_SMEM, st>>>(c, tsn, ttr, step, b);
  HIPCHK(hipGetLastError());
}

template <int BK>
static size_t potrf_smem()
{
  // P and the reciprocals of the diagonal
  return (BK * (BK + 1) + BK) * sizeof(double);
}

template <int BK>
static void launch_potrf(const CholDev& c, int nsn, const int* sns, int step, const Batch& b,
                         hipStream_t st)
{
  // set once per process; a function-local static initialises thread-safely
  static const bool attr = [] {
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_potrf_inv<BK>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(potrf_smem<BK>())));
    return true;
  }();
  (void)attr;
  k_potrf_inv<BK><<<dim3(nsn, b.count), 256, potrf_smem<BK>(), st>>>(c, sns, step, b);
  HIPCHK(hipGetLastError());
}

void potrf_inv(const CholDev& c, int nsn, const int* sns, int step, int max_b, const Batch& b,
               hipStream_t st)
{
  if (nsn <= 0 || b.count <= 0) return;
  if (max_b <= 32) launch_potrf<32>(c, nsn, sns, step, b, st);
  else if (max_b <= 64) launch_potrf<64>(c, nsn, sns, step, b, st);
  else launch_potrf<128>(c, nsn, sns, step, b, st);
}

}  // namespace dev
}  // namespace hykkt
