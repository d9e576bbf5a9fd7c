/*
 * hykkt_oracle.c -- CPU FP64 restatement of the HyKKT KKT solve.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline): nothing in the
 * product links or calls it.  Compiled with -ffp-contract=off so that every
 * fused multiply-add below is one written explicitly with fma().
 *
 * Arithmetic-order conventions that the GPU kernels reproduce bit for bit:
 *   - SpGEMM value of E(i,j) = alpha_s*D(i,j) + beta_s*(alpha_p*acc), where
 *     acc = fma(A(i,k), B(k,j), acc) over k ascending, starting from 0.
 *   - Ruiz: v *= (s_i * s_j); s_i = 1/sqrt(max) (IEEE-exact operations only).
 * Everything else (SpMV, Cholesky, CG) is compared within stated tolerances.
 */
#include "hykkt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void* xmalloc(size_t bytes)
{
  void* p = malloc(bytes ? bytes : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", bytes);
    abort();
  }
  return p;
}

static void* xcalloc(size_t n, size_t sz)
{
  void* p = calloc(n ? n : 1, sz ? sz : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

void orc_free(void* p) { free(p); }

/* ======================================================================
 * MatrixMarket ingest -- src/input_functions.cpp:16-217
 * ====================================================================== */

/* The size line is the LAST of the `skip` skipped lines
 * (input_functions.cpp:26-31). */
int orc_mm_header(const char* path, int skip, int* n, int* m, int* nnz)
{
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[256];
  line[0] = 0;
  for (int i = 0; i < skip; i++) {
    if (!fgets(line, sizeof line, f)) { fclose(f); return -2; }
  }
  fclose(f);
  if (sscanf(line, "%d %d %d", n, m, nnz) != 3) return -3;
  return 0;
}

/* Entries follow the size line, one "r c v" per line, 1-based
 * (input_functions.cpp:40-50). Returns the number of entries read. */
int orc_mm_read_coo(const char* path, int skip, int cap, int* rows, int* cols, double* vals)
{
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[256];
  for (int i = 0; i < skip; i++) {
    if (!fgets(line, sizeof line, f)) { fclose(f); return -2; }
  }
  int k = 0;
  while (fgets(line, sizeof line, f)) {
    int r, c;
    double v;
    if (sscanf(line, "%d %d %lf", &r, &c, &v) != 3) continue;
    if (k >= cap) { fclose(f); return -4; }
    rows[k] = r - 1;
    cols[k] = c - 1;
    vals[k] = v;
    k++;
  }
  fclose(f);
  return k;
}

/* input_functions.cpp:56-66: every off-diagonal stored entry is mirrored. */
int orc_sym_coo_nnz(int n, int nnz, const int* r, const int* c)
{
  (void)n;
  int t = 0;
  for (int i = 0; i < nnz; i++) t += (r[i] != c[i]) ? 2 : 1;
  return t;
}

/* Stable insertion sort of (idx, val) pairs by idx; rows are short. */
static void sort_row_pairs(int len, int* idx, double* val)
{
  for (int i = 1; i < len; i++) {
    int ki = idx[i];
    double kv = val[i];
    int j = i - 1;
    while (j >= 0 && idx[j] > ki) {
      idx[j + 1] = idx[j];
      val[j + 1] = val[j];
      j--;
    }
    idx[j + 1] = ki;
    val[j + 1] = kv;
  }
}

/* input_functions.cpp:54-119: mirror the stored triangle into full CSR,
 * entries per row in file order then sorted by column. */
void orc_sym_coo_to_csr(int n, int nnz, const int* r, const int* c, const double* v,
                        int* rp, int* ci, double* cv)
{
  int* cnt = (int*)xcalloc((size_t)n, sizeof(int));
  for (int i = 0; i < nnz; i++) {
    cnt[r[i]]++;
    if (r[i] != c[i]) cnt[c[i]]++;
  }
  rp[0] = 0;
  for (int i = 0; i < n; i++) rp[i + 1] = rp[i] + cnt[i];
  memset(cnt, 0, sizeof(int) * (size_t)n);
  for (int i = 0; i < nnz; i++) {
    int row = r[i];
    int p = rp[row] + cnt[row]++;
    ci[p] = c[i];
    cv[p] = v[i];
    if (r[i] != c[i]) {
      row = c[i];
      p = rp[row] + cnt[row]++;
      ci[p] = r[i];
      cv[p] = v[i];
    }
  }
  for (int i = 0; i < n; i++) sort_row_pairs(rp[i + 1] - rp[i], ci + rp[i], cv + rp[i]);
  free(cnt);
}

/* input_functions.cpp:121-175. */
void orc_coo_to_csr(int n, int nnz, const int* r, const int* c, const double* v,
                    int* rp, int* ci, double* cv)
{
  int* cnt = (int*)xcalloc((size_t)n, sizeof(int));
  for (int i = 0; i < nnz; i++) cnt[r[i]]++;
  rp[0] = 0;
  for (int i = 0; i < n; i++) rp[i + 1] = rp[i] + cnt[i];
  memset(cnt, 0, sizeof(int) * (size_t)n);
  for (int i = 0; i < nnz; i++) {
    int p = rp[r[i]] + cnt[r[i]]++;
    ci[p] = c[i];
    cv[p] = v[i];
  }
  for (int i = 0; i < n; i++) sort_row_pairs(rp[i + 1] - rp[i], ci + rp[i], cv + rp[i]);
  free(cnt);
}

/* input_functions.cpp:198-217: skip 2 lines, then one value per line. */
int orc_read_rhs(const char* path, int cap, double* out)
{
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[256];
  for (int i = 0; i < 2; i++) {
    if (!fgets(line, sizeof line, f)) { fclose(f); return -2; }
  }
  int k = 0;
  while (fgets(line, sizeof line, f)) {
    double v;
    if (sscanf(line, "%lf", &v) != 1) continue;
    if (k >= cap) { fclose(f); return -4; }
    out[k++] = v;
  }
  fclose(f);
  return k;
}

/* ======================================================================
 * Sparse building blocks
 * ====================================================================== */

/* CSR (n x m) -> CSR of the transpose (m x n); within a row of the transpose
 * entries are in increasing original-row order (cusparse_utils.cpp:49-98,
 * csr2csc semantics). */
void orc_transpose(int n, int m, const int* rp, const int* ci, const double* v,
                   int* trp, int* tci, double* tv)
{
  int* cnt = (int*)xcalloc((size_t)m + 1, sizeof(int));
  for (int p = 0; p < rp[n]; p++) cnt[ci[p] + 1]++;
  for (int j = 0; j < m; j++) cnt[j + 1] += cnt[j];
  memcpy(trp, cnt, sizeof(int) * ((size_t)m + 1));
  for (int i = 0; i < n; i++) {
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int q = cnt[ci[p]]++;
      tci[q] = i;
      if (tv) tv[q] = v[p];
    }
  }
  free(cnt);
}

/* y = alpha*A*x + beta*y, each row summed in storage order with fma
 * (matrix_vector_ops.cu:11-39 wraps cusparseSpMV with these semantics). */
void orc_spmv(int nrows, const int* rp, const int* ci, const double* v, double alpha,
              const double* x, double beta, double* y)
{
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nrows; i++) {
    double t = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; p++) t = fma(v[p], x[ci[p]], t);
    y[i] = (beta == 0.0) ? alpha * t : fma(alpha, t, beta * y[i]);
  }
}

/* Pattern of E = D + A*B (structural union, no numerical dropping);
 * SpgemmClass.hpp:4-5 / matrix_matrix_ops.cpp:13-146,219-287. */
int orc_spgemm_nnz(int n, int ncols, const int* arp, const int* aci, const int* brp,
                   const int* bci, const int* drp, const int* dci)
{
  int* mark = (int*)xmalloc(sizeof(int) * (size_t)ncols);
  for (int j = 0; j < ncols; j++) mark[j] = -1;
  int total = 0;
  for (int i = 0; i < n; i++) {
    for (int p = drp[i]; p < drp[i + 1]; p++) {
      if (mark[dci[p]] != i) { mark[dci[p]] = i; total++; }
    }
    for (int p = arp[i]; p < arp[i + 1]; p++) {
      int k = aci[p];
      for (int q = brp[k]; q < brp[k + 1]; q++) {
        if (mark[bci[q]] != i) { mark[bci[q]] = i; total++; }
      }
    }
  }
  free(mark);
  return total;
}

static int cmp_int(const void* a, const void* b)
{
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

void orc_spgemm_pattern(int n, int ncols, const int* arp, const int* aci, const int* brp,
                        const int* bci, const int* drp, const int* dci, int* erp, int* eci)
{
  int* mark = (int*)xmalloc(sizeof(int) * (size_t)ncols);
  for (int j = 0; j < ncols; j++) mark[j] = -1;
  int t = 0;
  erp[0] = 0;
  for (int i = 0; i < n; i++) {
    int start = t;
    for (int p = drp[i]; p < drp[i + 1]; p++) {
      if (mark[dci[p]] != i) { mark[dci[p]] = i; eci[t++] = dci[p]; }
    }
    for (int p = arp[i]; p < arp[i + 1]; p++) {
      int k = aci[p];
      for (int q = brp[k]; q < brp[k + 1]; q++) {
        if (mark[bci[q]] != i) { mark[bci[q]] = i; eci[t++] = bci[q]; }
      }
    }
    qsort(eci + start, (size_t)(t - start), sizeof(int), cmp_int);
    erp[i + 1] = t;
  }
  free(mark);
}

/* E = alpha_s*D + beta_s*(alpha_p*A*B) on the precomputed pattern of E
 * (matrix_matrix_ops.cpp:197-215 with beta=0, then :289-328).
 * acc(i,j) = fma(A(i,k), B(k,j), acc) over k ascending. */
void orc_spgemm_values(int n, int ncols, double alpha_p, double alpha_s, double beta_s,
                       const int* arp, const int* aci, const double* av,
                       const int* brp, const int* bci, const double* bv,
                       const int* drp, const int* dci, const double* dv,
                       const int* erp, const int* eci, double* ev)
{
#pragma omp parallel
  {
    double* acc = (double*)xcalloc((size_t)ncols, sizeof(double));
    double* dd = (double*)xcalloc((size_t)ncols, sizeof(double));
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
      for (int p = arp[i]; p < arp[i + 1]; p++) {
        int k = aci[p];
        double a = av[p];
        for (int q = brp[k]; q < brp[k + 1]; q++) acc[bci[q]] = fma(a, bv[q], acc[bci[q]]);
      }
      for (int p = drp[i]; p < drp[i + 1]; p++) dd[dci[p]] = dv[p];
      for (int p = erp[i]; p < erp[i + 1]; p++) {
        int j = eci[p];
        double c = alpha_p * acc[j];
        ev[p] = alpha_s * dd[j] + beta_s * c;
        acc[j] = 0.0;
        dd[j] = 0.0;
      }
    }
    free(acc);
    free(dd);
  }
}

/* Ruiz equilibration of K = [H Jt; J 0]: RuizClass.cpp:54-98 driving
 * adapt_row_max (matrix_vector_ops.cu:218-266) and adapt_diag_scale
 * (matrix_vector_ops.cu:146-186); max_d reset to 1 (:268-283). */
void orc_ruiz(int its, int n, int totn,
              const int* hrp, const int* hci, double* hv,
              const int* jtrp, const int* jtci, double* jtv,
              const int* jrp, const int* jci, double* jv,
              double* rhs1, double* rhs2, double* max_d)
{
  double* scale = (double*)xmalloc(sizeof(double) * (size_t)totn);
  for (int i = 0; i < totn; i++) max_d[i] = 1.0;
  for (int it = 0; it < its; it++) {
    for (int i = 0; i < totn; i++) {
      double max_l = 0.0, max_u = 0.0;
      if (i < n) {
        for (int p = hrp[i]; p < hrp[i + 1]; p++) {
          double e = fabs(hv[p]);
          if (e > max_l) max_l = e;
        }
        for (int p = jtrp[i]; p < jtrp[i + 1]; p++) {
          double e = fabs(jtv[p]);
          if (e > max_u) max_u = e;
        }
        scale[i] = (max_l > max_u) ? 1.0 / sqrt(max_l) : 1.0 / sqrt(max_u);
      } else {
        for (int p = jrp[i - n]; p < jrp[i - n + 1]; p++) {
          double e = fabs(jv[p]);
          if (e > max_l) max_l = e;
        }
        scale[i] = 1.0 / sqrt(max_l);
      }
    }
    for (int i = 0; i < totn; i++) {
      if (i < n) {
        for (int p = hrp[i]; p < hrp[i + 1]; p++) hv[p] *= scale[i] * scale[hci[p]];
        rhs1[i] *= scale[i];
        max_d[i] *= scale[i];
        for (int p = jtrp[i]; p < jtrp[i + 1]; p++) jtv[p] *= scale[i] * scale[n + jtci[p]];
      } else {
        for (int p = jrp[i - n]; p < jrp[i - n + 1]; p++) jv[p] *= scale[i] * scale[jci[p]];
        rhs2[i - n] *= scale[i];
        max_d[i] *= scale[i];
      }
    }
  }
  free(scale);
}

/* ======================================================================
 * Permutation maps -- src/permcheck.cu:78-195, PermClass.cpp:101-123
 * ====================================================================== */

/* permcheck.cu:126-132 */
void orc_perm_rev(int n, const int* perm, int* rev)
{
  for (int i = 0; i < n; i++) rev[perm[i]] = i;
}

/* Insertion sort keyed on a, carrying b (permcheck.cu:78-97). */
static void isort2(int len, int* a, int* b)
{
  for (int i = 1; i < len; i++) {
    int ka = a[i], kb = b[i];
    int j = i - 1;
    while (j >= 0 && a[j] > ka) {
      a[j + 1] = a[j];
      b[j + 1] = b[j];
      j--;
    }
    a[j + 1] = ka;
    b[j + 1] = kb;
  }
}

/* P*A*P^T structure + value map (permcheck.cu:162-195). */
void orc_map_rc(int n, const int* rows, const int* cols, const int* perm, const int* rev,
                int* prows, int* pcols, int* pmap)
{
  prows[0] = 0;
  int count = 0;
  for (int i = 0; i < n; i++) {
    int src = perm[i];
    int s = rows[src], len = rows[src + 1] - s;
    prows[i + 1] = prows[i] + len;
    for (int j = 0; j < len; j++) {
      pmap[count + j] = s + j;
      pcols[count + j] = rev[cols[s + j]];
    }
    isort2(len, pcols + count, pmap + count);
    count += len;
  }
}

/* A*P^T: columns relabelled by rev, rows kept (permcheck.cu:99-124). */
void orc_map_c(int n, const int* rows, const int* cols, const int* rev, int* pcols, int* pmap)
{
  for (int i = 0; i < n; i++) {
    int s = rows[i], len = rows[i + 1] - s;
    for (int j = 0; j < len; j++) {
      pmap[s + j] = s + j;
      pcols[s + j] = rev[cols[s + j]];
    }
    isort2(len, pcols + s, pmap + s);
  }
}

/* P*A: rows reordered, columns kept (permcheck.cu:134-160). */
void orc_map_r(int n, const int* rows, const int* cols, const int* perm, int* prows,
               int* pcols, int* pmap)
{
  prows[0] = 0;
  int count = 0;
  for (int i = 0; i < n; i++) {
    int src = perm[i];
    int s = rows[src], len = rows[src + 1] - s;
    prows[i + 1] = prows[i] + len;
    for (int j = 0; j < len; j++) {
      pmap[count + j] = s + j;
      pcols[count + j] = cols[s + j];
    }
    count += len;
  }
}

/* ======================================================================
 * Up-looking simplicial Cholesky of the LOWER triangle of a CSR matrix
 * (entries (i,j) with j <= i).  This is the matrix cusolverSpDcsrchol*
 * factors in CholeskyClass.cpp:31-101.  Written from the textbook
 * up-looking algorithm: row k of L solves L(0:k,0:k) l = A(0:k,k) over the
 * elimination-tree reach of row k (Liu 1990 for the tree).
 * ====================================================================== */
struct orc_chol {
  int n;
  const int* arp;   /* borrowed pattern of A */
  const int* aci;
  int* parent;
  long long* lp;    /* column pointers, diagonal first */
  int* li;
  double* lx;
  long long nnz;
  double flops;
};

/* Elimination-tree reach of row k restricted to columns < k; returns the
 * count, pattern written to s (unsorted). */
static int ereach(const orc_chol* c, int k, int* mark, int* s)
{
  int top = 0;
  mark[k] = k;
  for (int p = c->arp[k]; p < c->arp[k + 1]; p++) {
    int i = c->aci[p];
    if (i >= k) continue;
    while (mark[i] != k) {
      s[top++] = i;
      mark[i] = k;
      i = c->parent[i];
      if (i < 0) break;
    }
  }
  return top;
}

orc_chol* orc_chol_create(int n, const int* rp, const int* ci)
{
  orc_chol* c = (orc_chol*)xcalloc(1, sizeof(orc_chol));
  c->n = n;
  c->arp = rp;
  c->aci = ci;
  c->parent = (int*)xmalloc(sizeof(int) * (size_t)n);
  int* anc = (int*)xmalloc(sizeof(int) * (size_t)n);
  for (int k = 0; k < n; k++) {
    c->parent[k] = -1;
    anc[k] = -1;
    for (int p = rp[k]; p < rp[k + 1]; p++) {
      int i = ci[p];
      while (i != -1 && i < k) {
        int nxt = anc[i];
        anc[i] = k;
        if (nxt == -1) { c->parent[i] = k; break; }
        i = nxt;
      }
    }
  }
  free(anc);
  long long* cnt = (long long*)xcalloc((size_t)n + 1, sizeof(long long));
  int* mark = (int*)xmalloc(sizeof(int) * (size_t)n);
  int* s = (int*)xmalloc(sizeof(int) * (size_t)n);
  for (int k = 0; k < n; k++) mark[k] = -1;
  for (int k = 0; k < n; k++) {
    int top = ereach(c, k, mark, s);
    for (int t = 0; t < top; t++) cnt[s[t]]++;
    cnt[k]++;
  }
  c->lp = (long long*)xmalloc(sizeof(long long) * ((size_t)n + 1));
  c->lp[0] = 0;
  double fl = 0.0;
  for (int k = 0; k < n; k++) {
    c->lp[k + 1] = c->lp[k] + cnt[k];
    fl += (double)cnt[k] * (double)cnt[k];
  }
  c->nnz = c->lp[n];
  c->flops = fl;
  c->li = (int*)xmalloc(sizeof(int) * (size_t)c->nnz);
  c->lx = (double*)xmalloc(sizeof(double) * (size_t)c->nnz);
  free(cnt);
  free(mark);
  free(s);
  return c;
}

long long orc_chol_nnz(const orc_chol* c) { return c->nnz; }
double orc_chol_flops(const orc_chol* c) { return c->flops; }

/* Returns -1 on success, else the first row k whose pivot is not > tol
 * (cusolverSpDcsrcholZeroPivot semantics, CholeskyClass.cpp:83-89). */
int orc_chol_factor(orc_chol* c, const double* vals, double tol)
{
  int n = c->n;
  double* x = (double*)xcalloc((size_t)n, sizeof(double));
  long long* next = (long long*)xmalloc(sizeof(long long) * (size_t)n);
  int* mark = (int*)xmalloc(sizeof(int) * (size_t)n);
  int* s = (int*)xmalloc(sizeof(int) * (size_t)n);
  for (int k = 0; k < n; k++) { mark[k] = -1; next[k] = c->lp[k] + 1; }
  int fail = -1;
  for (int k = 0; k < n; k++) {
    double akk = 0.0;
    for (int p = c->arp[k]; p < c->arp[k + 1]; p++) {
      int j = c->aci[p];
      if (j < k) x[j] = vals[p];
      else if (j == k) akk = vals[p];
    }
    int top = ereach(c, k, mark, s);
    qsort(s, (size_t)top, sizeof(int), cmp_int);
    double d = akk;
    for (int t = 0; t < top; t++) {
      int j = s[t];
      double lkj = x[j] / c->lx[c->lp[j]];
      x[j] = 0.0;
      for (long long p = c->lp[j] + 1; p < next[j]; p++) x[c->li[p]] -= c->lx[p] * lkj;
      d -= lkj * lkj;
      c->li[next[j]] = k;
      c->lx[next[j]] = lkj;
      next[j]++;
    }
    if (!(d > 0.0) || !(sqrt(d) > tol)) { fail = k; break; }
    c->li[c->lp[k]] = k;
    c->lx[c->lp[k]] = sqrt(d);
  }
  free(x);
  free(next);
  free(mark);
  free(s);
  return fail;
}

/* L L^T x = b (CholeskyClass.cpp:93-101). */
void orc_chol_solve(const orc_chol* c, const double* b, double* x)
{
  int n = c->n;
  if (x != b) memcpy(x, b, sizeof(double) * (size_t)n);
  for (int j = 0; j < n; j++) {
    x[j] /= c->lx[c->lp[j]];
    double xj = x[j];
    for (long long p = c->lp[j] + 1; p < c->lp[j + 1]; p++) x[c->li[p]] -= c->lx[p] * xj;
  }
  for (int j = n - 1; j >= 0; j--) {
    double t = x[j];
    for (long long p = c->lp[j] + 1; p < c->lp[j + 1]; p++) t -= c->lx[p] * x[c->li[p]];
    x[j] = t / c->lx[c->lp[j]];
  }
}

void orc_chol_free(orc_chol* c)
{
  if (!c) return;
  free(c->parent);
  free(c->lp);
  free(c->li);
  free(c->lx);
  free(c);
}

/* ======================================================================
 * Schur-complement CG -- SchurComplementConjugateGradient.cpp:89-194.
 * S = J * Hg^{-1} * J^T (m x m);  x: in = x0, out = solution.
 * ====================================================================== */
static double dot(int n, const double* a, const double* b)
{
  double t = 0.0;
  for (int i = 0; i < n; i++) t = fma(a[i], b[i], t);
  return t;
}

int orc_schur_cg(int m, int n,
                 const int* jrp, const int* jci, const double* jv,
                 const int* jtrp, const int* jtci, const double* jtv,
                 const orc_chol* chol, double* x, const double* b,
                 double tol, int itmax, int* iters, double* final_norm)
{
  double* y = (double*)xcalloc((size_t)n, sizeof(double));
  double* z = (double*)xcalloc((size_t)n, sizeof(double));
  double* r = (double*)xmalloc(sizeof(double) * (size_t)m);
  double* w = (double*)xmalloc(sizeof(double) * (size_t)m);
  double* p = (double*)xmalloc(sizeof(double) * (size_t)m);
  double* s = (double*)xmalloc(sizeof(double) * (size_t)m);
  /* setup(): r = w = p = s = b, beta = 0 (:89-100) */
  memcpy(r, b, sizeof(double) * (size_t)m);
  memcpy(w, b, sizeof(double) * (size_t)m);
  memcpy(p, b, sizeof(double) * (size_t)m);
  memcpy(s, b, sizeof(double) * (size_t)m);
  double beta = 0.0;
  /* r = b - S x0 (:104-122) */
  orc_spmv(n, jtrp, jtci, jtv, 1.0, x, 0.0, y);
  orc_chol_solve(chol, y, z);
  orc_spmv(m, jrp, jci, jv, -1.0, z, 1.0, r);
  double gam = dot(m, r, r);
  double gam1 = gam;
  int i = 0;
  if (gam == 0.0) {
    /* Already exact: the recurrence would divide 0/0 (divergence from the
     * reference, which has no such guard). */
    *iters = 0;
    *final_norm = 0.0;
    free(y); free(z); free(r); free(w); free(p); free(s);
    return 0;
  }
  /* w = S r, delta = w.r, alpha = gam/delta (:124-144) */
  orc_spmv(n, jtrp, jtci, jtv, 1.0, r, 0.0, y);
  orc_chol_solve(chol, y, z);
  orc_spmv(m, jrp, jci, jv, 1.0, z, 0.0, w);
  double delta = dot(m, w, r);
  double alpha = gam / delta;
  for (i = 0; i < itmax; i++) {
    for (int k = 0; k < m; k++) {
      p[k] = beta * p[k] + r[k];
      s[k] = beta * s[k] + w[k];
    }
    for (int k = 0; k < m; k++) {
      x[k] = fma(alpha, p[k], x[k]);
      r[k] = fma(-alpha, s[k], r[k]);
    }
    gam1 = dot(m, r, r);
    if (sqrt(gam1) < tol) break;
    orc_spmv(n, jtrp, jtci, jtv, 1.0, r, 0.0, y);
    orc_chol_solve(chol, y, z);
    orc_spmv(m, jrp, jci, jv, 1.0, z, 0.0, w);
    delta = dot(m, w, r);
    beta = gam1 / gam;
    gam = gam1;
    alpha = gam / (delta - beta * gam / alpha);
  }
  *iters = (i < itmax) ? i + 1 : itmax;
  *final_norm = sqrt(gam1);
  free(y); free(z); free(r); free(w); free(p); free(s);
  return (i == itmax) ? 1 : 0;
}

/* ======================================================================
 * Residual of the 4x4 block KKT system.
 *   row1: rx - H dx - Jc^T dyc - Jd^T dyd     row2: rs - (Ds ds - dyd)
 *   row3: ry - Jc dx                           row4: ryd - (Jd dx - ds)
 * resid2 is the reference's rows-1-and-3 form (HykktSolver.cpp:846-930),
 * both relative to ||[rx; rs; ry; ryd]||.
 * ====================================================================== */
void orc_kkt_residual(int n, const int* hrp, const int* hci, const double* hv,
                      int d, const double* ds,
                      int m, const int* jcrp, const int* jcci, const double* jcv,
                      const int* jdrp, const int* jdci, const double* jdv, int nnz_jd,
                      const double* rx, const double* rs, const double* ry, const double* ryd,
                      const double* dx, const double* dsv, const double* dyc, const double* dyd,
                      double* resid2, double* resid4)
{
  double* t = (double*)xmalloc(sizeof(double) * (size_t)(n > m ? n : m) * 1 + sizeof(double) * (size_t)d);
  double nb = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0, r4 = 0.0;
  for (int i = 0; i < n; i++) nb += rx[i] * rx[i];
  for (int i = 0; i < d; i++) nb += rs[i] * rs[i] + ryd[i] * ryd[i];
  for (int i = 0; i < m; i++) nb += ry[i] * ry[i];
  /* row 1 */
  for (int i = 0; i < n; i++) {
    double acc = rx[i];
    for (int p = hrp[i]; p < hrp[i + 1]; p++) acc -= hv[p] * dx[hci[p]];
    t[i] = acc;
  }
  for (int k = 0; k < m; k++)
    for (int p = jcrp[k]; p < jcrp[k + 1]; p++) t[jcci[p]] -= jcv[p] * dyc[k];
  if (nnz_jd > 0)
    for (int k = 0; k < d; k++)
      for (int p = jdrp[k]; p < jdrp[k + 1]; p++) t[jdci[p]] -= jdv[p] * dyd[k];
  for (int i = 0; i < n; i++) r1 += t[i] * t[i];
  /* row 3 */
  for (int k = 0; k < m; k++) {
    double acc = ry[k];
    for (int p = jcrp[k]; p < jcrp[k + 1]; p++) acc -= jcv[p] * dx[jcci[p]];
    r3 += acc * acc;
  }
  /* rows 2 and 4 */
  for (int k = 0; k < d; k++) {
    double a2 = rs[k] - (ds[k] * dsv[k] - dyd[k]);
    double jx = 0.0;
    if (nnz_jd > 0)
      for (int p = jdrp[k]; p < jdrp[k + 1]; p++) jx += jdv[p] * dx[jdci[p]];
    double a4 = ryd[k] - (jx - dsv[k]);
    r2 += a2 * a2;
    r4 += a4 * a4;
  }
  free(t);
  double den = sqrt(nb);
  *resid2 = sqrt(r1 + r3) / den;
  *resid4 = sqrt(r1 + r2 + r3 + r4) / den;
}

/* ======================================================================
 * execute() -- HykktSolver.cpp:240-457 (+ check_error :846-930).
 * perm == NULL means the identity ordering.  Ds is passed as its value
 * vector (one diagonal entry per row, HykktSolver.cpp:660, 305-312).
 * ====================================================================== */
int orc_execute(const orc_opts* o,
                int n, const int* hrp, const int* hci, const double* hv,
                int d, const double* ds,
                int m, const int* jcrp, const int* jcci, const double* jcv,
                const int* jdrp, const int* jdci, const double* jdv, int nnz_jd,
                const double* rx, const double* rs, const double* ry, const double* ryd,
                const double* y0, const int* perm,
                double* dx, double* dsv, double* dyc, double* dyd,
                orc_stats* st, orc_intermediates* im)
{
  memset(st, 0, sizeof *st);
  st->pivot = -1;
  const int nnz_h = hrp[n];
  const int nnz_jc = jcrp[m];
  const int jd_flag = nnz_jd > 0;
  double t0 = now_s();
  double tsym = 0.0;

  /* ---- setup_parameters (:459-682): copies and Jc^T ---- */
  double* rx_til = (double*)xmalloc(sizeof(double) * (size_t)n);
  double* ryv = (double*)xmalloc(sizeof(double) * (size_t)m);
  memcpy(rx_til, rx, sizeof(double) * (size_t)n);
  memcpy(ryv, ry, sizeof(double) * (size_t)m);
  double* jcv_s = (double*)xmalloc(sizeof(double) * (size_t)nnz_jc);
  memcpy(jcv_s, jcv, sizeof(double) * (size_t)nnz_jc);
  int* jtrp = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
  int* jtci = (int*)xmalloc(sizeof(int) * (size_t)nnz_jc);
  double* jtv = (double*)xmalloc(sizeof(double) * (size_t)nnz_jc);
  orc_transpose(m, n, jcrp, jcci, jcv, jtrp, jtci, jtv);

  /* ---- compute_spgemm_htil (:289-337) ---- */
  int nnz_htil;
  int* htrp;
  int* htci;
  double* htv;
  if (jd_flag) {
    int* jdtrp = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
    int* jdtci = (int*)xmalloc(sizeof(int) * (size_t)nnz_jd);
    double* jdtv = (double*)xmalloc(sizeof(double) * (size_t)nnz_jd);
    orc_transpose(d, n, jdrp, jdci, jdv, jdtrp, jdtci, jdtv);
    double* jdvs = (double*)xmalloc(sizeof(double) * (size_t)nnz_jd);
    double* ryds = (double*)xmalloc(sizeof(double) * (size_t)d);
    for (int i = 0; i < d; i++) {                         /* row_scale :490-509 */
      for (int p = jdrp[i]; p < jdrp[i + 1]; p++) jdvs[p] = jdv[p] * ds[i];
      ryds[i] = ryd[i] * ds[i];
    }
    for (int i = 0; i < d; i++) ryds[i] += 1.0 * rs[i];   /* add_vecs :310-317 */
    orc_spmv(n, jdtrp, jdtci, jdtv, 1.0, ryds, 1.0, rx_til);
    double ts = now_s();
    nnz_htil = orc_spgemm_nnz(n, n, jdtrp, jdtci, jdrp, jdci, hrp, hci);
    htrp = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
    htci = (int*)xmalloc(sizeof(int) * (size_t)nnz_htil);
    orc_spgemm_pattern(n, n, jdtrp, jdtci, jdrp, jdci, hrp, hci, htrp, htci);
    tsym += now_s() - ts;
    htv = (double*)xmalloc(sizeof(double) * (size_t)nnz_htil);
    orc_spgemm_values(n, n, 1.0, 1.0, 1.0, jdtrp, jdtci, jdtv, jdrp, jdci, jdvs,
                      hrp, hci, hv, htrp, htci, htv);
    free(jdtrp); free(jdtci); free(jdtv); free(jdvs); free(ryds);
  } else {
    nnz_htil = nnz_h;
    htrp = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
    htci = (int*)xmalloc(sizeof(int) * (size_t)nnz_h);
    htv = (double*)xmalloc(sizeof(double) * (size_t)nnz_h);
    memcpy(htrp, hrp, sizeof(int) * ((size_t)n + 1));
    memcpy(htci, hci, sizeof(int) * (size_t)nnz_h);
    memcpy(htv, hv, sizeof(double) * (size_t)nnz_h);
  }
  st->nnz_htil = nnz_htil;
  if (im && im->htil_v) memcpy(im->htil_v, htv, sizeof(double) * (size_t)nnz_htil);

  /* ---- compute_ruiz_scaling (:340-349) ---- */
  double* max_d = (double*)xmalloc(sizeof(double) * (size_t)(n + m));
  orc_ruiz(o->ruiz_its, n, n + m, htrp, htci, htv, jtrp, jtci, jtv, jcrp, jcci, jcv_s,
           rx_til, ryv, max_d);
  if (im && im->max_d) memcpy(im->max_d, max_d, sizeof(double) * (size_t)(n + m));

  /* ---- compute_spgemm_hgamma (:370-383): Hg = Htil + gamma Jc^T Jc ---- */
  double ts = now_s();
  int nnz_hg = orc_spgemm_nnz(n, n, jtrp, jtci, jcrp, jcci, htrp, htci);
  int* hgrp = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
  int* hgci = (int*)xmalloc(sizeof(int) * (size_t)nnz_hg);
  orc_spgemm_pattern(n, n, jtrp, jtci, jcrp, jcci, htrp, htci, hgrp, hgci);
  tsym += now_s() - ts;
  double* hgv = (double*)xmalloc(sizeof(double) * (size_t)nnz_hg);
  orc_spgemm_values(n, n, o->gamma, 1.0, 1.0, jtrp, jtci, jtv, jcrp, jcci, jcv_s,
                    htrp, htci, htv, hgrp, hgci, hgv);
  st->nnz_hgam = nnz_hg;
  if (im && im->hgam_v) memcpy(im->hgam_v, hgv, sizeof(double) * (size_t)nnz_hg);
  double* rx_hat = (double*)xmalloc(sizeof(double) * (size_t)n);
  memcpy(rx_hat, rx_til, sizeof(double) * (size_t)n);
  orc_spmv(n, jtrp, jtci, jtv, o->gamma, ryv, 1.0, rx_hat);
  if (im && im->rx_hat) memcpy(im->rx_hat, rx_hat, sizeof(double) * (size_t)n);

  /* ---- setup_permutation / apply_permutation (:760-798, :386-395) ---- */
  ts = now_s();
  int* pm = (int*)xmalloc(sizeof(int) * (size_t)n);
  int* rev = (int*)xmalloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) pm[i] = perm ? perm[i] : i;
  orc_perm_rev(n, pm, rev);
  int* hgprp = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
  int* hgpci = (int*)xmalloc(sizeof(int) * (size_t)nnz_hg);
  int* hgmap = (int*)xmalloc(sizeof(int) * (size_t)nnz_hg);
  orc_map_rc(n, hgrp, hgci, pm, rev, hgprp, hgpci, hgmap);
  int* jcpci = (int*)xmalloc(sizeof(int) * (size_t)nnz_jc);
  int* jcmap = (int*)xmalloc(sizeof(int) * (size_t)nnz_jc);
  orc_map_c(m, jcrp, jcci, rev, jcpci, jcmap);
  int* jtprp = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
  int* jtpci = (int*)xmalloc(sizeof(int) * (size_t)nnz_jc);
  int* jtmap = (int*)xmalloc(sizeof(int) * (size_t)nnz_jc);
  orc_map_r(n, jtrp, jtci, pm, jtprp, jtpci, jtmap);
  orc_chol* ch = orc_chol_create(n, hgprp, hgpci);
  tsym += now_s() - ts;
  st->nnz_l = orc_chol_nnz(ch);
  double* hgpv = (double*)xmalloc(sizeof(double) * (size_t)nnz_hg);
  for (int q = 0; q < nnz_hg; q++) hgpv[q] = hgv[hgmap[q]];
  double* jcpv = (double*)xmalloc(sizeof(double) * (size_t)nnz_jc);
  double* jtpv = (double*)xmalloc(sizeof(double) * (size_t)nnz_jc);
  for (int q = 0; q < nnz_jc; q++) {
    jcpv[q] = jcv_s[jcmap[q]];
    jtpv[q] = jtv[jtmap[q]];
  }
  double* rxp = (double*)xmalloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) rxp[i] = rx_hat[pm[i]];

  /* ---- compute_hgamma_factorization (:398-401) ---- */
  int status = 0;
  int piv = orc_chol_factor(ch, hgpv, o->pivot_tol);
  st->pivot = piv;
  double* z = (double*)xmalloc(sizeof(double) * (size_t)n);
  double* schur = (double*)xmalloc(sizeof(double) * (size_t)m);
  double* yv = (double*)xmalloc(sizeof(double) * (size_t)m);
  if (piv >= 0) {
    status = 1;
    for (int i = 0; i < n; i++) dx[i] = NAN;
    for (int i = 0; i < m; i++) dyc[i] = NAN;
    for (int i = 0; i < d; i++) { dsv[i] = NAN; dyd[i] = NAN; }
    st->resid2 = st->resid4 = NAN;
  } else {
    /* ---- setup_conjugate_gradient (:814-843) ---- */
    orc_chol_solve(ch, rxp, z);
    memcpy(schur, ryv, sizeof(double) * (size_t)m);
    orc_spmv(m, jcrp, jcpci, jcpv, 1.0, z, -1.0, schur);
    if (im && im->schur_rhs) memcpy(im->schur_rhs, schur, sizeof(double) * (size_t)m);
    /* ---- compute_conjugate_gradient (:404-407) ---- */
    for (int i = 0; i < m; i++) yv[i] = y0 ? y0[i] : 0.0;
    double fn;
    st->cg_status = orc_schur_cg(m, n, jcrp, jcpci, jcpv, jtprp, jtpci, jtpv, ch, yv, schur,
                                 o->cg_tol, o->cg_itmax, &st->cg_iters, &fn);
    /* ---- recover_solution (:410-457) ---- */
    orc_spmv(n, jtprp, jtpci, jtpv, -1.0, yv, 1.0, rxp);
    orc_chol_solve(ch, rxp, z);
    for (int i = 0; i < n; i++) dx[i] = z[rev[i]];
    for (int i = 0; i < n; i++) dx[i] *= max_d[i];
    for (int i = 0; i < m; i++) dyc[i] = yv[i] * max_d[n + i];
    memcpy(dsv, ryd, sizeof(double) * (size_t)d);
    if (jd_flag) {
      orc_spmv(d, jdrp, jdci, jdv, 1.0, dx, -1.0, dsv);
    } else {
      for (int i = 0; i < d; i++) dsv[i] *= -1.0;
    }
    for (int i = 0; i < d; i++) {
      dyd[i] = dsv[i] * ds[i];
      dyd[i] += -1.0 * rs[i];
    }
    /* ---- check_error (:846-930), plus rows 2 and 4 ---- */
    orc_kkt_residual(n, hrp, hci, hv, d, ds, m, jcrp, jcci, jcv, jdrp, jdci, jdv, nnz_jd,
                     rx, rs, ry, ryd, dx, dsv, dyc, dyd, &st->resid2, &st->resid4);
    status = (st->resid2 < o->norm_tol) ? 0 : 1;
  }
  st->t_symbolic = tsym;
  st->t_numeric = now_s() - t0 - tsym;

  free(rx_til); free(ryv); free(jcv_s); free(jtrp); free(jtci); free(jtv);
  free(htrp); free(htci); free(htv); free(max_d);
  free(hgrp); free(hgci); free(hgv); free(rx_hat);
  free(pm); free(rev); free(hgprp); free(hgpci); free(hgmap);
  free(jcpci); free(jcmap); free(jtprp); free(jtpci); free(jtmap);
  free(hgpv); free(jcpv); free(jtpv); free(rxp); free(z); free(schur); free(yv);
  orc_chol_free(ch);
  return status;
}
