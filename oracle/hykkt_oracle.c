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
#ifdef _OPENMP
#include <omp.h>
#endif

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

/* Threads the OpenMP loops of this file run on. */
int orc_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

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
typedef struct orc_sn orc_sn;
static orc_sn* sn_create(int n, const int* rp, const int* ci, const int* parent,
                         long long* nnz, double* flops);
static int sn_factor(orc_sn* S, const double* vals, double tol);
static void sn_solve(const orc_sn* S, const double* b, double* x);
static void sn_free(orc_sn* S);

struct orc_chol {
  int n;
  int kind;         /* 0 up-looking, 1 supernodal left-looking */
  orc_sn* sn;
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
  return orc_chol_create_kind(n, rp, ci, 0);
}

orc_chol* orc_chol_create_kind(int n, const int* rp, const int* ci, int kind)
{
  orc_chol* c = (orc_chol*)xcalloc(1, sizeof(orc_chol));
  c->n = n;
  c->kind = kind;
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
  if (kind == 1) {
    c->sn = sn_create(n, rp, ci, c->parent, &c->nnz, &c->flops);
    return c;
  }
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
  if (c->kind == 1) return sn_factor(c->sn, vals, tol);
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
  if (c->kind == 1) {
    sn_solve(c->sn, b, x);
    return;
  }
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
  sn_free(c->sn);
  free(c->parent);
  free(c->lp);
  free(c->li);
  free(c->lx);
  free(c);
}

/* ======================================================================
 * Supernodal left-looking form of the same factorization: same matrix,
 * same ordering, another operation order, so the factor agrees with the
 * up-looking form to rounding.  It exists for speed (the 10,000-bus systems
 * take minutes up-looking).  Textbook construction (Liu 1990; Ng & Peyton
 * 1993): columns are relabelled by a postorder of the elimination tree,
 * which only relabels L; a supernode is a maximal run j, j+1, ... with
 * parent(j) = j+1 and colcount(j) = colcount(j+1) + 1, so all its columns
 * share one row list (own columns first).  Each supernode's panel is a dense
 * f x nc column-major block over that list.  Descendants are applied
 * left-looking through per-supernode link lists; panels are factored in
 * 64-column blocks; the dense products run on OpenMP threads.
 * ====================================================================== */
struct orc_sn {
  int n, nsn;
  int* post;        /* new label -> input label */
  int* sn_first;    /* nsn + 1 */
  int* sn_of;       /* new column -> supernode */
  long long* rowp;  /* nsn + 1 */
  int* rows;        /* row lists, new labels */
  long long* lxp;   /* nsn + 1 */
  double* lx;
  int* acp;         /* lower triangle of the relabelled matrix by column */
  int* ari;
  int* asrc;        /* position of the entry in the input value array */
  int maxf;
};

static void sn_free(orc_sn* S)
{
  if (!S) return;
  free(S->post); free(S->sn_first); free(S->sn_of); free(S->rowp); free(S->rows);
  free(S->lxp); free(S->lx); free(S->acp); free(S->ari); free(S->asrc);
  free(S);
}

/* Reach of row k in the relabelled tree `par`, over the row's entries
 * cols[0..len) (all < k); pattern to s, returns its size. */
static int sn_reach(const int* par, int k, const int* cols, int len, int* mark, int* s)
{
  int top = 0;
  mark[k] = k;
  for (int q = 0; q < len; q++) {
    int i = cols[q];
    while (i >= 0 && mark[i] != k) {
      s[top++] = i;
      mark[i] = k;
      i = par[i];
    }
  }
  return top;
}

static orc_sn* sn_create(int n, const int* rp, const int* ci, const int* parent,
                         long long* nnz_out, double* flops_out)
{
  orc_sn* S = (orc_sn*)xcalloc(1, sizeof(orc_sn));
  S->n = n;
  size_t N = (size_t)(n > 0 ? n : 1);
  /* postorder (children visited in increasing label order) */
  int* head = (int*)xmalloc(sizeof(int) * N);
  int* next = (int*)xmalloc(sizeof(int) * N);
  int* stack = (int*)xmalloc(sizeof(int) * N);
  int* post = (int*)xmalloc(sizeof(int) * N);
  int* ipost = (int*)xmalloc(sizeof(int) * N);
  for (int j = 0; j < n; j++) head[j] = -1;
  for (int j = n - 1; j >= 0; j--)
    if (parent[j] >= 0) { next[j] = head[parent[j]]; head[parent[j]] = j; }
  int k = 0;
  for (int r = 0; r < n; r++) {
    if (parent[r] >= 0) continue;
    int top = 0;
    stack[0] = r;
    while (top >= 0) {
      int p = stack[top], c = head[p];
      if (c == -1) { top--; post[k++] = p; }
      else { head[p] = next[c]; stack[++top] = c; }
    }
  }
  for (int j = 0; j < n; j++) ipost[post[j]] = j;
  int* par = (int*)xmalloc(sizeof(int) * N);
  for (int j = 0; j < n; j++) par[j] = parent[post[j]] < 0 ? -1 : ipost[parent[post[j]]];
  /* lower-triangle entries (input labels c <= r) in new labels: row
   * i = ipost[r], column j = ipost[c] <= i (c is a tree descendant of r). */
  int* brp = (int*)xcalloc(N + 1, sizeof(int));
  for (int r = 0; r < n; r++)
    for (int p = rp[r]; p < rp[r + 1]; p++)
      if (ci[p] <= r) brp[ipost[r] + 1]++;
  for (int i = 0; i < n; i++) brp[i + 1] += brp[i];
  int nb = brp[n];
  int* bci = (int*)xmalloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  int* bsrc = (int*)xmalloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  int* fill = (int*)xmalloc(sizeof(int) * N);
  for (int i = 0; i < n; i++) fill[i] = brp[i];
  for (int r = 0; r < n; r++)
    for (int p = rp[r]; p < rp[r + 1]; p++)
      if (ci[p] <= r) {
        int q = fill[ipost[r]]++;
        bci[q] = ipost[ci[p]];
        bsrc[q] = p;
      }
  /* the same entries by column, for assembly */
  S->acp = (int*)xcalloc(N + 1, sizeof(int));
  S->ari = (int*)xmalloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  S->asrc = (int*)xmalloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  for (int q = 0; q < nb; q++) S->acp[bci[q] + 1]++;
  for (int j = 0; j < n; j++) S->acp[j + 1] += S->acp[j];
  for (int j = 0; j < n; j++) fill[j] = S->acp[j];
  for (int i = 0; i < n; i++)
    for (int q = brp[i]; q < brp[i + 1]; q++) {
      int t = fill[bci[q]]++;
      S->ari[t] = i;
      S->asrc[t] = bsrc[q];
    }
  /* column counts from the row reaches; strictly-lower entries of each row
   * are passed to the reach (the diagonal is counted separately) */
  long long* cnt = (long long*)xcalloc(N, sizeof(long long));
  int* mark = (int*)xmalloc(sizeof(int) * N);
  int* sp = (int*)xmalloc(sizeof(int) * N);
  int* cols = (int*)xmalloc(sizeof(int) * N);
  for (int j = 0; j < n; j++) mark[j] = -1;
  for (int i = 0; i < n; i++) {
    int len = 0;
    for (int q = brp[i]; q < brp[i + 1]; q++)
      if (bci[q] < i) cols[len++] = bci[q];
    int top = sn_reach(par, i, cols, len, mark, sp);
    for (int t = 0; t < top; t++) cnt[sp[t]]++;
    cnt[i]++;
  }
  long long nnz = 0;
  double fl = 0.0;
  for (int j = 0; j < n; j++) {
    nnz += cnt[j];
    fl += (double)cnt[j] * (double)cnt[j];
  }
  *nnz_out = nnz;
  *flops_out = fl;
  /* supernodes */
  S->sn_first = (int*)xmalloc(sizeof(int) * (N + 1));
  S->sn_of = (int*)xmalloc(sizeof(int) * N);
  int ns = 0;
  for (int j = 0; j < n; j++) {
    if (!(j > 0 && par[j - 1] == j && cnt[j - 1] == cnt[j] + 1)) S->sn_first[ns++] = j;
    S->sn_of[j] = ns - 1;
  }
  S->sn_first[ns] = n;
  S->nsn = ns;
  S->rowp = (long long*)xmalloc(sizeof(long long) * ((size_t)ns + 1));
  S->lxp = (long long*)xmalloc(sizeof(long long) * ((size_t)ns + 1));
  S->rowp[0] = 0;
  S->lxp[0] = 0;
  S->maxf = 0;
  for (int s = 0; s < ns; s++) {
    int a = S->sn_first[s], nc = S->sn_first[s + 1] - a;
    long long f = cnt[a];
    if (f > S->maxf) S->maxf = (int)f;
    S->rowp[s + 1] = S->rowp[s] + f;
    S->lxp[s + 1] = S->lxp[s] + f * nc;
  }
  /* row lists: the diagonal first, then every later row whose reach holds
   * the supernode's first column (ascending, as rows are visited in order) */
  S->rows = (int*)xmalloc(sizeof(int) * (size_t)(S->rowp[ns] > 0 ? S->rowp[ns] : 1));
  long long* pos = (long long*)xmalloc(sizeof(long long) * ((size_t)ns + 1));
  for (int s = 0; s < ns; s++) {
    S->rows[S->rowp[s]] = S->sn_first[s];
    pos[s] = S->rowp[s] + 1;
  }
  for (int j = 0; j < n; j++) mark[j] = -1;
  for (int i = 0; i < n; i++) {
    int len = 0;
    for (int q = brp[i]; q < brp[i + 1]; q++)
      if (bci[q] < i) cols[len++] = bci[q];
    int top = sn_reach(par, i, cols, len, mark, sp);
    for (int t = 0; t < top; t++) {
      int j = sp[t], s = S->sn_of[j];
      if (S->sn_first[s] == j) S->rows[pos[s]++] = i;
    }
  }
  for (int s = 0; s < ns; s++)
    if (pos[s] != S->rowp[s + 1]) {
      fprintf(stderr, "oracle: supernode %d row list %lld != colcount %lld\n", s,
              pos[s] - S->rowp[s], S->rowp[s + 1] - S->rowp[s]);
      abort();
    }
  S->lx = (double*)xmalloc(sizeof(double) * (size_t)(S->lxp[ns] > 0 ? S->lxp[ns] : 1));
  S->post = post;
  free(pos); free(cnt); free(mark); free(sp); free(cols); free(fill);
  free(brp); free(bci); free(bsrc); free(par); free(ipost);
  free(head); free(next); free(stack);
  return S;
}

/* C[0:8, 0:4] -= sum_k A[0:8, k] B[0:4, k] (column-major, leading dims). */
static inline void sn_mk84(int K, const double* A, int lda, const double* B, int ldb, double* C,
                           int ldc)
{
  double acc[4][8];
  for (int j = 0; j < 4; j++)
    for (int i = 0; i < 8; i++) acc[j][i] = 0.0;
  for (int k = 0; k < K; k++) {
    const double* a = A + (size_t)k * lda;
    const double* b = B + (size_t)k * ldb;
    for (int j = 0; j < 4; j++) {
      double bj = b[j];
      for (int i = 0; i < 8; i++) acc[j][i] += a[i] * bj;
    }
  }
  for (int j = 0; j < 4; j++)
    for (int i = 0; i < 8; i++) C[i + (size_t)j * ldc] -= acc[j][i];
}

/* One tile: C[i0:i1, c0:c1] -= A[i0:i1, :] B[c0:c1, :]^T, K columns. */
static void sn_tile(int i0, int i1, int c0, int c1, int K, const double* A, int lda,
                    const double* B, int ldb, double* C, int ldc)
{
  const int KB = 256;
  for (int k0 = 0; k0 < K; k0 += KB) {
    int kb = K - k0 < KB ? K - k0 : KB;
    const double* Ak = A + (size_t)k0 * lda;
    const double* Bk = B + (size_t)k0 * ldb;
    int c = c0;
    for (; c + 4 <= c1; c += 4) {
      int i = i0;
      for (; i + 8 <= i1; i += 8) sn_mk84(kb, Ak + i, lda, Bk + c, ldb, C + i + (size_t)c * ldc, ldc);
      for (; i < i1; i++)
        for (int cc = c; cc < c + 4; cc++) {
          double t = 0.0;
          for (int k = 0; k < kb; k++) t += Ak[i + (size_t)k * lda] * Bk[cc + (size_t)k * ldb];
          C[i + (size_t)cc * ldc] -= t;
        }
    }
    for (; c < c1; c++)
      for (int i = i0; i < i1; i++) {
        double t = 0.0;
        for (int k = 0; k < kb; k++) t += Ak[i + (size_t)k * lda] * Bk[c + (size_t)k * ldb];
        C[i + (size_t)c * ldc] -= t;
      }
  }
}

/* C -= A B^T over the m x n block of C, skipping 64 x 64 tiles that lie
 * wholly above the diagonal (rows and columns of C share one origin). */
static void sn_gemm_lower(int m, int n, int K, const double* A, int lda, const double* B, int ldb,
                          double* C, int ldc)
{
  if (m <= 0 || n <= 0 || K <= 0) return;
  const int T = 64;
  int mt = (m + T - 1) / T, nt = (n + T - 1) / T;
  long long work = (long long)m * n * K;
#pragma omp parallel for schedule(dynamic, 1) collapse(2) if (work > 2000000)
  for (int tc = 0; tc < nt; tc++)
    for (int tr = 0; tr < mt; tr++) {
      int i0 = tr * T, i1 = i0 + T < m ? i0 + T : m;
      int c0 = tc * T, c1 = c0 + T < n ? c0 + T : n;
      if (i1 <= c0) continue;
      sn_tile(i0, i1, c0, c1, K, A, lda, B, ldb, C, ldc);
    }
}

/* Factor the panel P (f x nc, leading dim f) in place; returns the local
 * column of the first failing pivot or -1. */
static int sn_panel(double* P, int f, int nc, double tol)
{
  const int PB = 64;
  for (int j0 = 0; j0 < nc; j0 += PB) {
    int j1 = j0 + PB < nc ? j0 + PB : nc;
    for (int j = j0; j < j1; j++) {
      double d = P[j + (size_t)j * f];
      if (!(d > 0.0) || !(sqrt(d) > tol)) return j;
      d = sqrt(d);
      P[j + (size_t)j * f] = d;
      for (int i = j + 1; i < j1; i++) P[i + (size_t)j * f] /= d;
      for (int c = j + 1; c < j1; c++) {
        double lcj = P[c + (size_t)j * f];
        for (int i = c; i < j1; i++) P[i + (size_t)c * f] -= P[i + (size_t)j * f] * lcj;
      }
    }
    int rows = f - j1;
    if (rows > 0) {
      int chunks = (rows + 63) / 64;
#pragma omp parallel for schedule(dynamic, 1) if ((long long)rows * (j1 - j0) > 20000)
      for (int q = 0; q < chunks; q++) {
        int r0 = j1 + q * 64, r1 = r0 + 64 < f ? r0 + 64 : f;
        for (int j = j0; j < j1; j++) {
          double* pj = P + (size_t)j * f;
          for (int k = j0; k < j; k++) {
            double ljk = P[j + (size_t)k * f];
            const double* pk = P + (size_t)k * f;
            for (int i = r0; i < r1; i++) pj[i] -= pk[i] * ljk;
          }
          double djj = pj[j];
          for (int i = r0; i < r1; i++) pj[i] /= djj;
        }
      }
    }
    if (j1 < nc)
      sn_gemm_lower(f - j1, nc - j1, j1 - j0, P + j1 + (size_t)j0 * f, f, P + j1 + (size_t)j0 * f,
                    f, P + j1 + (size_t)j1 * f, f);
  }
  return -1;
}

static int sn_factor(orc_sn* S, const double* vals, double tol)
{
  const int n = S->n, ns = S->nsn;
  const int CB = 128;
  memset(S->lx, 0, sizeof(double) * (size_t)(S->lxp[ns] > 0 ? S->lxp[ns] : 1));
  int* relpos = (int*)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int* headl = (int*)xmalloc(sizeof(int) * (size_t)(ns > 0 ? ns : 1));
  int* nextl = (int*)xmalloc(sizeof(int) * (size_t)(ns > 0 ? ns : 1));
  long long* ptr = (long long*)xmalloc(sizeof(long long) * (size_t)(ns > 0 ? ns : 1));
  double* W = (double*)xmalloc(sizeof(double) * (size_t)(S->maxf > 0 ? S->maxf : 1) * CB);
  for (int s = 0; s < ns; s++) headl[s] = -1;
  int fail = -1;
  for (int s = 0; s < ns && fail < 0; s++) {
    const int a = S->sn_first[s], b = S->sn_first[s + 1], nc = b - a;
    const int f = (int)(S->rowp[s + 1] - S->rowp[s]);
    const int* R = S->rows + S->rowp[s];
    double* P = S->lx + S->lxp[s];
    for (int t = 0; t < f; t++) relpos[R[t]] = t;
    for (int j = a; j < b; j++)
      for (int q = S->acp[j]; q < S->acp[j + 1]; q++)
        P[relpos[S->ari[q]] + (size_t)(j - a) * f] = vals[S->asrc[q]];
    int D = headl[s];
    headl[s] = -1;
    while (D >= 0) {
      const int Dn = nextl[D];
      const int fD = (int)(S->rowp[D + 1] - S->rowp[D]);
      const int ncD = S->sn_first[D + 1] - S->sn_first[D];
      const int* RD = S->rows + S->rowp[D];
      const double* PD = S->lx + S->lxp[D];
      const int p1 = (int)ptr[D];
      int p2 = p1;
      while (p2 < fD && RD[p2] < b) p2++;
      for (int c0 = p1; c0 < p2; c0 += CB) {
        const int c1 = c0 + CB < p2 ? c0 + CB : p2, nw = c1 - c0, mw = fD - c0;
        memset(W, 0, sizeof(double) * (size_t)mw * nw);
        sn_gemm_lower(mw, nw, ncD, PD + c0, fD, PD + c0, fD, W, mw);
#pragma omp parallel for schedule(static) if ((long long)mw * nw > 50000)
        for (int c = 0; c < nw; c++) {
          double* pc = P + (size_t)(RD[c0 + c] - a) * f;
          const double* wc = W + (size_t)c * mw;
          for (int i = c; i < mw; i++) pc[relpos[RD[c0 + i]]] += wc[i];
        }
      }
      ptr[D] = p2;
      if (p2 < fD) {
        const int s2 = S->sn_of[RD[p2]];
        nextl[D] = headl[s2];
        headl[s2] = D;
      }
      D = Dn;
    }
    const int bad = sn_panel(P, f, nc, tol);
    if (bad >= 0) { fail = S->post[a + bad]; break; }
    if (f > nc) {
      ptr[s] = nc;
      const int s2 = S->sn_of[R[nc]];
      nextl[s] = headl[s2];
      headl[s2] = s;
    }
  }
  free(relpos); free(headl); free(nextl); free(ptr); free(W);
  return fail;
}

static void sn_solve(const orc_sn* S, const double* b, double* x)
{
  const int n = S->n, ns = S->nsn;
  double* y = (double*)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) y[j] = b[S->post[j]];
  for (int s = 0; s < ns; s++) {
    const int a = S->sn_first[s], nc = S->sn_first[s + 1] - a;
    const int f = (int)(S->rowp[s + 1] - S->rowp[s]);
    const int* R = S->rows + S->rowp[s];
    const double* P = S->lx + S->lxp[s];
    for (int j = 0; j < nc; j++) {
      const double* pj = P + (size_t)j * f;
      y[a + j] /= pj[j];
      const double yj = y[a + j];
      for (int i = j + 1; i < f; i++) y[R[i]] -= pj[i] * yj;
    }
  }
  for (int s = ns - 1; s >= 0; s--) {
    const int a = S->sn_first[s], nc = S->sn_first[s + 1] - a;
    const int f = (int)(S->rowp[s + 1] - S->rowp[s]);
    const int* R = S->rows + S->rowp[s];
    const double* P = S->lx + S->lxp[s];
    for (int j = nc - 1; j >= 0; j--) {
      const double* pj = P + (size_t)j * f;
      double t = y[a + j];
      for (int i = j + 1; i < f; i++) t -= pj[i] * y[R[i]];
      y[a + j] = t / pj[j];
    }
  }
  for (int j = 0; j < n; j++) x[S->post[j]] = y[j];
  free(y);
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
/* The residual rows themselves (ex n, es d, ey m, ed d; any may be NULL) and
 * the two relative norms. */
static void kkt_resid_vec(int n, const int* hrp, const int* hci, const double* hv,
                          int d, const double* ds,
                          int m, const int* jcrp, const int* jcci, const double* jcv,
                          const int* jdrp, const int* jdci, const double* jdv, int nnz_jd,
                          const double* rx, const double* rs, const double* ry,
                          const double* ryd, const double* dx, const double* dsv,
                          const double* dyc, const double* dyd, double* ex, double* es,
                          double* ey, double* ed, double* resid2, double* resid4)
{
  double* t = (double*)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
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
  if (ex) memcpy(ex, t, sizeof(double) * (size_t)n);
  /* row 3 */
  for (int k = 0; k < m; k++) {
    double acc = ry[k];
    for (int p = jcrp[k]; p < jcrp[k + 1]; p++) acc -= jcv[p] * dx[jcci[p]];
    r3 += acc * acc;
    if (ey) ey[k] = acc;
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
    if (es) es[k] = a2;
    if (ed) ed[k] = a4;
  }
  free(t);
  double den = sqrt(nb);
  *resid2 = sqrt(r1 + r3) / den;
  *resid4 = sqrt(r1 + r2 + r3 + r4) / den;
}

void orc_kkt_residual(int n, const int* hrp, const int* hci, const double* hv,
                      int d, const double* ds,
                      int m, const int* jcrp, const int* jcci, const double* jcv,
                      const int* jdrp, const int* jdci, const double* jdv, int nnz_jd,
                      const double* rx, const double* rs, const double* ry, const double* ryd,
                      const double* dx, const double* dsv, const double* dyc, const double* dyd,
                      double* resid2, double* resid4)
{
  kkt_resid_vec(n, hrp, hci, hv, d, ds, m, jcrp, jcci, jcv, jdrp, jdci, jdv, nnz_jd, rx, rs, ry,
                ryd, dx, dsv, dyc, dyd, NULL, NULL, NULL, NULL, resid2, resid4);
}

/* ======================================================================
 * execute() -- HykktSolver.cpp:240-457 (+ check_error :846-930).
 * perm == NULL means the identity ordering.  Ds is passed as its value
 * vector (one diagonal entry per row, HykktSolver.cpp:660, 305-312).
 * ====================================================================== */
/* Everything the solve phase of execute() needs once the factor exists. */
typedef struct {
  int n, m, d, jd_flag;
  const int *jdrp, *jdci;
  const double *jdv, *ds;
  const int *jtrp, *jtci;     /* scaled Jc^T, natural order */
  const double* jtv;
  const int *jcrp, *jcpci;    /* scaled Jc, permuted columns */
  const double* jcpv;
  const int *jtprp, *jtpci;   /* scaled Jc^T, permuted rows */
  const double* jtpv;
  const int *pm, *rev;
  const double* max_d;
  double gamma, cg_tol;
  int cg_itmax;
  const orc_chol* ch;
} solve_ctx;

/* rx_hat = rxt + gamma Jc^T ysc (:370-383), permuted (:386-395);
 * setup_conjugate_gradient (:814-843); CG (:404-407); recover_solution
 * (:410-457).  rxt, ysc: Ruiz-scaled right-hand sides; rs, ryd: unscaled. */
static int solve_phase(const solve_ctx* c, const double* rxt, const double* ysc,
                       const double* rs, const double* ryd, const double* y0,
                       double* dx, double* dsv, double* dyc, double* dyd, double* schur_out,
                       int* iters)
{
  const int n = c->n, m = c->m, d = c->d;
  double* rx_hat = (double*)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double* rxp = (double*)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double* z = (double*)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double* schur = (double*)xmalloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* yv = (double*)xmalloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  memcpy(rx_hat, rxt, sizeof(double) * (size_t)n);
  orc_spmv(n, c->jtrp, c->jtci, c->jtv, c->gamma, ysc, 1.0, rx_hat);
  for (int i = 0; i < n; i++) rxp[i] = rx_hat[c->pm[i]];
  orc_chol_solve(c->ch, rxp, z);
  memcpy(schur, ysc, sizeof(double) * (size_t)m);
  orc_spmv(m, c->jcrp, c->jcpci, c->jcpv, 1.0, z, -1.0, schur);
  if (schur_out) memcpy(schur_out, schur, sizeof(double) * (size_t)m);
  for (int i = 0; i < m; i++) yv[i] = y0 ? y0[i] : 0.0;
  double fn;
  int cg_status = orc_schur_cg(m, n, c->jcrp, c->jcpci, c->jcpv, c->jtprp, c->jtpci, c->jtpv,
                               c->ch, yv, schur, c->cg_tol, c->cg_itmax, iters, &fn);
  orc_spmv(n, c->jtprp, c->jtpci, c->jtpv, -1.0, yv, 1.0, rxp);
  orc_chol_solve(c->ch, rxp, z);
  for (int i = 0; i < n; i++) dx[i] = z[c->rev[i]];
  for (int i = 0; i < n; i++) dx[i] *= c->max_d[i];
  for (int i = 0; i < m; i++) dyc[i] = yv[i] * c->max_d[n + i];
  memcpy(dsv, ryd, sizeof(double) * (size_t)d);
  if (c->jd_flag) {
    orc_spmv(d, c->jdrp, c->jdci, c->jdv, 1.0, dx, -1.0, dsv);
  } else {
    for (int i = 0; i < d; i++) dsv[i] *= -1.0;
  }
  for (int i = 0; i < d; i++) {
    dyd[i] = dsv[i] * c->ds[i];
    dyd[i] += -1.0 * rs[i];
  }
  free(rx_hat); free(rxp); free(z); free(schur); free(yv);
  return cg_status;
}

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
  orc_chol* ch = orc_chol_create_kind(n, hgprp, hgpci, o->chol_kind);
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

  /* add_diag(delta) (:392): the reference passes zero; a caller's delta
     regularizes the factorized matrix, H_gamma + delta I in scaled terms */
  if (o->delta != 0.0)
    for (int i = 0; i < n; i++)
      for (int q = hgprp[i]; q < hgprp[i + 1]; q++)
        if (hgpci[q] == i) hgpv[q] += o->delta;

  /* ---- compute_hgamma_factorization (:398-401) ---- */
  int status = 0;
  int piv = orc_chol_factor(ch, hgpv, o->pivot_tol);
  st->pivot = piv;
  if (piv >= 0) {
    status = 1;
    for (int i = 0; i < n; i++) dx[i] = NAN;
    for (int i = 0; i < m; i++) dyc[i] = NAN;
    for (int i = 0; i < d; i++) { dsv[i] = NAN; dyd[i] = NAN; }
    st->resid2 = st->resid4 = NAN;
  } else {
    solve_ctx c = {n, m, d, jd_flag, jdrp, jdci, jdv, ds, jtrp, jtci, jtv, jcrp, jcpci, jcpv,
                   jtprp, jtpci, jtpv, pm, rev, max_d, o->gamma, o->cg_tol, o->cg_itmax, ch};
    st->cg_status = solve_phase(&c, rx_til, ryv, rs, ryd, y0, dx, dsv, dyc, dyd,
                                im ? im->schur_rhs : NULL, &st->cg_iters);
    /* ---- check_error (:846-930), plus rows 2 and 4 ---- */
    orc_kkt_residual(n, hrp, hci, hv, d, ds, m, jcrp, jcci, jcv, jdrp, jdci, jdv, nnz_jd,
                     rx, rs, ry, ryd, dx, dsv, dyc, dyd, &st->resid2, &st->resid4);
    st->resid4_first = st->resid4;
    /* ---- iterative refinement on the 4x4 block KKT system (an addition:
     * the reference stops after one pass).  The residual rows go through the
     * same right-hand-side transform (r~x = ex + Jd^T (Ds ed + es), Ruiz
     * scaling by max_d) and solve phase; a correction is kept only if it
     * lowers the 4-row residual. ---- */
    if (o->refine_max > 0 && st->resid4 > o->refine_tol) {
      size_t N = (size_t)(n > 0 ? n : 1), M = (size_t)(m > 0 ? m : 1), D = (size_t)(d > 0 ? d : 1);
      double* ex = (double*)xmalloc(sizeof(double) * N);
      double* es = (double*)xmalloc(sizeof(double) * D);
      double* ey = (double*)xmalloc(sizeof(double) * M);
      double* ed = (double*)xmalloc(sizeof(double) * D);
      double* t = (double*)xmalloc(sizeof(double) * D);
      double* cx = (double*)xmalloc(sizeof(double) * N);
      double* cs = (double*)xmalloc(sizeof(double) * D);
      double* cy = (double*)xmalloc(sizeof(double) * M);
      double* cd = (double*)xmalloc(sizeof(double) * D);
      int* jdtrp = NULL;
      int* jdtci = NULL;
      double* jdtv = NULL;
      if (jd_flag) {
        jdtrp = (int*)xmalloc(sizeof(int) * (N + 1));
        jdtci = (int*)xmalloc(sizeof(int) * (size_t)nnz_jd);
        jdtv = (double*)xmalloc(sizeof(double) * (size_t)nnz_jd);
        orc_transpose(d, n, jdrp, jdci, jdv, jdtrp, jdtci, jdtv);
      }
      for (int k = 0; k < o->refine_max && st->resid4 > o->refine_tol; k++) {
        double r2, r4;
        kkt_resid_vec(n, hrp, hci, hv, d, ds, m, jcrp, jcci, jcv, jdrp, jdci, jdv, nnz_jd, rx,
                      rs, ry, ryd, dx, dsv, dyc, dyd, ex, es, ey, ed, &r2, &r4);
        if (jd_flag) {
          for (int i = 0; i < d; i++) t[i] = ed[i] * ds[i] + es[i];
          orc_spmv(n, jdtrp, jdtci, jdtv, 1.0, t, 1.0, ex);
        }
        for (int i = 0; i < n; i++) ex[i] *= max_d[i];
        for (int i = 0; i < m; i++) ey[i] *= max_d[n + i];
        int it;
        solve_phase(&c, ex, ey, es, ed, NULL, cx, cs, cy, cd, NULL, &it);
        for (int i = 0; i < n; i++) cx[i] += dx[i];
        for (int i = 0; i < d; i++) { cs[i] += dsv[i]; cd[i] += dyd[i]; }
        for (int i = 0; i < m; i++) cy[i] += dyc[i];
        kkt_resid_vec(n, hrp, hci, hv, d, ds, m, jcrp, jcci, jcv, jdrp, jdci, jdv, nnz_jd, rx,
                      rs, ry, ryd, cx, cs, cy, cd, NULL, NULL, NULL, NULL, &r2, &r4);
        if (!(r4 < st->resid4)) break;
        memcpy(dx, cx, sizeof(double) * (size_t)n);
        memcpy(dsv, cs, sizeof(double) * (size_t)d);
        memcpy(dyc, cy, sizeof(double) * (size_t)m);
        memcpy(dyd, cd, sizeof(double) * (size_t)d);
        st->resid2 = r2;
        st->resid4 = r4;
        st->refine_steps++;
      }
      free(ex); free(es); free(ey); free(ed); free(t);
      free(cx); free(cs); free(cy); free(cd);
      free(jdtrp); free(jdtci); free(jdtv);
    }
    status = (st->resid2 < o->norm_tol) ? 0 : 1;
  }
  st->t_symbolic = tsym;
  st->t_numeric = now_s() - t0 - tsym;

  free(rx_til); free(ryv); free(jcv_s); free(jtrp); free(jtci); free(jtv);
  free(htrp); free(htci); free(htv); free(max_d);
  free(hgrp); free(hgci); free(hgv); free(rx_hat);
  free(pm); free(rev); free(hgprp); free(hgpci); free(hgmap);
  free(jcpci); free(jcmap); free(jtprp); free(jtpci); free(jtmap);
  free(hgpv); free(jcpv); free(jtpv);
  orc_chol_free(ch);
  return status;
}
