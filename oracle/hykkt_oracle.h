/*
 * hykkt_oracle.h -- CPU FP64 restatement of the HyKKT KKT solve.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (hykkt_amd/, include/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker / CPU baseline.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the ORNL/hykkt source tree).
 *
 * Pinning status (what checks each function against something independent):
 *   ingest (orc_mm_*, orc_*coo_to_csr, orc_read_rhs)  pinned: reference sources built
 *        under oracle/_ref + fixture closed-form counts.
 *   orc_map_rc/_c/_r, orc_perm_rev                      pinned: perm_driver.cpp KAT + _ref.
 *   orc_transpose, orc_spgemm_*                         pinned: dense brute force (tests).
 *   orc_ruiz                                           pinned: Ruiz_driver.cpp KAT (6 values).
 *   orc_chol_* (kind 0, up-looking)                    pinned: cuSolver_driver_cholesky.cpp KAT
 *        + dense numpy Cholesky / residual at small n.
 *   orc_chol_* (kind 1, supernodal)                    pinned: same KAT + agreement with
 *        kind 0 (solutions, nnz(L), flop count) on sparse SPD systems up to n = 4,000.
 *   orc_schur_cg                                       pinned: cuSolver_driver_schur_cg.cpp KAT
 *        + dense direct Schur solve.
 *   orc_execute                                        pinned: dense LU of the full 4x4 block
 *        KKT system at small n + residual bounds of the reference's CTest drivers.
 */
#ifndef HYKKT_ORACLE_H
#define HYKKT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

int  orc_threads(void);   /* OpenMP threads used by the parallel loops */

/* ---------------- MatrixMarket ingest (src/input_functions.cpp) ---------------- */
int  orc_mm_header(const char* path, int skip, int* n, int* m, int* nnz);
int  orc_mm_read_coo(const char* path, int skip, int cap, int* rows, int* cols, double* vals);
int  orc_sym_coo_nnz(int n, int nnz, const int* r, const int* c);
void orc_sym_coo_to_csr(int n, int nnz, const int* r, const int* c, const double* v,
                        int* rp, int* ci, double* cv);
void orc_coo_to_csr(int n, int nnz, const int* r, const int* c, const double* v,
                    int* rp, int* ci, double* cv);
int  orc_read_rhs(const char* path, int cap, double* out);

/* ---------------- sparse building blocks ---------------- */
void orc_transpose(int n, int m, const int* rp, const int* ci, const double* v,
                   int* trp, int* tci, double* tv);
void orc_spmv(int nrows, const int* rp, const int* ci, const double* v, double alpha,
              const double* x, double beta, double* y);
int  orc_spgemm_nnz(int n, int ncols, const int* arp, const int* aci, const int* brp,
                    const int* bci, const int* drp, const int* dci);
void orc_spgemm_pattern(int n, int ncols, const int* arp, const int* aci, const int* brp,
                        const int* bci, const int* drp, const int* dci, int* erp, int* eci);
void orc_spgemm_values(int n, int ncols, double alpha_p, double alpha_s, double beta_s,
                       const int* arp, const int* aci, const double* av,
                       const int* brp, const int* bci, const double* bv,
                       const int* drp, const int* dci, const double* dv,
                       const int* erp, const int* eci, double* ev);
void orc_ruiz(int its, int n, int totn,
              const int* hrp, const int* hci, double* hv,
              const int* jtrp, const int* jtci, double* jtv,
              const int* jrp, const int* jci, double* jv,
              double* rhs1, double* rhs2, double* max_d);

/* ---------------- permutation maps (src/permcheck.cu host part) ---------------- */
void orc_perm_rev(int n, const int* perm, int* rev);
void orc_map_rc(int n, const int* rows, const int* cols, const int* perm, const int* rev,
                int* prows, int* pcols, int* pmap);
void orc_map_c(int n, const int* rows, const int* cols, const int* rev, int* pcols, int* pmap);
void orc_map_r(int n, const int* rows, const int* cols, const int* perm, int* prows,
               int* pcols, int* pmap);

/* ---------------- sparse Cholesky of the lower triangle ---------------- */
typedef struct orc_chol orc_chol;
orc_chol* orc_chol_create(int n, const int* rp, const int* ci);
/* kind 0: up-looking (the form above); kind 1: supernodal left-looking on
 * OpenMP threads -- same factor to rounding, for the large systems. */
orc_chol* orc_chol_create_kind(int n, const int* rp, const int* ci, int kind);
int       orc_chol_factor(orc_chol* c, const double* vals, double tol);
void      orc_chol_solve(const orc_chol* c, const double* b, double* x);
long long orc_chol_nnz(const orc_chol* c);
double    orc_chol_flops(const orc_chol* c);
void      orc_chol_free(orc_chol* c);

/* ---------------- Schur-complement CG ---------------- */
int orc_schur_cg(int m, int n,
                 const int* jrp, const int* jci, const double* jv,
                 const int* jtrp, const int* jtci, const double* jtv,
                 const orc_chol* chol, double* x, const double* b,
                 double tol, int itmax, int* iters, double* final_norm);

/* ---------------- full pipeline (src/HykktSolver.cpp execute()) ---------------- */
typedef struct {
  double gamma;
  int    ruiz_its;
  double cg_tol;
  int    cg_itmax;
  double pivot_tol;
  double norm_tol;
  int    chol_kind;   /* orc_chol_create_kind() kind */
  int    refine_max;  /* iterative-refinement passes at most (0: none, as the reference) */
  double refine_tol;  /* refine while the 4-row residual is above this */
  double delta;       /* added to the diagonal of the scaled, permuted H_gamma
                         before factorization (add_diag hook, HykktSolver.cpp:392) */
} orc_opts;

typedef struct {
  double resid2;      /* reference 2-block-row relative residual */
  double resid4;      /* all 4 block rows */
  int    cg_iters;
  int    cg_status;   /* 0 converged, 1 itmax */
  int    pivot;       /* -1 ok, else first failing pivot (permuted index) */
  int    nnz_htil;
  int    nnz_hgam;
  long long nnz_l;
  double resid4_first; /* 4-row residual before refinement */
  int    refine_steps;  /* refinement passes kept */
  double t_symbolic;  /* seconds */
  double t_numeric;   /* seconds: everything a repeated solve does */
} orc_stats;

/* Optional intermediate outputs (any may be NULL). Array sizes: htil nnz_htil,
 * hgam nnz_hgam (natural order), maxd n+m, rxhat n, schur_rhs m. */
typedef struct {
  double* htil_v;
  double* hgam_v;
  double* max_d;
  double* rx_hat;
  double* schur_rhs;
} orc_intermediates;

int orc_execute(const orc_opts* o,
                int n, const int* hrp, const int* hci, const double* hv,
                int d, const double* ds,
                int m, const int* jcrp, const int* jcci, const double* jcv,
                const int* jdrp, const int* jdci, const double* jdv, int nnz_jd,
                const double* rx, const double* rs, const double* ry, const double* ryd,
                const double* y0, const int* perm,
                double* dx, double* dsv, double* dyc, double* dyd,
                orc_stats* st, orc_intermediates* im);

/* Residual of the full 4x4 block KKT system (HykktSolver.cpp:846-930 plus rows 2, 4). */
void orc_kkt_residual(int n, const int* hrp, const int* hci, const double* hv,
                      int d, const double* ds,
                      int m, const int* jcrp, const int* jcci, const double* jcv,
                      const int* jdrp, const int* jdci, const double* jdv, int nnz_jd,
                      const double* rx, const double* rs, const double* ry, const double* ryd,
                      const double* dx, const double* dsv, const double* dyc, const double* dyd,
                      double* resid2, double* resid4);

void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
