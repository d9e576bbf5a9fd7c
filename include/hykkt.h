/*
 * hykkt.h -- C ABI of the MI355X-native HyKKT KKT solver.
 *
 * Solves the 4x4 block KKT system of ORNL/hykkt (README.md:11-26)
 *
 *   [H   0   Jc^T  Jd^T] [dx ]   [rx ]
 *   [0   Ds  0     -I  ] [ds ] = [rs ]
 *   [Jc  0   0     0   ] [dyc]   [ry ]
 *   [Jd  -I  0     0   ] [dyd]   [ryd]
 *
 * by the reference's method: eliminate ds/dyd, Ruiz-equilibrate, gamma-augment,
 * permute, Cholesky-factor H_gamma, solve the Schur complement by CG, recover.
 * All matrices are CSR, 0-based, int32 indices, FP64 values.
 *
 * Every function returns 0 (HYKKT_OK) on success, 1 where the reference's
 * int execute() returns 1, and a negative error code otherwise (the reference
 * terminates the process instead, cuda_check_errors.hpp:3-14).  The message
 * of the last failing call on the calling thread is hykkt_last_error().
 */
#ifndef HYKKT_H
#define HYKKT_H

#ifdef __cplusplus
extern "C" {
#endif

#define HYKKT_OK            0
#define HYKKT_RESIDUAL_FAIL 1
#define HYKKT_E_ARG        -1
#define HYKKT_E_STATE      -2
#define HYKKT_E_HIP        -3
#define HYKKT_E_PATTERN    -4
#define HYKKT_E_PIVOT      -5
#define HYKKT_E_IO         -6
#define HYKKT_E_NOMEM      -7

typedef struct hykkt_solver_s* hykkt_solver;

/* Library version string. */
const char* hykkt_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* hykkt_last_error(void);

/* ---------------------------------------------------------------------------
 * HykktSolver façade (src/HykktSolver.hpp:17-137).
 * ------------------------------------------------------------------------- */

/* HykktSolver::HykktSolver(double gamma), HykktSolver.cpp:18-34. */
int hykkt_create(double gamma, hykkt_solver* out);
/* HykktSolver::~HykktSolver(), HykktSolver.cpp:36-105. */
int hykkt_destroy(hykkt_solver s);

/* set_H / set_Ds / set_Jc / set_Jd, HykktSolver.cpp:163-192 and :932-949.
 * Pointers are borrowed (not copied) and read at every hykkt_execute, exactly
 * as the reference reads them at every execute() (HykktSolver.cpp:635-666).
 * H is n x n, symmetric, both triangles stored, sorted columns.  Ds is d x d
 * with exactly one diagonal entry per row.  Jc is m x n, Jd is d x n.
 * With the option inputs_on_device the pointers are device pointers. */
int hykkt_set_H(hykkt_solver s, const int* rowptrs, const int* colidx, const double* vals,
                int n, int m, int nnz);
int hykkt_set_Ds(hykkt_solver s, const int* rowptrs, const int* colidx, const double* vals,
                 int n, int m, int nnz);
int hykkt_set_Jc(hykkt_solver s, const int* rowptrs, const int* colidx, const double* vals,
                 int n, int m, int nnz);
int hykkt_set_Jd(hykkt_solver s, const int* rowptrs, const int* colidx, const double* vals,
                 int n, int m, int nnz);

/* set_rx / set_rs / set_ry / set_ryd, HykktSolver.cpp:194-212 (borrowed). */
int hykkt_set_rx(hykkt_solver s, const double* vals, int n);
int hykkt_set_rs(hykkt_solver s, const double* vals, int n);
int hykkt_set_ry(hykkt_solver s, const double* vals, int n);
int hykkt_set_ryd(hykkt_solver s, const double* vals, int n);

/* set_x_host / set_s_host / set_y_host / set_yd_host, HykktSolver.cpp:214-232.
 * Initial guesses; only y is used (the Schur CG starting point). */
int hykkt_set_x_host(hykkt_solver s, const double* vals, int n);
int hykkt_set_s_host(hykkt_solver s, const double* vals, int n);
int hykkt_set_y_host(hykkt_solver s, const double* vals, int n);
int hykkt_set_yd_host(hykkt_solver s, const double* vals, int n);

/* HykktSolver::set_gamma, HykktSolver.cpp:235-238. */
int hykkt_set_gamma(hykkt_solver s, double gamma);

/* HykktSolver::read_matrix_files, HykktSolver.cpp:113-161 (input_functions.cpp
 * semantics: the last of skip_lines skipped lines is the size line; vectors
 * skip 2 lines).  The solver owns the data read. */
int hykkt_read_matrix_files(hykkt_solver s, const char* h_file, const char* ds_file,
                            const char* jc_file, const char* jd_file, const char* rx_file,
                            const char* rs_file, const char* ry_file, const char* ryd_file,
                            int skip_lines);

/* HykktSolver::execute, HykktSolver.cpp:240-286.  0 if the relative residual
 * (rows 1 and 3, HykktSolver.cpp:846-930) is below norm_tol, 1 if not or if
 * the pattern changed, negative on error.  The first call (and the first after
 * a pattern change on the file path) runs the symbolic analysis; later calls
 * reuse it and refactor numerically. */
int hykkt_execute(hykkt_solver s);

/* ---------------------------------------------------------------------------
 * Additions (no counterpart in the reference API).
 * ------------------------------------------------------------------------- */
typedef struct {
  int    ruiz_its;          /* default 2     (HykktSolver.hpp:367) */
  double norm_tol;          /* default 1e-2  (HykktSolver.hpp:368) */
  double pivot_tol;         /* default 1e-12 (HykktSolver.hpp:369) */
  double cg_tol;            /* default 1e-12 (SchurComplementConjugateGradient.hpp:126) */
  int    cg_itmax;          /* default 100   (SchurComplementConjugateGradient.hpp:125) */
  double delta;             /* diagonal shift added to H_gamma (add_diag hook,
                               HykktSolver.cpp:392); default 0 */
  int    inputs_on_device;  /* 1: pointers given to the setters are device pointers */
  int    cg_check_every;    /* the stopping test runs on the device every iteration; the host
                               reads whether systems are still iterating every k iterations
                               (k > 0), or first after as many iterations as the last solve
                               took and then every iteration (0, the default) */
  int    verbose;           /* print the reference's progress lines (default 0) */
  int    refine_max;        /* iterative-refinement passes on the 4x4 KKT system after the
                               reference's single pass (default 3; 0 = the reference) */
  double refine_tol;        /* refine while resid4 > refine_tol (default 9e-9); a pass is
                               kept only if it lowers resid4 */
  int    autotune;          /* 1 (default): at the first solve of a pattern, time the numeric
                               factorization under several launch shapes (hykkt_tuning) and
                               keep the fastest; results are bit-identical under all of them.
                               The timing runs are left out of the stats' times and of the
                               kernel profile.  HYKKT_PAIR_MIN / HYKKT_GROUP_MAX in the
                               environment fix the shapes instead */
  int    sweep_share;       /* 1 (default): the triangular-solve sweeps may hold every
                               workgroup slot of the device; k > 1: each sweep launch holds at
                               most 1/k of them, leaving compute units to other solvers'
                               kernels (many solvers sharing the device); same results */
} hykkt_options;

typedef struct {
  double resid2;            /* relative residual, block rows 1 and 3 (reference form) */
  double resid4;            /* relative residual over all four block rows */
  double cg_final_norm;     /* ||r|| at CG exit */
  int    cg_iters;
  int    cg_status;         /* 0 converged, 1 itmax reached */
  int    pivot;             /* -1, or first non-positive pivot (internal order) */
  int    n, m, d;
  int    nnz_htil, nnz_hgam;
  long long nnz_l;          /* stored factor entries incl. amalgamation zeros */
  int    n_supernodes, n_levels, max_front;
  double factor_flops;      /* sum of squared column counts */
  double t_symbolic_ms;     /* last symbolic analysis (host) */
  double t_solve_ms;        /* last execute, device work, event-timed */
  double t_factor_ms;       /* numeric factorization part of t_solve_ms */
  double t_cg_ms;           /* Schur CG part of t_solve_ms */
  int    n_solves;
  double resid4_first;      /* resid4 after the first pass, before refinement */
  int    refine_steps;      /* refinement passes kept in the last execute */
  double t_refine_ms;       /* refinement part of t_solve_ms */
} hykkt_stats;

/* Launch shapes of the numeric factorization the autotuner chose for the
 * current pattern (no counterpart in the reference). */
typedef struct {
  int    short_max;         /* TRSM tiles (of 128 rows, over the systems of a launch) up to which
                               a step runs its TRSM and next-column update in 32-row tiles */
  int    group_max;         /* most consecutive steps sharing one trailing update */
  int    pair_min;          /* trailing-update tiles a group of G steps needs: pair_min*G*G/4 */
  int    count;             /* systems per launch the shapes were tuned for */
  int    candidates;        /* shapes timed (0: not tuned, defaults kept) */
  double factor_ms;         /* factorization time with the chosen shapes */
  double default_ms;        /* factorization time with the default shapes */
} hykkt_tuning;

int hykkt_default_options(hykkt_options* o);
int hykkt_get_tuning(hykkt_solver s, hykkt_tuning* t);
int hykkt_set_options(hykkt_solver s, const hykkt_options* o);
int hykkt_get_options(hykkt_solver s, hykkt_options* o);
int hykkt_get_stats(hykkt_solver s, hykkt_stats* st);

/* Copy the last solution to host arrays (any may be NULL).  The reference
 * leaves it on the device (the copy-back is disabled, HykktSolver.cpp:451-456). */
int hykkt_get_solution(hykkt_solver s, double* dx, double* ds, double* dyc, double* dyd);
/* Device pointers of the last solution (valid until the next execute/destroy). */
int hykkt_get_solution_device(hykkt_solver s, const double** dx, const double** ds,
                              const double** dyc, const double** dyd);

/* ---------------------------------------------------------------------------
 * Batched form (no counterpart in the reference, whose execute() solves one
 * system; its numeric refactorisation per system is HykktSolver.cpp:240-286,
 * CholeskyClass.cpp:71-101).  A batch solver holds up to `capacity` systems
 * that share one sparsity pattern: the pattern is given by the hykkt_set_H /
 * _Ds / _Jc / _Jd setters (their value arrays are not used), each system's
 * values by hykkt_bind_system, and hykkt_execute_batch solves systems
 * 0 .. count-1 together, every kernel launch covering all of them.  Each
 * system gets exactly the arithmetic it gets alone through hykkt_execute, so
 * its results are bit-identical to a solve of that system alone.  Pointers
 * follow the option inputs_on_device (device or host arrays) and are read at
 * each hykkt_execute_batch.
 * ------------------------------------------------------------------------- */
int hykkt_create_batch(double gamma, int capacity, hykkt_solver* out);
int hykkt_batch_capacity(hykkt_solver s, int* capacity);
/* Values of system k < capacity: H, Ds, Jc, Jd values in the setters'
 * patterns, the four right-hand sides, and the Schur CG start y0 (may be
 * NULL: zero); jd_vals may be NULL when Jd is empty. */
int hykkt_bind_system(hykkt_solver s, int k, double gamma, const double* h_vals,
                      const double* ds_vals, const double* jc_vals, const double* jd_vals,
                      const double* rx, const double* rs, const double* ry, const double* ryd,
                      const double* y0);
/* status[k] (count entries): what hykkt_execute would return for system k
 * alone (0, 1, or a negative error code of that system, e.g. HYKKT_E_ARG for
 * a row without a nonzero value). */
int hykkt_execute_batch(hykkt_solver s, int count, int* status);
int hykkt_get_stats_k(hykkt_solver s, int k, hykkt_stats* st);
int hykkt_get_solution_k(hykkt_solver s, int k, double* dx, double* ds, double* dyc, double* dyd);
int hykkt_get_solution_device_k(hykkt_solver s, int k, const double** dx, const double** ds,
                                const double** dyc, const double** dyd);

/* The fill-reducing ordering used (perm[k] = original index of pivot k). */
int hykkt_get_permutation(hykkt_solver s, int* perm, int n);
/* Replace the ordering before the first execute (PermClass::add_perm,
 * PermClass.cpp:63-68); perm must be a permutation of 0..n-1. */
int hykkt_set_permutation(hykkt_solver s, const int* perm, int n);

/* Internal arrays after the last execute, copied to the host, for parity
 * checks against the oracle.  `what` selects (sizes in elements):
 *   1 H~ values after Ruiz (nnz_htil)    2 H~ pattern rowptr (n+1)   3 colidx (nnz_htil)
 *   4 max_d (n+m)                        5 Jc values after Ruiz (nnz Jc)
 *   6 r~x after Ruiz (n)                 7 ry after Ruiz (m)
 *   8 H_gamma pattern rowptr (n+1)       9 H_gamma colidx (nnz_hgam)
 *  10 H_gamma as the factorization receives it: the last execute's assembly
 *     (Ruiz-scaled H~ + gamma Jc^T Jc, permuted, plus delta on the diagonal)
 *     is run again into the factor storage and read back from there, in
 *     P*Hg*P^T CSR order (selectors 15, 16), NaN at upper-triangle positions
 *     (nnz_hgam).  The factor is overwritten; the solution stays valid.
 *  11 H~ values before Ruiz (nnz_htil)  12 r^x permuted = rxp (n)
 *  13 Schur right-hand side (m)        14 Schur CG solution before unscaling (m)
 *  15 P*Hg*P^T rowptr (n+1)  16 its colidx (nnz_hgam)  17 its value map (nnz_hgam)
 *  18 Jc*P^T colidx (nnz Jc)  19 its map (nnz Jc)
 *  20 P*Jc^T rowptr (n+1)    21 its colidx (nnz Jc)   22 its map (nnz Jc)
 * Returns the number of elements copied or a negative error. */
long long hykkt_debug_copy(hykkt_solver s, int what, void* host_out, long long capacity);

/* MatrixMarket ingest with the reference's semantics: read_mm_file_into_coo
 * followed by sym_coo_to_csr (symmetric = 1: a lower-triangle file expanded
 * to both triangles) or coo_to_csr (input_functions.cpp:16-200); the last of
 * the skip_lines skipped lines is the "rows cols nnz" line.  Call once with
 * NULL arrays to learn n, m and nnz, then with arrays of n+1, nnz, nnz. */
int hykkt_mm_read_csr(const char* path, int skip_lines, int symmetric, int* n, int* m, int* nnz,
                      int* rowptr, int* colidx, double* vals);
/* read_rhs (input_functions.cpp:202-217): two header lines, then one value
 * per line.  NULL vals: only *n is set. */
int hykkt_mm_read_rhs(const char* path, int* n, double* vals);

/* Device memory helpers for callers without a HIP runtime of their own (for
 * example a ctypes binding), so that inputs can be made resident in HBM. */
int hykkt_device_malloc(void** ptr, unsigned long long bytes);
int hykkt_device_free(void* ptr);
int hykkt_memcpy_htod(void* dst, const void* src, unsigned long long bytes);
int hykkt_memcpy_dtoh(void* dst, const void* src, unsigned long long bytes);
int hykkt_device_synchronize(void);
/* dst[i] = src[i], i < count, by a plain streaming copy kernel (null stream):
 * the known-size reference against which HBM byte counters are scaled. */
int hykkt_device_copy_f64(double* dst, const double* src, long long count);
/* Free and total device memory in bytes (hipMemGetInfo); the served form's
 * HBM accounting reads it around each slot it creates.  No counterpart in the
 * reference. */
int hykkt_device_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes);

/* Per-kernel timing of the factorization and triangular-solve launches
 * (no counterpart in the reference, which times whole phases with
 * std::chrono in its examples).  While enabled, every launch of the kernels
 * below is bracketed by HIP events on the solver's stream and its duration
 * is summed with the algorithmic flops and HBM bytes it performs:
 *   k_dense_nt<128> / k_dense_nt<32>  MFMA tiles of 128 or 32 rows x 128
 *                 columns: TRSM (r*b^2), panel update and U = -L21 L21^T
 *                 (2*K per lower-triangle entry); the 32-row tiles for the
 *                 short launches on the factorization's critical path
 *   k_potrf_inv   128-column diagonal blocks (b^3/3)
 *   k_extend_add  update-matrix assembly (1 flop, 24 bytes per entry)
 *   k_solve_fwd / k_solve_bwd  one sweep each (2*nnz(L) flops, 8*nnz(L) bytes)
 *   k_spmv        CSR products with Jc, Jc^T, Jd^T (2*nnz flops; indices, values,
 *                 x and y each moved once)
 *   k_spgemm_values / k_spgemm_scatter  H~ and H_gamma values (2 flops and 24
 *                 bytes per term, plus each output entry's indices, D term and
 *                 store)
 * hykkt_profile_enable clears the totals; hykkt_profile_get writes up to
 * `capacity` entries, one per kernel launched since, and sets *count to the
 * number of those kernels. */
typedef struct {
  char name[32];
  long long launches;
  double ms;      /* sum of launch durations */
  double flops;   /* algorithmic flops of those launches */
  double bytes;   /* algorithmic HBM bytes of those launches */
} hykkt_kernel_time;
int hykkt_profile_enable(hykkt_solver s, int on);
int hykkt_profile_get(hykkt_solver s, hykkt_kernel_time* out, int capacity, int* count);

/* ---------------------------------------------------------------------------
 * Component entry points mirroring the reference's step classes, used by
 * their known-answer tests.  Host arrays in and out.
 * ------------------------------------------------------------------------- */

/* RuizClass (RuizClass.cpp:54-98): in-place scaling of H (n x n), Jt (n x m2)
 * and J (m2 x n) plus rhs1 (n), rhs2 (m2); max_d (n+m2) receives the
 * accumulated scaling.  Runs on the GPU. */
int hykkt_ruiz_scale(int ruiz_its, int n, int m2,
                     const int* h_rp, const int* h_ci, double* h_v, int h_nnz,
                     const int* jt_rp, const int* jt_ci, double* jt_v,
                     const int* j_rp, const int* j_ci, double* j_v, int j_nnz,
                     double* rhs1, double* rhs2, double* max_d);

/* hykkt_ruiz_scale without any device allocation (no counterpart in the
 * reference: for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same arguments, semantics and results as hykkt_ruiz_scale; the device
 * copies of the arrays and the scaling vectors live in a fixed 32 MiB device
 * scratch that each GPU holds until the process ends.  Systems that need more
 * (the index and value arrays, rhs, scale and max_d together, each rounded up
 * to 256 bytes) are refused with HYKKT_E_ARG.  Calls that overlap on one GPU
 * share that scratch and race. */
int hykkt_ruiz_scale_noalloc(int ruiz_its, int n, int m2,
                             const int* h_rp, const int* h_ci, double* h_v, int h_nnz,
                             const int* jt_rp, const int* jt_ci, double* jt_v,
                             const int* j_rp, const int* j_ci, double* j_v, int j_nnz,
                             double* rhs1, double* rhs2, double* max_d);

/* Host permutation maps (permcheck.cu:99-195). */
int hykkt_perm_reverse(int n, const int* perm, int* rev);
int hykkt_perm_map_rc(int n, const int* rows, const int* cols, const int* perm, const int* rev,
                      int* prows, int* pcols, int* pmap);
int hykkt_perm_map_c(int n, const int* rows, const int* cols, const int* rev, int* pcols,
                     int* pmap);
int hykkt_perm_map_r(int n, const int* rows, const int* cols, const int* perm, int* prows,
                     int* pcols, int* pmap);
/* Fill-reducing ordering of a symmetric pattern (PermClass::symamd,
 * PermClass.cpp:71-99 -- an independent AMD, see amd.cpp). */
int hykkt_amd_order(int n, const int* rp, const int* ci, int* perm);

/* Host-only symbolic statistics of the supernodal Cholesky of the lower
 * triangle of an n x n CSR pattern (no device needed): nnz(L) without
 * amalgamation zeros, flops = sum of squared column counts, supernode and
 * level counts, largest front. */
int hykkt_chol_analyze_host(int n, const int* rp, const int* ci, long long* nnz_l,
                            double* flops, int* n_supernodes, int* n_levels, int* max_front);

/* CholeskyClass (CholeskyClass.cpp:9-101): sparse Cholesky of the LOWER
 * triangle of an n x n CSR matrix (entries with col <= row). */
typedef struct hykkt_chol_s* hykkt_chol;
int hykkt_chol_create(int n, int nnz, const int* rp, const int* ci, hykkt_chol* out);
/* Numeric factorization; *singularity = -1 or the first failing pivot
 * (cusolverSpDcsrcholZeroPivot semantics).  Returns HYKKT_E_PIVOT on failure. */
int hykkt_chol_factor(hykkt_chol c, const double* vals, double tol, int* singularity);
int hykkt_chol_solve(hykkt_chol c, const double* b, double* x);
int hykkt_chol_info(hykkt_chol c, long long* nnz_l, int* n_supernodes, int* n_levels,
                    double* flops);
int hykkt_chol_destroy(hykkt_chol c);

/* SchurComplementConjugateGradient::solve (…ConjugateGradient.cpp:89-194):
 * CG on S = J * A^{-1} * J^T with A factored in `c` (J is m x n host CSR).
 * x: initial guess in, solution out.  Returns 0 converged, 1 itmax. */
int hykkt_schur_cg(hykkt_chol c, int m, int n, const int* j_rp, const int* j_ci,
                   const double* j_v, double* x, const double* b, double tol, int itmax,
                   int* iters, double* final_norm);

#ifdef __cplusplus
}
#endif
#endif
