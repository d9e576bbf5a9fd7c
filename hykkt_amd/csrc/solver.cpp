// solver.cpp -- HykktSolver::execute() on one MI355X.
//
// Step order follows HykktSolver.cpp:240-457 (+ check_error :846-930).
// Everything per solve runs on one HIP stream; the host only launches and
// reads back the CG convergence scalar (as the reference does) and the
// final residual.
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace hykkt {

struct Solver::Sym {
  int n = 0, m = 0, d = 0, nnz_h = 0, nnz_jc = 0, nnz_jd = 0, nnz_ds = 0;
  bool has_jd = false;
  Pattern H, Jc, Jd;
  TransposeMap jct, jdt;
  SpgemmPlan htp, hgp;
  std::vector<int> perm, rev;
  Pattern hgp_pat, jcp_pat, jctp_pat;
  std::vector<int> hgp_map, jcp_map, jctp_map;
  std::unique_ptr<CholEngine> chol;
  std::vector<int> phat, ipos;
  DevBuf<int> h_rp, h_ci, jc_rp, jc_ci, jd_rp, jd_ci;
  DevBuf<int> jct_rp, jct_ci, jct_map, jdt_rp, jdt_ci, jdt_map;
  DevBuf<int> ht_rp, ht_ci, ht_dsrc, ht_tp, ht_ta, ht_tb;
  DevBuf<int> hg_dsrc, hg_tp, hg_ta, hg_tb, ent_src;
  DevBuf<int> jch_rp, jch_ci, jch_map, jcth_rp, jcth_ci, jcth_map;
  DevBuf<int> phat_d, ipos_d;
  DevBuf<double> h_v, jc_v, jd_v, ds_v, rx, rs, ry, ryd, y0;
  DevBuf<double> jd_vs, ryd_s, rx_til, htil, jcs, ry_s, max_d, scale;
  DevBuf<double> jch_v, jcth_v, bhat, rxp, tmp_n, t1;
  DevBuf<double> cx, cr, cw, cp, cs, schur, sc, part;
  DevBuf<double> dx, dsv, dyc, dyd, sq_n, sq_m, sq_d2, sq_d4, res;
  // refinement: signed residual rows, scaled rhs, correction / candidate
  DevBuf<double> e_x, e_s, e_y, e_d, e_rxt, e_ysc, e_rxp, e_schur, e_cx;
  DevBuf<double> c_dx, c_dsv, c_dyc, c_dyd;
  DevBuf<int> zero_row;   // smallest all-zero row seen by the Ruiz sweeps (INT_MAX: none)
  // current value pointers (device)
  const double *hv = nullptr, *jcv = nullptr, *jdv = nullptr, *dsd = nullptr;
  const double *drx = nullptr, *drs = nullptr, *dry = nullptr, *dryd = nullptr;
};

namespace {

template <class T>
void upload_or_dummy(DevBuf<T>& b, const std::vector<T>& v)
{
  if (v.empty()) b.upload(std::vector<T>{T{}});
  else b.upload(v);
}

}  // namespace

Solver::Solver(double gamma) : gamma_(gamma)
{
  hykkt_default_options(&opts);
  HIPCHK(hipStreamCreate(&st_));
  for (auto& e : ev_) HIPCHK(hipEventCreate(&e));
}

Solver::~Solver()
{
  sym_.reset();
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (st_) (void)hipStreamDestroy(st_);
}

void Solver::set_block(int which, const int* rp, const int* ci, const double* v, int n, int m,
                       int nnz)
{
  HYKKT_REQUIRE(which >= 0 && which < 4, HYKKT_E_ARG, "bad block id");
  HYKKT_REQUIRE(n >= 0 && m >= 0 && nnz >= 0, HYKKT_E_ARG, "negative dimension");
  HYKKT_REQUIRE(rp || n == 0, HYKKT_E_ARG, "null row pointer");
  HYKKT_REQUIRE((ci && v) || nnz == 0, HYKKT_E_ARG, "null column/value pointer");
  if (which == BLK_H || which == BLK_DS)
    HYKKT_REQUIRE(n == m, HYKKT_E_ARG, "H and Ds must be square");   // HykktSolver.cpp:165,172
  blk_[which] = BlockIn{rp, ci, v, n, m, nnz, true};
  file_path_ = false;
  if (which == BLK_JD) {
    // HykktSolver.cpp:188-191: the API path cannot detect pattern changes.
    jd_flag_ = nnz > 0;
    status_ = true;
  }
}

void Solver::set_vec(int which, const double* v, int n)
{
  HYKKT_REQUIRE(which >= 0 && which < 8, HYKKT_E_ARG, "bad vector id");
  HYKKT_REQUIRE(v || n == 0, HYKKT_E_ARG, "null vector");
  vecs_[which] = v;
  vec_n_[which] = n;
}

void Solver::set_perm(const int* perm, int n)
{
  std::vector<int> p(perm, perm + n), seen(n, 0);
  for (int v : p) {
    HYKKT_REQUIRE(v >= 0 && v < n && !seen[v], HYKKT_E_ARG, "not a permutation");
    seen[v] = 1;
  }
  user_perm_ = std::move(p);
  sym_.reset();   // re-analyse with the new ordering
}

void Solver::read_files(const char* const files[8], int skip)
{
  // HykktSolver::read_matrix_files, HykktSolver.cpp:113-161
  fblk_[BLK_H] = read_mm_csr(files[0], skip, true);
  fblk_[BLK_DS] = read_mm_csr(files[1], skip, false);
  fblk_[BLK_JC] = read_mm_csr(files[2], skip, false);
  fblk_[BLK_JD] = read_mm_csr(files[3], skip, false);
  for (int k = 0; k < 4; k++) fvec_[k] = read_rhs(files[4 + k]);
  HYKKT_REQUIRE(static_cast<int>(fvec_[0].size()) == fblk_[BLK_H].n &&
                    static_cast<int>(fvec_[1].size()) == fblk_[BLK_DS].n &&
                    static_cast<int>(fvec_[2].size()) == fblk_[BLK_JC].n &&
                    static_cast<int>(fvec_[3].size()) == fblk_[BLK_JD].n,
                HYKKT_E_IO, "rhs length does not match its block");
  bool jd_flag = fblk_[BLK_JD].nnz() > 0;
  status_ = (jd_flag_ == jd_flag) || !sym_;
  jd_flag_ = jd_flag;
  for (int k = 0; k < 4; k++) {
    CsrHost& b = fblk_[k];
    blk_[k] = BlockIn{b.rp.data(), b.ci.empty() ? nullptr : b.ci.data(),
                      b.v.empty() ? nullptr : b.v.data(), b.n, b.m, b.nnz(), true};
  }
  for (int k = 0; k < 4; k++) {
    vecs_[k] = fvec_[k].data();
    vec_n_[k] = static_cast<int>(fvec_[k].size());
  }
  if (!vecs_[VEC_Y] || vec_n_[VEC_Y] != fblk_[BLK_JC].n) {
    fvec_[VEC_Y].assign(fblk_[BLK_JC].n, 0.0);   // h_y_ zeros (HykktSolver.cpp:153)
    vecs_[VEC_Y] = fvec_[VEC_Y].data();
    vec_n_[VEC_Y] = fblk_[BLK_JC].n;
  }
  file_path_ = true;
}

bool Solver::pattern_matches() const
{
  const Sym& S = *sym_;
  return blk_[BLK_H].n == S.n && blk_[BLK_H].nnz == S.nnz_h && blk_[BLK_JC].n == S.m &&
         blk_[BLK_JC].nnz == S.nnz_jc && blk_[BLK_JD].nnz == S.nnz_jd &&
         blk_[BLK_DS].n == S.d && blk_[BLK_DS].nnz == S.nnz_ds;
}

// ------------------------------------------------------------------ symbolic
void Solver::analyze()
{
  auto t0 = std::chrono::steady_clock::now();
  sym_.reset();
  auto S = std::make_unique<Sym>();
  const BlockIn &H = blk_[BLK_H], &Ds = blk_[BLK_DS], &Jc = blk_[BLK_JC], &Jd = blk_[BLK_JD];
  const int n = H.n, m = Jc.n, d = Ds.n;
  HYKKT_REQUIRE(Jc.m == n, HYKKT_E_ARG, "Jc must have n columns");
  HYKKT_REQUIRE(Jd.n == d && (Jd.m == n || Jd.nnz == 0), HYKKT_E_ARG,
                "Jd must be d x n with d = size of Ds");
  HYKKT_REQUIRE(Ds.nnz == d, HYKKT_E_ARG, "Ds must hold one diagonal entry per row");
  S->n = n;
  S->m = m;
  S->d = d;
  S->nnz_h = H.nnz;
  S->nnz_jc = Jc.nnz;
  S->nnz_jd = Jd.nnz;
  S->nnz_ds = Ds.nnz;
  S->has_jd = Jd.nnz > 0;
  const bool on_dev = opts.inputs_on_device != 0;
  auto fetch = [&](const BlockIn& b, Pattern& p, int rows, int cols) {
    p.n = rows;
    p.m = cols;
    p.rp.resize(static_cast<size_t>(rows) + 1);
    p.ci.resize(b.nnz);
    if (on_dev) {
      HIPCHK(hipMemcpy(p.rp.data(), b.rp, sizeof(int) * (rows + 1), hipMemcpyDeviceToHost));
      if (b.nnz) HIPCHK(hipMemcpy(p.ci.data(), b.ci, sizeof(int) * b.nnz, hipMemcpyDeviceToHost));
    } else {
      std::memcpy(p.rp.data(), b.rp, sizeof(int) * (rows + 1));
      if (b.nnz) std::memcpy(p.ci.data(), b.ci, sizeof(int) * b.nnz);
    }
    HYKKT_REQUIRE(p.rp[0] == 0 && p.rp[rows] == b.nnz, HYKKT_E_ARG, "bad CSR row pointer");
    for (int i = 0; i < rows; i++) {
      HYKKT_REQUIRE(p.rp[i + 1] >= p.rp[i], HYKKT_E_ARG, "CSR row pointer not monotone");
      for (int q = p.rp[i]; q < p.rp[i + 1]; q++) {
        HYKKT_REQUIRE(p.ci[q] >= 0 && p.ci[q] < cols, HYKKT_E_ARG, "column index out of range");
        HYKKT_REQUIRE(q == p.rp[i] || p.ci[q] > p.ci[q - 1], HYKKT_E_ARG,
                      "CSR columns must be sorted and unique");
      }
    }
  };
  fetch(H, S->H, n, n);
  fetch(Jc, S->Jc, m, n);
  Pattern dsp;
  fetch(Ds, dsp, d, d);
  for (int i = 0; i < d; i++)
    HYKKT_REQUIRE(dsp.rp[i + 1] - dsp.rp[i] == 1 && dsp.ci[dsp.rp[i]] == i, HYKKT_E_ARG,
                  "Ds must be diagonal with one entry per row");
  if (S->has_jd) fetch(Jd, S->Jd, d, n);
  else {
    S->Jd.n = d;
    S->Jd.m = n;
    S->Jd.rp.assign(static_cast<size_t>(d) + 1, 0);
  }

  // transposes (csr2csc every solve in the reference, HykktSolver.cpp:292-303,670-681)
  S->jct = transpose_map(S->Jc);
  S->jdt = transpose_map(S->Jd);
  // H~ = H + Jd^T (Ds Jd)  (HykktSolver.cpp:289-337, 685-697)
  if (S->has_jd) {
    S->htp = spgemm_plan(S->jdt.at, S->jdt.map, S->Jd, S->H);
  } else {
    S->htp.e = S->H;
    S->htp.d_src.resize(S->H.nnz());
    std::iota(S->htp.d_src.begin(), S->htp.d_src.end(), 0);
    S->htp.term_ptr.assign(static_cast<size_t>(S->H.nnz()) + 1, 0);
  }
  // H_gamma = H~ + gamma Jc^T Jc  (HykktSolver.cpp:352-383)
  S->hgp = spgemm_plan(S->jct.at, S->jct.map, S->Jc, S->htp.e);
  const Pattern& HG = S->hgp.e;
  // ordering + reference maps (HykktSolver.cpp:760-798)
  if (!user_perm_.empty()) {
    HYKKT_REQUIRE(static_cast<int>(user_perm_.size()) == n, HYKKT_E_ARG,
                  "permutation length differs from n");
    S->perm = user_perm_;
  } else {
    S->perm = amd_order(HG);
  }
  perm_reverse(S->perm, S->rev);
  perm_map_rc(HG, S->perm, S->rev, S->hgp_pat, S->hgp_map);
  perm_map_c(S->Jc, S->rev, S->jcp_pat, S->jcp_map);
  perm_map_r(S->jct.at, S->perm, S->jctp_pat, S->jctp_map);
  // Cholesky analysis of the lower triangle of P Hg P^T (CholeskyClass.cpp:31-58)
  S->chol = std::make_unique<CholEngine>(S->hgp_pat, CholOptions{}, st_);
  S->chol->set_profiler(&prof);
  const CholPlan& pl = S->chol->plan();
  S->phat.resize(n);
  S->ipos.resize(n);
  for (int k = 0; k < n; k++) {
    S->phat[k] = S->perm[pl.post[k]];
    S->ipos[S->phat[k]] = k;
  }
  std::vector<int> ent(pl.a_src.size());
  for (size_t e = 0; e < ent.size(); e++) ent[e] = S->hgp_map[pl.a_src[e]];
  // Jc with internal column order, and its transpose; maps into Jc positions
  Pattern jch;
  std::vector<int> jch_map;
  perm_map_c(S->Jc, S->ipos, jch, jch_map);
  TransposeMap jcht = transpose_map(jch);
  std::vector<int> jcth_map(jcht.map.size());
  for (size_t q = 0; q < jcth_map.size(); q++) jcth_map[q] = jch_map[jcht.map[q]];

  // ---- device index arrays ----
  S->h_rp.upload(S->H.rp);
  upload_or_dummy(S->h_ci, S->H.ci);
  S->jc_rp.upload(S->Jc.rp);
  upload_or_dummy(S->jc_ci, S->Jc.ci);
  S->jd_rp.upload(S->Jd.rp);
  upload_or_dummy(S->jd_ci, S->Jd.ci);
  S->jct_rp.upload(S->jct.at.rp);
  upload_or_dummy(S->jct_ci, S->jct.at.ci);
  upload_or_dummy(S->jct_map, S->jct.map);
  S->jdt_rp.upload(S->jdt.at.rp);
  upload_or_dummy(S->jdt_ci, S->jdt.at.ci);
  upload_or_dummy(S->jdt_map, S->jdt.map);
  S->ht_rp.upload(S->htp.e.rp);
  upload_or_dummy(S->ht_ci, S->htp.e.ci);
  upload_or_dummy(S->ht_dsrc, S->htp.d_src);
  S->ht_tp.upload(S->htp.term_ptr);
  upload_or_dummy(S->ht_ta, S->htp.ta);
  upload_or_dummy(S->ht_tb, S->htp.tb);
  upload_or_dummy(S->hg_dsrc, S->hgp.d_src);
  S->hg_tp.upload(S->hgp.term_ptr);
  upload_or_dummy(S->hg_ta, S->hgp.ta);
  upload_or_dummy(S->hg_tb, S->hgp.tb);
  upload_or_dummy(S->ent_src, ent);
  S->jch_rp.upload(jch.rp);
  upload_or_dummy(S->jch_ci, jch.ci);
  upload_or_dummy(S->jch_map, jch_map);
  S->jcth_rp.upload(jcht.at.rp);
  upload_or_dummy(S->jcth_ci, jcht.at.ci);
  upload_or_dummy(S->jcth_map, jcth_map);
  S->phat_d.upload(S->phat);
  S->ipos_d.upload(S->ipos);

  // ---- device work buffers ----
  auto A = [](DevBuf<double>& b, size_t k) { b.alloc(std::max<size_t>(k, 1)); };
  A(S->jd_vs, S->nnz_jd);
  A(S->ryd_s, d);
  A(S->rx_til, n);
  A(S->htil, S->htp.e.nnz());
  A(S->jcs, S->nnz_jc);
  A(S->ry_s, m);
  A(S->max_d, n + m);
  A(S->scale, n + m);
  A(S->jch_v, S->nnz_jc);
  A(S->jcth_v, S->nnz_jc);
  A(S->bhat, n);
  A(S->rxp, n);
  A(S->tmp_n, n);
  A(S->t1, n);
  for (DevBuf<double>* b : {&S->cx, &S->cr, &S->cw, &S->cp, &S->cs, &S->schur}) A(*b, m);
  A(S->sc, 8);
  A(S->part, dev::reduce_partials());
  A(S->dx, n);
  A(S->dsv, d);
  A(S->dyc, m);
  A(S->dyd, d);
  A(S->sq_n, n);
  A(S->sq_m, m);
  A(S->sq_d2, d);
  A(S->sq_d4, d);
  A(S->res, 16);
  for (DevBuf<double>* b : {&S->e_x, &S->e_rxt, &S->e_rxp, &S->c_dx}) A(*b, n);
  for (DevBuf<double>* b : {&S->e_y, &S->e_ysc, &S->e_schur, &S->e_cx, &S->c_dyc}) A(*b, m);
  for (DevBuf<double>* b : {&S->e_s, &S->e_d, &S->c_dsv, &S->c_dyd}) A(*b, d);
  A(S->y0, m);
  S->zero_row.alloc(1);
  HIPCHK(hipDeviceSynchronize());

  stats.n = n;
  stats.m = m;
  stats.d = d;
  stats.nnz_htil = S->htp.e.nnz();
  stats.nnz_hgam = HG.nnz();
  stats.nnz_l = pl.nnz_l_stored;
  stats.n_supernodes = pl.ns;
  stats.n_levels = static_cast<int>(pl.level_ptr.size()) - 1;
  stats.max_front = pl.max_front;
  stats.factor_flops = pl.flops;
  stats.t_symbolic_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  sym_ = std::move(S);
}

// ------------------------------------------------------------------ values
const double* Solver::vec(int which) const { return vecs_[which]; }

void Solver::upload_values()
{
  Sym& S = *sym_;
  auto need = [&](int which, int len, const char* name) {
    HYKKT_REQUIRE(vecs_[which] != nullptr, HYKKT_E_STATE, std::string(name) + " not set");
    HYKKT_REQUIRE(vec_n_[which] == 0 || vec_n_[which] >= len, HYKKT_E_ARG,
                  std::string(name) + " too short");
  };
  need(VEC_RX, S.n, "rx");
  need(VEC_RS, S.d, "rs");
  need(VEC_RY, S.m, "ry");
  need(VEC_RYD, S.d, "ryd");
  if (opts.inputs_on_device) {
    S.hv = blk_[BLK_H].v;
    S.jcv = blk_[BLK_JC].v;
    S.jdv = blk_[BLK_JD].v;
    S.dsd = blk_[BLK_DS].v;
    S.drx = vecs_[VEC_RX];
    S.drs = vecs_[VEC_RS];
    S.dry = vecs_[VEC_RY];
    S.dryd = vecs_[VEC_RYD];
    if (vecs_[VEC_Y]) dev::copy(S.y0.p, vecs_[VEC_Y], S.m, st_);
    else dev::fill(S.y0.p, 0.0, S.m, st_);
  } else {
    // value-only H2D per solve (HykktSolver.cpp:635-666), into staging
    // buffers made on first use (the option inputs_on_device may change
    // between solves without a new analysis)
    auto h2d = [&](DevBuf<double>& b, const double* src, size_t cnt) {
      b.ensure(std::max<size_t>(cnt, 1));
      if (cnt) HIPCHK(hipMemcpyAsync(b.p, src, cnt * sizeof(double), hipMemcpyHostToDevice, st_));
    };
    h2d(S.h_v, blk_[BLK_H].v, S.nnz_h);
    h2d(S.jc_v, blk_[BLK_JC].v, S.nnz_jc);
    h2d(S.jd_v, blk_[BLK_JD].v, S.nnz_jd);
    h2d(S.ds_v, blk_[BLK_DS].v, S.d);
    h2d(S.rx, vecs_[VEC_RX], S.n);
    h2d(S.rs, vecs_[VEC_RS], S.d);
    h2d(S.ry, vecs_[VEC_RY], S.m);
    h2d(S.ryd, vecs_[VEC_RYD], S.d);
    if (vecs_[VEC_Y]) HIPCHK(hipMemcpyAsync(S.y0.p, vecs_[VEC_Y], S.m * sizeof(double),
                                            hipMemcpyHostToDevice, st_));
    else dev::fill(S.y0.p, 0.0, S.m, st_);
    S.hv = S.h_v.p;
    S.jcv = S.jc_v.p;
    S.jdv = S.jd_v.p;
    S.dsd = S.ds_v.p;
    S.drx = S.rx.p;
    S.drs = S.rs.p;
    S.dry = S.ry.p;
    S.dryd = S.ryd.p;
  }
}

void Solver::chol_solve(double* x) { sym_->chol->solve(x); }

// compute_spgemm_hgamma + apply_permutation + add_diag (HykktSolver.cpp:370-395):
// the lower triangle of P*Hg*P^T, each value evaluated from the Ruiz-scaled
// H~ and Jc and written straight to its slot in the factor storage, then the
// diagonal shift.
void Solver::assemble_hgamma()
{
  Sym& S = *sym_;
  CholEngine& ch = *S.chol;
  ch.zero_L();
  dev::spgemm_scatter(ch.n_assembled(), S.ent_src.p, ch.a_dst(), S.hg_dsrc.p, S.htil.p,
                      S.hg_tp.p, S.hg_ta.p, S.hg_tb.p, S.jcs.p, S.jcs.p, gamma_, ch.L(), st_);
  ch.shift_diag(opts.delta);   // add_diag(delta) hook (HykktSolver.cpp:392)
}

// ------------------------------------------------------------------ execute
int Solver::execute()
{
  // HykktSolver.cpp:243-250: a changed Jd emptiness on the file path is an error
  if (file_path_ && !status_ && sym_) {
    if (opts.verbose) std::printf("\n\nERROR: USING HYKKT WITH NEW NONZERO STRUCTURE\n\n");
    return HYKKT_RESIDUAL_FAIL;
  }
  for (int k = 0; k < 4; k++)
    HYKKT_REQUIRE(blk_[k].set, HYKKT_E_STATE, "a matrix block was not set");
  if (!sym_ || !pattern_matches() || sym_->has_jd != (blk_[BLK_JD].nnz > 0)) analyze();
  Sym& S = *sym_;
  const int n = S.n, m = S.m, d = S.d;
  const double gam = gamma_;
  HIPCHK(hipEventRecord(ev_[0], st_));
  upload_values();

  // ---- compute_spgemm_htil (HykktSolver.cpp:289-337) ----
  if (S.has_jd) {
    dev::row_scale(d, S.jd_rp.p, S.jdv, S.dsd, S.dryd, S.drs, S.jd_vs.p, S.ryd_s.p, st_);
    dev::spmv(n, S.jdt_rp.p, S.jdt_ci.p, S.jdt_map.p, S.jdv, 1.0, S.ryd_s.p, 1.0, S.drx,
              S.rx_til.p, st_);
    dev::spgemm_values(S.htp.e.nnz(), S.ht_dsrc.p, S.hv, S.ht_tp.p, S.ht_ta.p, S.ht_tb.p, S.jdv,
                       S.jd_vs.p, 1.0, S.htil.p, st_);
  } else {
    dev::copy(S.rx_til.p, S.drx, n, st_);
    dev::copy(S.htil.p, S.hv, S.nnz_h, st_);
  }
  dev::copy(S.jcs.p, S.jcv, S.nnz_jc, st_);
  dev::copy(S.ry_s.p, S.dry, m, st_);

  // ---- compute_ruiz_scaling (HykktSolver.cpp:340-349; RuizClass.cpp:54-98) ----
  if (opts.ruiz_its <= 0) dev::fill(S.max_d.p, 1.0, n + m, st_);
  dev::fill_i(S.zero_row.p, INT_MAX, 1, st_);
  for (int it = 0; it < opts.ruiz_its; it++) {
    dev::ruiz_rowmax(n, n + m, S.ht_rp.p, S.htil.p, S.jct_rp.p, S.jct_map.p, S.jcs.p, S.jc_rp.p,
                     S.jcs.p, S.scale.p, st_, S.zero_row.p);
    dev::ruiz_scale(n, n + m, S.ht_rp.p, S.ht_ci.p, S.htil.p, S.jct_rp.p, S.jct_ci.p, nullptr,
                    S.jc_rp.p, S.jc_ci.p, S.jcs.p, S.scale.p, S.rx_til.p, S.ry_s.p, S.max_d.p,
                    it == 0, st_);
  }

  // ---- H_gamma straight into the factor storage (HykktSolver.cpp:370-395) ----
  CholEngine& ch = *S.chol;
  assemble_hgamma();

  // ---- compute_hgamma_factorization (HykktSolver.cpp:398-401) ----
  ch.factor(opts.pivot_tol);
  HIPCHK(hipEventRecord(ev_[1], st_));

  // permuted Jc and Jc^T values (apply_permutation, HykktSolver.cpp:386-395)
  dev::gather(S.jch_v.p, S.jcs.p, S.jch_map.p, S.nnz_jc, st_);
  dev::gather(S.jcth_v.p, S.jcs.p, S.jcth_map.p, S.nnz_jc, st_);

  // ---- first pass: the reference's solve (HykktSolver.cpp:404-457) ----
  Pass first{S.rxp.p, S.schur.p, S.cx.p};
  int iters = 0, cg_status = 0;
  double fnorm = 0.0;
  // CG start y_host, or zero when none is bound (then no start-residual solve)
  solve_phase(first, S.rx_til.p, S.ry_s.p, S.drs, S.dryd, vecs_[VEC_Y] ? S.y0.p : nullptr, S.dx.p,
              S.dsv.p, S.dyc.p, S.dyd.p, &iters, &fnorm, &cg_status, ev_[2]);

  // ---- check_error (HykktSolver.cpp:846-930) + rows 2 and 4 ----
  double hr[8];
  int zrow = INT_MAX;
  HIPCHK(hipMemcpyAsync(&zrow, S.zero_row.p, sizeof(int), hipMemcpyDeviceToHost, st_));
  residual(S.dx.p, S.dsv.p, S.dyc.p, S.dyd.p, true, ev_[3], hr);
  // matrix_vector_ops.cu:258-264 scales such a row by 1/sqrt(0) = inf and the
  // solve turns to NaN; here it is an argument error
  if (zrow != INT_MAX)
    throw Error(HYKKT_E_ARG, zrow < n ? "row " + std::to_string(zrow) +
                                            " of [H~ Jc^T] has no nonzero value (Ruiz scale 1/sqrt(0))"
                                      : "row " + std::to_string(zrow - n) +
                                            " of Jc has no nonzero value (Ruiz scale 1/sqrt(0))");
  const int hpiv = ch.pivot();
  HYKKT_REQUIRE(!ch.solve_error(), HYKKT_E_STATE,
                "triangular-solve sweep exceeded its wait bound (results invalid)");
  const double nb = hr[4] + hr[5] + hr[6] + hr[7];
  stats.resid2 = std::sqrt(hr[0] + hr[1]) / std::sqrt(nb);
  stats.resid4 = std::sqrt(hr[0] + hr[1] + hr[2] + hr[3]) / std::sqrt(nb);
  stats.resid4_first = stats.resid4;
  stats.refine_steps = 0;

  // ---- iterative refinement on the 4x4 block KKT system (an addition: the
  // reference stops after one pass).  The signed residual rows go through the
  // right-hand-side half of formation (r~x = ex + Jd^T (Ds ed + es), then the
  // Ruiz scaling max_d) and the same solve phase with a zero CG start; a
  // correction is kept only if it lowers resid4. ----
  Pass corr{S.e_rxp.p, S.e_schur.p, S.e_cx.p};
  for (int k = 0; hpiv < 0 && k < opts.refine_max && stats.resid4 > opts.refine_tol; k++) {
    if (S.has_jd) {
      dev::row_scale(d, S.jd_rp.p, S.jdv, S.dsd, S.e_d.p, S.e_s.p, nullptr, S.ryd_s.p, st_);
      dev::spmv(n, S.jdt_rp.p, S.jdt_ci.p, S.jdt_map.p, S.jdv, 1.0, S.ryd_s.p, 1.0, S.e_x.p,
                S.e_rxt.p, st_);
    } else {
      dev::copy(S.e_rxt.p, S.e_x.p, n, st_);
    }
    dev::vec_scale(S.e_rxt.p, S.max_d.p, n, st_);
    dev::copy(S.e_ysc.p, S.e_y.p, m, st_);
    dev::vec_scale(S.e_ysc.p, S.max_d.p + n, m, st_);
    int it = 0, cgs = 0;
    double fn = 0.0;
    solve_phase(corr, S.e_rxt.p, S.e_ysc.p, S.e_s.p, S.e_d.p, nullptr, S.c_dx.p, S.c_dsv.p,
                S.c_dyc.p, S.c_dyd.p, &it, &fn, &cgs, nullptr);
    dev::add(S.c_dx.p, S.c_dx.p, S.dx.p, n, st_);
    dev::add(S.c_dsv.p, S.c_dsv.p, S.dsv.p, d, st_);
    dev::add(S.c_dyc.p, S.c_dyc.p, S.dyc.p, m, st_);
    dev::add(S.c_dyd.p, S.c_dyd.p, S.dyd.p, d, st_);
    double cr[8];
    residual(S.c_dx.p, S.c_dsv.p, S.c_dyc.p, S.c_dyd.p, true, nullptr, cr);
    HYKKT_REQUIRE(!ch.solve_error(), HYKKT_E_STATE,
                  "triangular-solve sweep exceeded its wait bound (results invalid)");
    const double r4 = std::sqrt(cr[0] + cr[1] + cr[2] + cr[3]) / std::sqrt(nb);
    if (!(r4 < stats.resid4)) break;   // keep the better solution; the rows in e_* are stale
    S.dx.swap(S.c_dx);
    S.dsv.swap(S.c_dsv);
    S.dyc.swap(S.c_dyc);
    S.dyd.swap(S.c_dyd);
    stats.resid2 = std::sqrt(cr[0] + cr[1]) / std::sqrt(nb);
    stats.resid4 = r4;
    stats.refine_steps++;
  }
  HIPCHK(hipEventRecord(ev_[4], st_));
  HIPCHK(hipEventSynchronize(ev_[4]));
  prof.collect();

  stats.cg_iters = iters;
  stats.cg_status = cg_status;
  stats.cg_final_norm = fnorm;
  stats.pivot = hpiv;
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev_[0], ev_[4]));
  stats.t_solve_ms = ms;
  HIPCHK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  stats.t_factor_ms = ms;
  HIPCHK(hipEventElapsedTime(&ms, ev_[1], ev_[2]));
  stats.t_cg_ms = ms;
  HIPCHK(hipEventElapsedTime(&ms, ev_[3], ev_[4]));
  stats.t_refine_ms = ms;
  stats.n_solves++;
  if (opts.verbose) {
    std::printf("||Ax-b||/||b|| = %32.32g\n\n", stats.resid2);
    std::printf(stats.resid2 < opts.norm_tol ? "Residual test passed\n" : "Residual test failed\n");
  }
  if (hpiv >= 0) return HYKKT_RESIDUAL_FAIL;
  return (stats.resid2 < opts.norm_tol) ? HYKKT_OK : HYKKT_RESIDUAL_FAIL;
}

// setup_conjugate_gradient (HykktSolver.cpp:814-843), the Schur CG
// (SchurComplementConjugateGradient.cpp:89-194) and recover_solution
// (HykktSolver.cpp:410-457) for Ruiz-scaled right-hand sides rxt (n, natural
// order) and ysc (m); rs, ryd are the unscaled d-rows used by the recovery.
void Solver::solve_phase(const Pass& w, const double* rxt, const double* ysc, const double* rs,
                         const double* ryd, const double* x0, double* dx, double* dsv,
                         double* dyc, double* dyd, int* iters, double* fnorm, int* cg_status,
                         hipEvent_t after_cg)
{
  Sym& S = *sym_;
  const int n = S.n, m = S.m, d = S.d;
  // r^x = r~x + gamma Jc^T ry in the factor's internal order
  dev::gather(S.tmp_n.p, rxt, S.phat_d.p, n, st_);
  dev::spmv(n, S.jcth_rp.p, S.jcth_ci.p, nullptr, S.jcth_v.p, gamma_, ysc, 1.0, S.tmp_n.p, w.rxp,
            st_);
  dev::copy(S.t1.p, w.rxp, n, st_);
  chol_solve(S.t1.p);
  dev::spmv(m, S.jch_rp.p, S.jch_ci.p, nullptr, S.jch_v.p, 1.0, S.t1.p, -1.0, ysc, w.schur, st_);

  SchurOps ops;
  ops.m = m;
  ops.n = n;
  ops.jh_rp = S.jch_rp.p;
  ops.jh_ci = S.jch_ci.p;
  ops.jh_v = S.jch_v.p;
  ops.jth_rp = S.jcth_rp.p;
  ops.jth_ci = S.jcth_ci.p;
  ops.jth_v = S.jcth_v.p;
  ops.chol = S.chol.get();
  ops.st = st_;
  ops.t1 = S.t1.p;
  ops.x = w.cx;
  ops.r = S.cr.p;
  ops.w = S.cw.p;
  ops.p = S.cp.p;
  ops.s = S.cs.p;
  ops.sc = S.sc.p;
  ops.part = S.part.p;
  if (x0) dev::copy(w.cx, x0, m, st_);
  else dev::fill(w.cx, 0.0, m, st_);
  *cg_status = schur_cg_run(ops, w.schur, opts.cg_tol, opts.cg_itmax, opts.cg_check_every, iters,
                            fnorm, x0 == nullptr);
  if (after_cg) HIPCHK(hipEventRecord(after_cg, st_));

  dev::spmv(n, S.jcth_rp.p, S.jcth_ci.p, nullptr, S.jcth_v.p, -1.0, w.cx, 1.0, w.rxp, S.t1.p,
            st_);
  chol_solve(S.t1.p);
  dev::gather_scale(dx, S.t1.p, S.ipos_d.p, S.max_d.p, n, st_);
  dev::copy(dyc, w.cx, m, st_);
  dev::vec_scale(dyc, S.max_d.p + n, m, st_);
  dev::recover_d(d, S.jd_rp.p, S.jd_ci.p, S.jdv, S.has_jd ? 1 : 0, dx, ryd, S.dsd, rs, dsv, dyd,
                 st_);
}

// KKT residual of (dx, dsv, dyc, dyd) against the bound right-hand sides:
// out[0..3] = squared norms of rows 1, 3, 2, 4; out[4..7] = squared norms of
// rx, rs, ry, ryd.  rows: also keep the signed rows in e_x, e_y, e_s, e_d.
// Synchronizes the stream.
void Solver::residual(const double* dx, const double* dsv, const double* dyc, const double* dyd,
                      bool rows, hipEvent_t done, double out[8])
{
  Sym& S = *sym_;
  const int n = S.n, m = S.m, d = S.d;
  dev::resid_rows_n(n, S.h_rp.p, S.h_ci.p, S.hv, S.jct_rp.p, S.jct_ci.p, S.jct_map.p, S.jcv,
                    S.has_jd ? S.jdt_rp.p : nullptr, S.jdt_ci.p, S.jdt_map.p, S.jdv, S.drx, dx,
                    dyc, dyd, S.sq_n.p, rows ? S.e_x.p : nullptr, st_);
  dev::resid_rows_m(m, S.jc_rp.p, S.jc_ci.p, S.jcv, S.dry, dx, S.sq_m.p,
                    rows ? S.e_y.p : nullptr, st_);
  dev::resid_rows_d(d, S.jd_rp.p, S.jd_ci.p, S.jdv, S.has_jd ? 1 : 0, S.dsd, S.drs, S.dryd, dx,
                    dsv, dyd, S.sq_d2.p, S.sq_d4.p, rows ? S.e_s.p : nullptr,
                    rows ? S.e_d.p : nullptr, st_);
  double* r = S.res.p;
  dev::dot(n, S.sq_n.p, nullptr, S.part.p, r + 0, st_);
  dev::dot(m, S.sq_m.p, nullptr, S.part.p, r + 1, st_);
  dev::dot(d, S.sq_d2.p, nullptr, S.part.p, r + 2, st_);
  dev::dot(d, S.sq_d4.p, nullptr, S.part.p, r + 3, st_);
  dev::dot(n, S.drx, S.drx, S.part.p, r + 4, st_);
  dev::dot(d, S.drs, S.drs, S.part.p, r + 5, st_);
  dev::dot(m, S.dry, S.dry, S.part.p, r + 6, st_);
  dev::dot(d, S.dryd, S.dryd, S.part.p, r + 7, st_);
  if (done) HIPCHK(hipEventRecord(done, st_));
  HIPCHK(hipMemcpyAsync(out, r, 8 * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
}

void Solver::get_solution(double* dx, double* ds, double* dyc, double* dyd)
{
  HYKKT_REQUIRE(sym_ && stats.n_solves > 0, HYKKT_E_STATE, "no solution yet");
  Sym& S = *sym_;
  auto d2h = [&](double* dst, const DevBuf<double>& b, int cnt) {
    if (dst && cnt) HIPCHK(hipMemcpy(dst, b.p, cnt * sizeof(double), hipMemcpyDeviceToHost));
  };
  d2h(dx, S.dx, S.n);
  d2h(ds, S.dsv, S.d);
  d2h(dyc, S.dyc, S.m);
  d2h(dyd, S.dyd, S.d);
}

void Solver::get_solution_device(const double** dx, const double** ds, const double** dyc,
                                 const double** dyd) const
{
  HYKKT_REQUIRE(sym_ && stats.n_solves > 0, HYKKT_E_STATE, "no solution yet");
  if (dx) *dx = sym_->dx.p;
  if (ds) *ds = sym_->dsv.p;
  if (dyc) *dyc = sym_->dyc.p;
  if (dyd) *dyd = sym_->dyd.p;
}

void Solver::get_perm(int* perm, int n) const
{
  HYKKT_REQUIRE(sym_, HYKKT_E_STATE, "no analysis yet");
  HYKKT_REQUIRE(n == sym_->n, HYKKT_E_ARG, "wrong permutation length");
  std::memcpy(perm, sym_->perm.data(), sizeof(int) * n);
}

long long Solver::debug_copy(int what, void* out, long long cap)
{
  HYKKT_REQUIRE(sym_, HYKKT_E_STATE, "no analysis yet");
  Sym& S = *sym_;
  HIPCHK(hipStreamSynchronize(st_));
  auto dcopy = [&](const double* src, long long cnt) -> long long {
    HYKKT_REQUIRE(cnt <= cap, HYKKT_E_ARG, "debug_copy: buffer too small");
    if (cnt) HIPCHK(hipMemcpy(out, src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    return cnt;
  };
  auto icopy = [&](const std::vector<int>& v) -> long long {
    HYKKT_REQUIRE(static_cast<long long>(v.size()) <= cap, HYKKT_E_ARG,
                  "debug_copy: buffer too small");
    std::memcpy(out, v.data(), v.size() * sizeof(int));
    return static_cast<long long>(v.size());
  };
  DevBuf<double> tmp;
  switch (what) {
    case 1: return dcopy(S.htil.p, S.htp.e.nnz());
    case 2: return icopy(S.htp.e.rp);
    case 3: return icopy(S.htp.e.ci);
    case 4: return dcopy(S.max_d.p, S.n + S.m);
    case 5: return dcopy(S.jcs.p, S.nnz_jc);
    case 6: return dcopy(S.rx_til.p, S.n);
    case 7: return dcopy(S.ry_s.p, S.m);
    case 8: return icopy(S.hgp.e.rp);
    case 9: return icopy(S.hgp.e.ci);
    case 10: {
      // The assembly of the last execute, run again into the factor storage
      // (the factor itself is recomputed by every execute), and read back
      // from there through the assembly map: entry a_src[e] of P*Hg*P^T is
      // at L[a_dst[e]]; upper-triangle positions are NaN.
      const long long nz = S.hgp_pat.nnz();
      HYKKT_REQUIRE(nz <= cap, HYKKT_E_ARG, "debug_copy: buffer too small");
      CholEngine& ch = *S.chol;
      const int na = ch.n_assembled();
      tmp.alloc(std::max(1, na));
      assemble_hgamma();
      dev::gather64(tmp.p, ch.L(), ch.a_dst(), na, st_);
      std::vector<double> got(std::max(1, na));
      HIPCHK(hipMemcpyAsync(got.data(), tmp.p, sizeof(double) * na, hipMemcpyDeviceToHost, st_));
      HIPCHK(hipStreamSynchronize(st_));
      double* o = static_cast<double*>(out);
      std::fill(o, o + nz, std::nan(""));
      const std::vector<int>& src = ch.plan().a_src;
      for (int e = 0; e < na; e++) o[src[e]] = got[e];
      return nz;
    }
    case 11:   // H~ before Ruiz, recomputed from the current inputs
      tmp.alloc(std::max(1, S.htp.e.nnz()));
      if (S.has_jd)
        dev::spgemm_values(S.htp.e.nnz(), S.ht_dsrc.p, S.hv, S.ht_tp.p, S.ht_ta.p, S.ht_tb.p,
                           S.jdv, S.jd_vs.p, 1.0, tmp.p, st_);
      else
        dev::copy(tmp.p, S.hv, S.nnz_h, st_);
      HIPCHK(hipStreamSynchronize(st_));
      return dcopy(tmp.p, S.htp.e.nnz());
    case 12: {  // r^x in natural order
      std::vector<double> a(S.n), b(S.n);
      HIPCHK(hipMemcpy(a.data(), S.rxp.p, S.n * sizeof(double), hipMemcpyDeviceToHost));
      for (int k = 0; k < S.n; k++) b[S.phat[k]] = a[k];
      HYKKT_REQUIRE(S.n <= cap, HYKKT_E_ARG, "debug_copy: buffer too small");
      std::memcpy(out, b.data(), S.n * sizeof(double));
      return S.n;
    }
    case 13: return dcopy(S.schur.p, S.m);
    case 14: return dcopy(S.cx.p, S.m);
    case 15: return icopy(S.hgp_pat.rp);
    case 16: return icopy(S.hgp_pat.ci);
    case 17: return icopy(S.hgp_map);
    case 18: return icopy(S.jcp_pat.ci);
    case 19: return icopy(S.jcp_map);
    case 20: return icopy(S.jctp_pat.rp);
    case 21: return icopy(S.jctp_pat.ci);
    case 22: return icopy(S.jctp_map);
    default: throw Error(HYKKT_E_ARG, "debug_copy: unknown selector");
  }
}

}  // namespace hykkt
