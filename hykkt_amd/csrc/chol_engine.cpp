// chol_engine.cpp -- host orchestration of the device multifrontal Cholesky:
// builds the per-level task lists once per pattern, then launches
// assembly -> (extend-add, blocked panel factorization, update-matrix SYRK,
// extend-add of update matrices) per tree level, and the level-scheduled
// supernodal triangular solves.
#include <algorithm>
#include <climits>

#include "device.hpp"

namespace hykkt {

namespace {
constexpr int NB = 64;
constexpr int TS = 64;
inline int cdiv(int a, int b) { return (a + b - 1) / b; }
}  // namespace

CholEngine::CholEngine(const Pattern& a, const CholOptions& opt, hipStream_t st) : st_(st)
{
  pl_ = chol_analyze(a, opt);
  sn_first_.upload(pl_.sn_first);
  rows_ptr_.upload(pl_.rows_ptr);
  rows_.upload(pl_.rows.empty() ? std::vector<int>{0} : pl_.rows);
  rel_.upload(pl_.rel.empty() ? std::vector<int>{0} : pl_.rel);
  child_ptr_.upload(pl_.child_ptr);
  child_list_.upload(pl_.child_list.empty() ? std::vector<int>{0} : pl_.child_list);
  post_.upload(pl_.post);
  l_off_.upload(pl_.l_off);
  u_off_.upload(pl_.u_off);
  a_src_.upload(pl_.a_src);
  a_dst_.upload(pl_.a_dst);
  std::vector<int64_t> dpos(pl_.n), uvo(pl_.ns + 1, 0);
  for (int s = 0; s < pl_.ns; s++) {
    int f = pl_.nc(s) + pl_.nr(s);
    for (int j = pl_.sn_first[s]; j < pl_.sn_first[s + 1]; j++) {
      int64_t lj = j - pl_.sn_first[s];
      dpos[j] = pl_.l_off[s] + lj * f + lj;
    }
    uvo[s + 1] = uvo[s] + pl_.nr(s);
  }
  diag_pos_.upload(dpos);
  uvec_off_.upload(uvo);
  L_.alloc(static_cast<size_t>(std::max<int64_t>(pl_.l_off[pl_.ns], 1)));
  U_.alloc(static_cast<size_t>(std::max<int64_t>(pl_.u_off[pl_.ns], 1)));
  ubuf_.alloc(static_cast<size_t>(std::max<int64_t>(uvo[pl_.ns], 1)));
  pivot_.alloc(1);
  build_tasks();
  wbuf_.alloc(static_cast<size_t>(max_level_width_) * std::max(pl_.max_front, 1));
  cd_.n = pl_.n;
  cd_.ns = pl_.ns;
  cd_.sn_first = sn_first_.p;
  cd_.rows_ptr = rows_ptr_.p;
  cd_.rows = rows_.p;
  cd_.rel = rel_.p;
  cd_.l_off = l_off_.p;
  cd_.u_off = u_off_.p;
  cd_.child_ptr = child_ptr_.p;
  cd_.child_list = child_list_.p;
  cd_.L = L_.p;
  cd_.U = U_.p;
  cd_.pivot = pivot_.p;
}

void CholEngine::build_tasks()
{
  const int nlev = static_cast<int>(pl_.level_ptr.size()) - 1;
  lv_.assign(nlev, Level{});
  std::vector<int> sns, eap, eacol, eacptr{0}, eacc, eaccol;
  std::vector<int> potrf, trsm_sn, trsm_tile, upd_sn, upd_tr, upd_tc, syrk_sn, syrk_tr, syrk_tc;
  struct Contrib {
    int t, c, r2;
  };
  std::vector<Contrib> con;
  max_level_width_ = 1;
  for (int l = 0; l < nlev; l++) {
    Level& L = lv_[l];
    const int b0 = pl_.level_ptr[l], b1 = pl_.level_ptr[l + 1];
    max_level_width_ = std::max(max_level_width_, b1 - b0);
    L.sns.off = static_cast<int>(sns.size());
    for (int q = b0; q < b1; q++) sns.push_back(pl_.level_sn[q]);
    L.sns.cnt = b1 - b0;
    // extend-add tasks: panel part and update-matrix part
    for (int part = 0; part < 2; part++) {
      Launch& la = part == 0 ? L.ea_panel : L.ea_u;
      la.off = static_cast<int>(eap.size());
      for (int q = b0; q < b1; q++) {
        const int p = pl_.level_sn[q], ncp = pl_.nc(p);
        con.clear();
        for (int t = pl_.child_ptr[p]; t < pl_.child_ptr[p + 1]; t++) {
          const int c = pl_.child_list[t];
          for (int k = pl_.rows_ptr[c]; k < pl_.rows_ptr[c + 1]; k++) {
            int col = pl_.rel[k];
            if ((part == 0) == (col < ncp)) con.push_back({col, c, k - pl_.rows_ptr[c]});
          }
        }
        std::stable_sort(con.begin(), con.end(),
                         [](const Contrib& x, const Contrib& y) { return x.t < y.t; });
        for (size_t i = 0; i < con.size();) {
          size_t j = i;
          while (j < con.size() && con[j].t == con[i].t) {
            eacc.push_back(con[j].c);
            eaccol.push_back(con[j].r2);
            j++;
          }
          eap.push_back(p);
          eacol.push_back(con[i].t);
          eacptr.push_back(static_cast<int>(eacc.size()));
          i = j;
        }
      }
      la.cnt = static_cast<int>(eap.size()) - la.off;
    }
    // blocked panel factorization steps
    int nsteps = 0;
    for (int q = b0; q < b1; q++) nsteps = std::max(nsteps, cdiv(pl_.nc(pl_.level_sn[q]), NB));
    L.potrf.resize(nsteps);
    L.trsm.resize(nsteps);
    L.upd.resize(nsteps);
    for (int j = 0; j < nsteps; j++) {
      L.potrf[j].off = static_cast<int>(potrf.size());
      L.trsm[j].off = static_cast<int>(trsm_sn.size());
      L.upd[j].off = static_cast<int>(upd_sn.size());
      for (int q = b0; q < b1; q++) {
        const int s = pl_.level_sn[q];
        const int nc = pl_.nc(s), f = nc + pl_.nr(s);
        const int k0 = j * NB;
        if (k0 >= nc) continue;
        const int b = std::min(NB, nc - k0);
        potrf.push_back(s);
        const int below = f - (k0 + b);
        for (int t = 0; t < cdiv(below, 64); t++) {
          trsm_sn.push_back(s);
          trsm_tile.push_back(t);
        }
        const int ctiles = cdiv(nc - (k0 + b), TS), rtiles = cdiv(f - (k0 + b), TS);
        for (int tc = 0; tc < ctiles; tc++)
          for (int tr = tc; tr < rtiles; tr++) {
            upd_sn.push_back(s);
            upd_tr.push_back(tr);
            upd_tc.push_back(tc);
          }
      }
      L.potrf[j].cnt = static_cast<int>(potrf.size()) - L.potrf[j].off;
      L.trsm[j].cnt = static_cast<int>(trsm_sn.size()) - L.trsm[j].off;
      L.upd[j].cnt = static_cast<int>(upd_sn.size()) - L.upd[j].off;
    }
    L.syrk.off = static_cast<int>(syrk_sn.size());
    for (int q = b0; q < b1; q++) {
      const int s = pl_.level_sn[q], nt = cdiv(pl_.nr(s), TS);
      for (int tc = 0; tc < nt; tc++)
        for (int tr = tc; tr < nt; tr++) {
          syrk_sn.push_back(s);
          syrk_tr.push_back(tr);
          syrk_tc.push_back(tc);
        }
    }
    L.syrk.cnt = static_cast<int>(syrk_sn.size()) - L.syrk.off;
  }
  auto up = [](DevBuf<int>& b, std::vector<int>& v) {
    if (v.empty()) v.push_back(0);
    b.upload(v);
  };
  up(t_sns_, sns);
  up(t_ea_p_, eap);
  up(t_ea_col_, eacol);
  up(t_ea_cptr_, eacptr);
  up(t_ea_cc_, eacc);
  up(t_ea_ccol_, eaccol);
  up(t_potrf_, potrf);
  up(t_trsm_sn_, trsm_sn);
  up(t_trsm_tile_, trsm_tile);
  up(t_upd_sn_, upd_sn);
  up(t_upd_tr_, upd_tr);
  up(t_upd_tc_, upd_tc);
  up(t_syrk_sn_, syrk_sn);
  up(t_syrk_tr_, syrk_tr);
  up(t_syrk_tc_, syrk_tc);
}

void CholEngine::zero_L() { dev::fill(L_.p, 0.0, static_cast<int64_t>(L_.n), st_); }

void CholEngine::assemble(const double* vals)
{
  zero_L();
  dev::scatter(L_.p, a_dst_.p, vals, a_src_.p, n_assembled(), st_);
}

void CholEngine::shift_diag(double delta)
{
  if (delta != 0.0) dev::add_at(L_.p, diag_pos_.p, pl_.n, delta, st_);
}

void CholEngine::factor(double pivot_tol)
{
  cd_.pivot_tol = pivot_tol;
  dev::fill_i(pivot_.p, INT_MAX, 1, st_);
  for (const Level& L : lv_) {
    dev::extend_add(cd_, L.ea_panel.cnt, t_ea_p_.p + L.ea_panel.off, t_ea_col_.p + L.ea_panel.off,
                    t_ea_cptr_.p + L.ea_panel.off, t_ea_cc_.p, t_ea_ccol_.p, st_);
    for (size_t j = 0; j < L.potrf.size(); j++) {
      dev::potrf_block(cd_, L.potrf[j].cnt, t_potrf_.p + L.potrf[j].off, static_cast<int>(j), st_);
      dev::trsm_block(cd_, L.trsm[j].cnt, t_trsm_sn_.p + L.trsm[j].off,
                      t_trsm_tile_.p + L.trsm[j].off, static_cast<int>(j), st_);
      dev::panel_update(cd_, L.upd[j].cnt, t_upd_sn_.p + L.upd[j].off, t_upd_tr_.p + L.upd[j].off,
                        t_upd_tc_.p + L.upd[j].off, static_cast<int>(j), st_);
    }
    dev::syrk_update(cd_, L.syrk.cnt, t_syrk_sn_.p + L.syrk.off, t_syrk_tr_.p + L.syrk.off,
                     t_syrk_tc_.p + L.syrk.off, st_);
    dev::extend_add(cd_, L.ea_u.cnt, t_ea_p_.p + L.ea_u.off, t_ea_col_.p + L.ea_u.off,
                    t_ea_cptr_.p + L.ea_u.off, t_ea_cc_.p, t_ea_ccol_.p, st_);
  }
}

int CholEngine::pivot()
{
  int h = INT_MAX;
  HIPCHK(hipMemcpyAsync(&h, pivot_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return h == INT_MAX ? -1 : h;
}

void CholEngine::solve(double* x)
{
  const int nlev = static_cast<int>(lv_.size());
  for (int l = 0; l < nlev; l++)
    dev::trsv_forward(cd_, lv_[l].sns.cnt, t_sns_.p + lv_[l].sns.off, x, ubuf_.p, uvec_off_.p,
                      wbuf_.p, pl_.max_front, st_);
  for (int l = nlev - 1; l >= 0; l--)
    dev::trsv_backward(cd_, lv_[l].sns.cnt, t_sns_.p + lv_[l].sns.off, x, wbuf_.p, pl_.max_front,
                       st_);
}

}  // namespace hykkt
