// chol_engine.cpp -- host orchestration of the device multifrontal Cholesky.
//
// Numeric factorization, per tree level (leaves first):
//   extend-add of the children's update matrices into the panels,
//   blocked right-looking panel factorization in 128-column steps
//     (factor + invert the diagonal block; TRSM of the rows below with the
//      inverted block on the matrix cores -- for 128-wide blocks by blocked
//      substitution with the inverses of the 16x16 diagonal blocks, the
//      whole inverse formed after the level for the solves; MFMA update of
//      the panel's trailing columns),
//   U = -L21 L21^T on the matrix cores,
//   extend-add of the children's update matrices into U.
// Solves: one forward and one backward sweep over the whole tree.
#include <algorithm>
#include <queue>
#include <climits>
#include <cstdio>
#include <cstdlib>

#include "device.hpp"

namespace hykkt {

namespace {
constexpr int NB = 128;   // panel step, MFMA tile edge, solve block
constexpr int SR = 32;    // rows of a short tile
constexpr int GMAX = 8;   // most steps one grouped trailing update can cover
constexpr int BWD_OWN_CHUNKS = 2;   // see build_solve
inline int cdiv(int a, int b) { return (a + b - 1) / b; }
}  // namespace

CholEngine::CholEngine(const Pattern& a, const CholOptions& opt, hipStream_t st) : st_(st)
{
  HIPCHK(hipStreamCreate(&st2_));   // default priority, like the solver's stream
  // the two streams' events order work on this device only: a device-scope
  // release, no system-scope cache writeback when they are recorded
  HIPCHK(hipEventCreateWithFlags(&ev_main_, hipEventDisableTiming | hipEventReleaseToDevice));
  HIPCHK(hipEventCreateWithFlags(&ev_rest_, hipEventDisableTiming | hipEventReleaseToDevice));
  HIPCHK(hipEventCreateWithFlags(&ev_zero_, hipEventDisableTiming | hipEventReleaseToDevice));
  pl_ = chol_analyze(a, opt);
  if (const char* pm = std::getenv("HYKKT_PAIR_MIN")) {
    tun_.pair_min = std::max(1, std::atoi(pm));
    pinned_ = true;
  }
  if (const char* gm = std::getenv("HYKKT_GROUP_MAX")) {
    tun_.group_max = std::clamp(std::atoi(gm), 1, GMAX);
    pinned_ = true;
  }
  {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&n_cu_, hipDeviceAttributeMultiprocessorCount, dev));
    n_cu_ = std::max(n_cu_, 1);
  }
  auto upi = [](DevBuf<int>& b, const std::vector<int>& v) {
    if (v.empty()) b.upload(std::vector<int>{0});
    else b.upload(v);
  };
  upi(sn_first_, pl_.sn_first);
  upi(rows_ptr_, pl_.rows_ptr);
  upi(rows_, pl_.rows);
  upi(rel_, pl_.rel);
  upi(child_ptr_, pl_.child_ptr);
  upi(child_list_, pl_.child_list);
  upi(post_, pl_.post);
  upi(sn_parent_, pl_.sn_parent);
  l_off_.upload(pl_.l_off);
  u_off_.upload(pl_.u_off);
  upi(a_src_, pl_.a_src);
  a_dst_.upload(pl_.a_dst.empty() ? std::vector<int64_t>{0} : pl_.a_dst);
  std::vector<int64_t> dpos(std::max(pl_.n, 1)), uvo(pl_.ns + 1, 0), dvo(pl_.ns + 1, 0);
  for (int s = 0; s < pl_.ns; s++) {
    int f = pl_.nc(s) + pl_.nr(s);
    for (int j = pl_.sn_first[s]; j < pl_.sn_first[s + 1]; j++) {
      int64_t lj = j - pl_.sn_first[s];
      dpos[j] = pl_.l_off[s] + lj * f + lj;
    }
    uvo[s + 1] = uvo[s] + pl_.nr(s);
    dvo[s + 1] = dvo[s] + static_cast<int64_t>(cdiv(pl_.nc(s), NB)) * NB * NB;
  }
  diag_pos_.upload(dpos);
  uvec_off_.upload(uvo);
  dinv_off_.upload(dvo);
  n_ubuf_ = uvo[pl_.ns];
  n_dinv_ = dvo[pl_.ns];
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
  cd_.sn_parent = sn_parent_.p;
  cd_.dinv_off = dinv_off_.p;
  build_tasks();
  build_solve();
}

CholEngine::Sizes CholEngine::sizes() const
{
  auto at_least1 = [](int64_t v) { return std::max<int64_t>(v, 1); };
  return Sizes{at_least1(pl_.l_off[pl_.ns]), at_least1(pl_.u_off[pl_.ns]), at_least1(n_dinv_),
               at_least1(pl_.n), at_least1(n_ubuf_), at_least1(n_bpart_)};
}

void CholEngine::set_tuning(const Tuning& t, int count)
{
  HIPCHK(hipStreamSynchronize(st_));   // the task lists of running launches are replaced
  tun_ = t;
  tcount_ = std::max(count, 1);
  build_tasks();
}

void CholEngine::bind(const Storage& s, int64_t vs, int cap)
{
  quiesce();   // a prezero_L of the previous storage may be in flight
  clean_ = 0;
  cd_.L = s.L;
  cd_.U = s.U;
  cd_.dinv = s.dinv;
  cd_.pivot = s.pivot;
  ybuf_ = s.y;
  sd_.ubuf = s.ubuf;
  sd_.bpart = s.bpart;
  vs_ = vs;
  cap_ = cap;
  // only the b x b lower part of an inverted diagonal block is ever written
  dev::fill(cd_.dinv, 0.0, n_dinv_, dev::Batch{cap, vs, nullptr}, st_);
}

void CholEngine::own(int cap)
{
  const Sizes z = sizes();
  auto al = [](int64_t v) { return (v + 31) / 32 * 32; };   // 256-byte slots
  const int64_t vs = al(z.L) + al(z.U) + al(z.dinv) + al(z.y) + al(z.ubuf) + al(z.bpart);
  own_.alloc(static_cast<size_t>(vs) * cap);
  own_piv_.alloc(static_cast<size_t>(cap));
  double* p = own_.p;
  Storage s{};
  s.L = p;
  p += al(z.L);
  s.U = p;
  p += al(z.U);
  s.dinv = p;
  p += al(z.dinv);
  s.y = p;
  p += al(z.y);
  s.ubuf = p;
  p += al(z.ubuf);
  s.bpart = p;
  s.pivot = own_piv_.p;
  bind(s, vs, cap);
}

CholEngine::~CholEngine()
{
  if (st2_) quiesce();
  if (ev_main_) (void)hipEventDestroy(ev_main_);
  if (ev_rest_) (void)hipEventDestroy(ev_rest_);
  if (ev_zero_) (void)hipEventDestroy(ev_zero_);
  if (st2_) (void)hipStreamDestroy(st2_);
}

void CholEngine::build_tasks()
{
  // Launch shapes (see Tuning), counted over the tcount_ systems a launch
  // covers: a step whose TRSM has at most short_max tiles of 128 rows runs
  // its TRSM and next-column update in 32-row tiles (four times the
  // workgroups, each with a quarter of the work, for launches on the critical
  // path); up to group_max consecutive steps share one trailing update with
  // K = G * NB (each C tile read and written once for the group) while that
  // update has at least pair_min * G^2 / 4 tiles of 128 x 128 beyond the G
  // block columns next in line (a group's inner steps update their next
  // block column with K up to (G-1) NB on the critical path, which only a big
  // trailing update repays).
  const int64_t tc = std::max(tcount_, 1);
  const int pair_min = std::max(1, tun_.pair_min);
  const int gmax = std::clamp(tun_.group_max, 1, GMAX);
  const int nlev = static_cast<int>(pl_.level_ptr.size()) - 1;
  lv_.assign(nlev, Level{});
  std::vector<int> eap, eacol, eacptr{0}, eacc, eaccol;
  std::vector<int> potrf, trsm_sn, trsm_tr, trsm_tc, upd_sn, upd_tr, upd_tc;
  std::vector<int> syrk_sn, syrk_tr, syrk_tc, inv_sn, inv_step;
  struct Contrib {
    int t, c, r2;
  };
  std::vector<Contrib> con;
  for (int l = 0; l < nlev; l++) {
    Level& L = lv_[l];
    const int b0 = pl_.level_ptr[l], b1 = pl_.level_ptr[l + 1];
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
            const double e = pl_.nr(con[j].c) - con[j].r2;   // entries of that U column
            la.flops += e;
            la.bytes += 24.0 * e;   // read the child entry, read + write the parent's
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
    int nsteps = 0;
    for (int q = b0; q < b1; q++) nsteps = std::max(nsteps, cdiv(pl_.nc(pl_.level_sn[q]), NB));
    L.potrf.resize(nsteps);
    L.trsm.resize(nsteps);
    L.upd_next.resize(nsteps);
    L.upd_rest.resize(nsteps);
    // Groups: steps a .. a+G-1 defer their blocks' updates of the columns
    // after block a+G-1 and apply them together at the group's last step, K =
    // G NB, while that update is big (its tiles beyond the next G block
    // columns).  Step a+i of a group first updates block column a+i+1 by
    // blocks a .. a+i (K = (i+1) NB: the same tiles as a single step's next-
    // column update).  Single steps (G = 1) update after their own block.
    L.gpos.assign(nsteps, 0);
    L.gsize.assign(nsteps, 1);
    for (int j = 0; j < nsteps;) {
      int G = 1;
      for (int g = std::min(gmax, nsteps - j); g >= 2 && G == 1; g--) {
        int64_t rest = 0;
        for (int q = b0; q < b1; q++) {
          const int s = pl_.level_sn[q];
          const int nc = pl_.nc(s), f = nc + pl_.nr(s), k0 = j * NB;
          if (k0 >= nc) continue;
          const int cs = k0 + std::min(g * NB, nc - k0);
          const int rt = cdiv(f - cs, NB), ct = cdiv(nc - cs, NB);
          for (int tc = g; tc < ct; tc++) rest += rt - tc;
        }
        if (rest * tc >= static_cast<int64_t>(pair_min) * g * g / 4) G = g;
      }
      for (int i = 0; i < G; i++) {
        L.gpos[j + i] = i;
        L.gsize[j + i] = G;
      }
      j += G;
    }
    for (int j = 0; j < nsteps; j++) {
      const int gi = L.gpos[j], G = L.gsize[j];
      const bool last = gi == G - 1;   // the group's update of the columns after it
      L.potrf[j].off = static_cast<int>(potrf.size());
      L.trsm[j].off = static_cast<int>(trsm_sn.size());
      int full_tiles = 0;
      for (int q = b0; q < b1; q++) {
        const int s = pl_.level_sn[q];
        const int nc = pl_.nc(s), f = nc + pl_.nr(s), k0 = j * NB;
        if (k0 < nc) full_tiles += cdiv(f - (k0 + std::min(NB, nc - k0)), NB);
      }
      const bool shrt = full_tiles * tc <= tun_.short_max;
      const int rows_unit = shrt ? SR : NB;
      L.trsm[j].short_rows = shrt;
      L.upd_next[j].short_rows = shrt && !(last && G > 1);
      for (int q = b0; q < b1; q++) {
        const int s = pl_.level_sn[q];
        const int nc = pl_.nc(s), f = nc + pl_.nr(s);
        const int k0 = j * NB;
        if (k0 >= nc) continue;
        const int b = std::min(NB, nc - k0);
        potrf.push_back(s);
        L.potrf[j].maxb = std::max(L.potrf[j].maxb, b);
        // Algorithmic (LAPACK) counts: POTRF b^3/3, TRSM r*b^2 for the r rows
        // below the block, and the lower-trapezoid rank-b update of the c
        // trailing panel columns; bytes read each operand and write each
        // result once.
        const double db = b, r = f - (k0 + b), cc = nc - (k0 + b);
        const double tri = db * (db + 1) / 2, pairs = cc * r - cc * (cc - 1) / 2;
        L.potrf[j].flops += db * db * db / 3 + db * db / 2 + db / 6;
        L.potrf[j].bytes += 16.0 * tri;
        L.trsm[j].flops += r * db * db;
        L.trsm[j].bytes += 8.0 * (2 * r * db + tri);
        // the update splits into the next block column (tile column 0) and
        // the rest, so the next step can start before the rest is done.  In a
        // group, an inner step updates the next block column by the group's
        // blocks so far (K = kw) and the last step the columns after the group
        // by all its blocks: the next G block columns, then the rest
        const int ka = k0 - gi * NB;   // the group's first column
        if (last && G > 1) {
          const int kw = std::min(G * NB, nc - ka);
          const double dk = kw, r2 = f - (ka + kw), c2 = nc - (ka + kw);
          const double p2 = c2 * r2 - c2 * (c2 - 1) / 2;
          const double cn = std::min<double>(G * NB, c2);
          const double pn = cn * r2 - cn * (cn - 1) / 2;
          L.upd_next[j].flops += 2.0 * dk * pn;
          L.upd_next[j].bytes += 8.0 * (r2 * dk + 2 * pn);
          L.upd_rest[j].flops += 2.0 * dk * (p2 - pn);
          L.upd_rest[j].bytes += 8.0 * 2 * (p2 - pn);
        } else {
          const double dk = std::min((gi + 1) * NB, nc - ka);
          const double cn = std::min<double>(NB, cc);
          const double pn = cn * r - cn * (cn - 1) / 2;
          L.upd_next[j].flops += 2.0 * dk * pn;
          L.upd_next[j].bytes += 8.0 * (r * dk + 2 * pn);
          if (G == 1) {
            L.upd_rest[j].flops += 2.0 * db * (pairs - pn);
            L.upd_rest[j].bytes += 8.0 * 2 * (pairs - pn);
          }
        }
        for (int t = 0; t < cdiv(f - (k0 + b), rows_unit); t++) {
          trsm_sn.push_back(s);
          trsm_tr.push_back(t);
          trsm_tc.push_back(0);
        }
      }
      L.potrf[j].cnt = static_cast<int>(potrf.size()) - L.potrf[j].off;
      L.trsm[j].cnt = static_cast<int>(trsm_sn.size()) - L.trsm[j].off;
      for (int part = 0; part < 2; part++) {
        Launch& la = part == 0 ? L.upd_next[j] : L.upd_rest[j];
        la.off = static_cast<int>(upd_sn.size());
        for (int q = b0; q < b1; q++) {
          const int s = pl_.level_sn[q];
          const int nc = pl_.nc(s), f = nc + pl_.nr(s), k0 = j * NB;
          if (k0 >= nc) continue;
          if (G > 1 && !last && part == 1) continue;   // deferred to the group's update
          if (G > 1 && last) {
            // columns after the group's G blocks, K = all of them: the next G
            // block columns (part 0), then the rest (part 1)
            const int ka = k0 - gi * NB, cs = ka + std::min(G * NB, nc - ka);
            const int rtiles = cdiv(f - cs, NB), ctiles = cdiv(nc - cs, NB);
            for (int tc = part == 0 ? 0 : G; tc < (part == 0 ? std::min(G, ctiles) : ctiles); tc++)
              for (int tr = tc; tr < rtiles; tr++) {
                upd_sn.push_back(s);
                upd_tr.push_back(tr);
                upd_tc.push_back(tc);
              }
            continue;
          }
          const int b = std::min(NB, nc - k0);
          const int rtiles = cdiv(f - (k0 + b), NB), ctiles = cdiv(nc - (k0 + b), NB);
          if (part == 0 && shrt) {   // next block column in 32-row tiles
            if (ctiles > 0)
              for (int tr = 0; tr < cdiv(f - (k0 + b), SR); tr++) {
                upd_sn.push_back(s);
                upd_tr.push_back(tr);
                upd_tc.push_back(0);
              }
            continue;
          }
          for (int tc = part == 0 ? 0 : 1; tc < (part == 0 ? std::min(1, ctiles) : ctiles); tc++)
            for (int tr = tc; tr < rtiles; tr++) {
              upd_sn.push_back(s);
              upd_tr.push_back(tr);
              upd_tc.push_back(tc);
            }
        }
        la.cnt = static_cast<int>(upd_sn.size()) - la.off;
      }
    }
    L.syrk.off = static_cast<int>(syrk_sn.size());
    for (int q = b0; q < b1; q++) {
      const int s = pl_.level_sn[q], nt = cdiv(pl_.nr(s), NB);
      const double nr = pl_.nr(s), nc = pl_.nc(s), pairs = nr * (nr + 1) / 2;
      L.syrk.flops += 2.0 * nc * pairs;
      L.syrk.bytes += 8.0 * (nr * nc + pairs);
      for (int tc = 0; tc < nt; tc++)
        for (int tr = tc; tr < nt; tr++) {
          syrk_sn.push_back(s);
          syrk_tr.push_back(tr);
          syrk_tc.push_back(tc);
        }
    }
    L.syrk.cnt = static_cast<int>(syrk_sn.size()) - L.syrk.off;
    // the whole inverses of the blocks of the steps factored without them
    // (128-wide launches): b^3/3 flops, the triangle read twice and written
    L.inv.off = static_cast<int>(inv_sn.size());
    for (int j = 0; j < nsteps; j++) {
      if (L.potrf[j].maxb <= 64) continue;
      for (int q = b0; q < b1; q++) {
        const int s = pl_.level_sn[q], nc = pl_.nc(s);
        if (j * NB >= nc) continue;
        const double db = std::min(NB, nc - j * NB);
        inv_sn.push_back(s);
        inv_step.push_back(j);
        L.inv.flops += db * db * db / 3;
        L.inv.bytes += 24.0 * db * (db + 1) / 2;
      }
    }
    L.inv.cnt = static_cast<int>(inv_sn.size()) - L.inv.off;
  }
  auto up = [](DevBuf<int>& b, std::vector<int>& v) {
    if (v.empty()) v.push_back(0);
    b.upload(v);
  };
  up(t_ea_p_, eap);
  up(t_ea_col_, eacol);
  up(t_ea_cptr_, eacptr);
  up(t_ea_cc_, eacc);
  up(t_ea_ccol_, eaccol);
  up(t_potrf_, potrf);
  up(t_trsm_sn_, trsm_sn);
  up(t_trsm_tr_, trsm_tr);
  up(t_trsm_tc_, trsm_tc);
  up(t_upd_sn_, upd_sn);
  up(t_upd_tr_, upd_tr);
  up(t_upd_tc_, upd_tc);
  up(t_syrk_sn_, syrk_sn);
  up(t_syrk_tr_, syrk_tr);
  up(t_syrk_tc_, syrk_tc);
  up(t_inv_sn_, inv_sn);
  up(t_inv_step_, inv_step);
}

void CholEngine::build_solve()
{
  const int ns = pl_.ns;
  // top of the tree: supernodes wider than one block or with fronts taller
  // than TOP_F rows, and all their ancestors; the rest is a forest closed
  // under descendants
  constexpr int TOP_F = 512;
  std::vector<char> top(ns, 0);
  for (int s = 0; s < ns; s++)
    if (pl_.nc(s) > NB || pl_.nc(s) + pl_.nr(s) > TOP_F) top[s] = 1;
  for (int s = 0; s < ns; s++)
    if (top[s])
      for (int p = pl_.sn_parent[s]; p >= 0 && !top[p]; p = pl_.sn_parent[p]) top[p] = 1;
  std::vector<int> fwd, bwd, fwd_top, bwd_top;
  solve_lower_[0] = solve_lower_[1] = 0.0;
  solve_n_[0] = solve_n_[1] = 0;
  for (int s = 0; s < ns; s++) {   // postorder: children first
    std::vector<int>& f = top[s] ? fwd_top : fwd;
    for (int b = 0; b < cdiv(pl_.nc(s), NB); b++) f.insert(f.end(), {s, 0, b});
    for (int u = 0; u < cdiv(pl_.nr(s), NB); u++) f.insert(f.end(), {s, 1, u});
    const double nc = pl_.nc(s), nr = pl_.nr(s);
    solve_lower_[top[s]] += nc * (nc + 1) / 2 + nr * nc;
    solve_n_[top[s]] += pl_.nc(s);
  }
  std::vector<int64_t> bpo(ns + 1, 0);
  for (int s = ns - 1; s >= 0; s--) {   // parents first: row chunks, then pivot blocks
    std::vector<int>& b = top[s] ? bwd_top : bwd;
    const int np = cdiv(pl_.nc(s), NB), nu = cdiv(pl_.nr(s), NB);
    if (!top[s] && nu <= BWD_OWN_CHUNKS) {
      // a forest supernode (one block) with few row chunks: its pivot item
      // forms the chunks' products itself, one dependency link instead of two
      b.insert(b.end(), {s, 2, 0});
      continue;
    }
    for (int u = 0; u < nu; u++)
      for (int k = 0; k < np; k++) b.insert(b.end(), {s, 0, u * np + k});
    for (int k = np - 1; k >= 0; k--) b.insert(b.end(), {s, 1, k});
  }
  for (int s = 0; s < ns; s++)
    bpo[s + 1] = bpo[s] + static_cast<int64_t>(cdiv(pl_.nr(s), NB)) * pl_.nc(s);
  // pull lists: front row of a parent <- entries of its children's update vectors
  std::vector<int64_t> foff(ns + 1, 0);
  for (int s = 0; s < ns; s++) foff[s + 1] = foff[s] + pl_.nc(s) + pl_.nr(s);
  std::vector<int> cnt(static_cast<size_t>(foff[ns]) + 1, 0);
  std::vector<int64_t> uvo(ns + 1, 0);
  for (int s = 0; s < ns; s++) uvo[s + 1] = uvo[s] + pl_.nr(s);
  for (int c = 0; c < ns; c++) {
    int p = pl_.sn_parent[c];
    if (p < 0) continue;
    for (int k = pl_.rows_ptr[c]; k < pl_.rows_ptr[c + 1]; k++) cnt[foff[p] + pl_.rel[k] + 1]++;
  }
  for (size_t i = 0; i + 1 < cnt.size(); i++) cnt[i + 1] += cnt[i];
  std::vector<int> src(cnt.back() > 0 ? cnt.back() : 1), fill(cnt.begin(), cnt.end() - 1);
  for (int p = 0; p < ns; p++)
    for (int t = pl_.child_ptr[p]; t < pl_.child_ptr[p + 1]; t++) {
      const int c = pl_.child_list[t];
      for (int k = pl_.rows_ptr[c]; k < pl_.rows_ptr[c + 1]; k++)
        src[fill[foff[p] + pl_.rel[k]]++] =
            static_cast<int>(uvo[c] + (k - pl_.rows_ptr[c]));
    }
  auto upi = [](DevBuf<int>& b, std::vector<int>& v) {
    if (v.empty()) v.push_back(0);
    b.upload(v);
  };
  sd_.n_fwd = static_cast<int>(fwd.size() / 3);
  sd_.n_bwd = static_cast<int>(bwd.size() / 3);
  sd_.n_fwd_top = static_cast<int>(fwd_top.size() / 3);
  sd_.n_bwd_top = static_cast<int>(bwd_top.size() / 3);
  upi(fwd_items_, fwd);
  upi(bwd_items_, bwd);
  upi(fwd_top_items_, fwd_top);
  upi(bwd_top_items_, bwd_top);
  upi(pull_ptr_, cnt);
  upi(pull_src_, src);
  front_off_.upload(foff);
  bpart_off_.upload(bpo);
  flags_.alloc(5);   // wait-bound error word, four sweep item counters
  sd_.fwd_items = fwd_items_.p;
  sd_.bwd_items = bwd_items_.p;
  sd_.fwd_items_top = fwd_top_items_.p;
  sd_.bwd_items_top = bwd_top_items_.p;
  sd_.uvec_off = uvec_off_.p;
  sd_.pull_ptr = pull_ptr_.p;
  sd_.pull_src = pull_src_.p;
  sd_.front_off = front_off_.p;
  sd_.bpart_off = bpart_off_.p;
  sd_.err = flags_.p;
  sd_.ticket = flags_.p + 1;
  grid_fwd_full_ = n_cu_ * dev::forest_blocks_per_cu();
  sd_.grid_fwd = grid_fwd_full_;
  share_ = 1;
  n_bpart_ = bpo[ns];
  dev::fill_i(flags_.p, 0, 1, st_);
}

void CholEngine::zero_L(const dev::Batch& b)
{
  if (b.count <= clean_) HIPCHK(hipStreamWaitEvent(st_, ev_zero_, 0));
  else dev::fill(cd_.L, 0.0, pl_.l_off[pl_.ns], b, st_);
  clean_ = 0;
}

void CholEngine::prezero_L(int count)
{
  count = std::min(count, cap_);
  dev::fill(cd_.L, 0.0, pl_.l_off[pl_.ns], dev::Batch{count, vs_, nullptr}, st2_);
  HIPCHK(hipEventRecord(ev_zero_, st2_));
  clean_ = count;
}

void CholEngine::assemble(const double* vals, const dev::Batch& b)
{
  zero_L(b);
  dev::scatter(cd_.L, a_dst_.p, vals, a_src_.p, n_assembled(), b, st_);
}

void CholEngine::shift_diag(double delta, const dev::Batch& b)
{
  if (delta != 0.0) dev::add_at(cd_.L, diag_pos_.p, pl_.n, delta, b, st_);
}

void CholEngine::factor(double pivot_tol, const dev::Batch& bt)
{
  HYKKT_REQUIRE(cd_.L && bt.count <= cap_, HYKKT_E_STATE, "factor storage not bound");
  const dev::Batch b{bt.count, vs_, nullptr};
  cd_.pivot_tol = pivot_tol;
  dev::fill_i(cd_.pivot, INT_MAX, b.count, st_);
  Profiler* pf = prof_ && prof_->on() ? prof_ : nullptr;
  // Bracket one launch (skipped when it has no tasks, like the launchers);
  // charged with the work of all the systems it covers.
  auto timed_on = [&](hipStream_t st, int kind, const Launch& la, auto&& launch) {
    if (la.cnt <= 0) return;
    if (pf) pf->begin(st);
    launch();
    if (pf) pf->end(kind, st, la.flops * b.count, la.bytes * b.count);
  };
  auto timed = [&](int kind, const Launch& la, auto&& launch) { timed_on(st_, kind, la, launch); };
  for (const Level& L : lv_) {
    timed(PK_EXTEND_ADD, L.ea_panel, [&] {
      dev::extend_add(cd_, L.ea_panel.cnt, t_ea_p_.p + L.ea_panel.off,
                      t_ea_col_.p + L.ea_panel.off, t_ea_cptr_.p + L.ea_panel.off, t_ea_cc_.p,
                      t_ea_ccol_.p, b, st_);
    });
    // Blocked steps with look-ahead.  Main stream: diagonal blocks, TRSM,
    // update of the next block column(s); second stream: the rest of the
    // trailing update, which overlaps the next steps' diagonal blocks and
    // TRSMs.  A main-stream update waits for the pending rests (ev_rest_)
    // only when it writes a block column they write (rest_lo: the lowest
    // such column); a rest waits for its step's TRSM and next-column update
    // (ev_main_).  Both events are re-recorded after their waits were
    // enqueued.
    bool rest_pending = false;
    int rest_lo = INT_MAX;
    for (size_t j = 0; j < L.potrf.size(); j++) {
      const int step = static_cast<int>(j);
      // in a group of G starting at step a: step a+i updates block column
      // step+1 by blocks a .. a+i (mode i+2 for K = i+1 blocks from block a,
      // mode 0 for one block); the last step block columns step+1 .. step+G
      // by all G blocks (mode G+1), its rest step+G+1 on
      const int gi = L.gpos[j], G = L.gsize[j], a = step - gi;
      const bool last = gi == G - 1;
      const int kblk = last ? G : gi + 1;
      const int umode = kblk == 1 ? 0 : kblk + 1, ustep = a;
      const int upd_hi = (last && G > 1) ? step + G : step + 1;
      timed(PK_POTRF, L.potrf[j], [&] {
        dev::potrf_inv(cd_, L.potrf[j].cnt, t_potrf_.p + L.potrf[j].off, step, L.potrf[j].maxb, b,
                       st_);
      });
      if (L.potrf[j].maxb > 64) {
        // 128-wide blocks: blocked substitution, no whole inverse needed yet
        timed(L.trsm[j].short_rows ? PK_TRSM_SHORT : PK_TRSM, L.trsm[j], [&] {
          dev::trsm(cd_, L.trsm[j].cnt, t_trsm_sn_.p + L.trsm[j].off,
                    t_trsm_tr_.p + L.trsm[j].off, step, b, st_, L.trsm[j].short_rows);
        });
      } else {
        timed(L.trsm[j].short_rows ? PK_DENSE_NT_SHORT : PK_DENSE_NT, L.trsm[j], [&] {
          dev::dense_nt(cd_, L.trsm[j].cnt, t_trsm_sn_.p + L.trsm[j].off,
                        t_trsm_tr_.p + L.trsm[j].off, t_trsm_tc_.p + L.trsm[j].off, 2, step, b,
                        st_, L.trsm[j].short_rows);
        });
      }
      if (rest_pending && upd_hi >= rest_lo) {
        HIPCHK(hipStreamWaitEvent(st_, ev_rest_, 0));
        rest_pending = false;
        rest_lo = INT_MAX;
      }
      // the rest waits for this step's next-column update: its event is
      // recorded by that launch's own completion where there is one
      const bool rest = L.upd_rest[j].cnt > 0;
      const bool on_launch = rest && L.upd_next[j].cnt > 0;
      timed(L.upd_next[j].short_rows ? PK_DENSE_NT_SHORT : PK_DENSE_NT, L.upd_next[j], [&] {
        dev::dense_nt(cd_, L.upd_next[j].cnt, t_upd_sn_.p + L.upd_next[j].off,
                      t_upd_tr_.p + L.upd_next[j].off, t_upd_tc_.p + L.upd_next[j].off, umode,
                      ustep, b, st_, L.upd_next[j].short_rows, on_launch ? ev_main_ : nullptr);
      });
      if (rest) {
        if (!on_launch) HIPCHK(hipEventRecord(ev_main_, st_));
        HIPCHK(hipStreamWaitEvent(st2_, ev_main_, 0));
        timed_on(st2_, PK_DENSE_NT, L.upd_rest[j], [&] {
          dev::dense_nt(cd_, L.upd_rest[j].cnt, t_upd_sn_.p + L.upd_rest[j].off,
                        t_upd_tr_.p + L.upd_rest[j].off, t_upd_tc_.p + L.upd_rest[j].off, umode,
                        ustep, b, st2_, false, ev_rest_);
        });
        rest_pending = true;
        rest_lo = std::min(rest_lo, upd_hi + 1);   // the rest starts after the columns just updated
      }
    }
    if (rest_pending) HIPCHK(hipStreamWaitEvent(st_, ev_rest_, 0));
    timed(PK_INV_BLOCKS, L.inv, [&] {
      dev::inv_blocks(cd_, L.inv.cnt, t_inv_sn_.p + L.inv.off, t_inv_step_.p + L.inv.off, b, st_);
    });
    timed(PK_DENSE_NT, L.syrk, [&] {
      dev::dense_nt(cd_, L.syrk.cnt, t_syrk_sn_.p + L.syrk.off, t_syrk_tr_.p + L.syrk.off,
                    t_syrk_tc_.p + L.syrk.off, 1, 0, b, st_);
    });
    timed(PK_EXTEND_ADD, L.ea_u, [&] {
      dev::extend_add(cd_, L.ea_u.cnt, t_ea_p_.p + L.ea_u.off, t_ea_col_.p + L.ea_u.off,
                      t_ea_cptr_.p + L.ea_u.off, t_ea_cc_.p, t_ea_ccol_.p, b, st_);
    });
  }
}

std::vector<int> CholEngine::pivots(int count)
{
  std::vector<int> h(std::max(count, 1), INT_MAX);
  HIPCHK(hipMemcpyAsync(h.data(), cd_.pivot, sizeof(int) * count, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  for (int& v : h) v = (v == INT_MAX) ? -1 : v;
  h.resize(count);
  return h;
}

void CholEngine::clear_solve_error() { dev::fill_i(sd_.err, 0, 1, st_); }

void CholEngine::set_sweep_share(int share)
{
  share = std::max(share, 1);
  if (share == share_) return;
  share_ = share;
  if (share == 1) {
    sd_.grid_fwd = grid_fwd_full_;
    sd_.cap_forest = sd_.cap_top = 0;
    return;
  }
  // the forest sweeps: their share of the 256-thread workgroups the device
  // holds; the top of the tree: of two 1024-thread workgroups per CU
  sd_.grid_fwd = std::max(1, grid_fwd_full_ / share);
  sd_.cap_forest = sd_.grid_fwd;
  sd_.cap_top = std::max(1, 2 * n_cu_ / share);
}

int CholEngine::solve_error()
{
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, sd_.err, sizeof(int), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  if (h) dev::fill_i(sd_.err, 0, 1, st_);
  return h;
}

void CholEngine::solve(double* x, const dev::Batch& bt)
{
  const dev::Batch b{bt.count, vs_, bt.mask};
  Profiler* pf = prof_ && prof_->on() ? prof_ : nullptr;
  const int64_t nu = n_ubuf_;
  // Algorithmic, per part of the tree: each sweep reads that part's factor
  // entries (lower trapezoids of its panels) once, 2 flops per entry, and
  // reads + writes its vector entries; charged for every system of the
  // launch (those a mask skips included).
  auto sweep = [&](int kind, int part, bool fwd) {
    if ((part == 0 ? (fwd ? sd_.n_fwd : sd_.n_bwd) : (fwd ? sd_.n_fwd_top : sd_.n_bwd_top)) <= 0)
      return;
    if (pf) pf->begin(st_);
    if (fwd) dev::solve_forward(cd_, sd_, x, ybuf_, part, b, st_);
    else dev::solve_backward(cd_, sd_, ybuf_, x, part, b, st_);
    if (pf)
      pf->end(kind, st_, 2.0 * solve_lower_[part] * b.count,
              8.0 * (solve_lower_[part] + 2.0 * solve_n_[part]) * b.count);
  };
  // forward: b in x -> y in ybuf; backward: y in ybuf -> x.  x holds b until
  // the forward sweep is done, so it is filled after it; the other buffers
  // (and the item counters) before it, in one launch.
  dev::fill_sentinel3(ybuf_, pl_.n, sd_.ubuf, nu, sd_.bpart, n_bpart_, b, st_, sd_.ticket);
  sweep(PK_SOLVE_FWD_FOREST, 0, true);
  sweep(PK_SOLVE_FWD_TOP, 1, true);
  dev::fill_sentinel(x, pl_.n, b, st_);
  sweep(PK_SOLVE_BWD_TOP, 1, false);
  sweep(PK_SOLVE_BWD_FOREST, 0, false);
}

}  // namespace hykkt
