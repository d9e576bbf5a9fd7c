// cg.cpp -- Schur-complement conjugate gradient on the device, for a batch of
// same-pattern systems.
//
// Same recurrences as SchurComplementConjugateGradient.cpp:102-194
// (Chronopoulos-Gear form: one S-apply and two dots per iteration):
//   p = beta p + r;  s = beta s + w;  x += alpha p;  r -= alpha s;
//   gam1 = r.r;  stop if sqrt(gam1) < tol;  w = S r;  delta = w.r;
//   beta = gam1/gam;  gam = gam1;  alpha = gam/(delta - beta gam/alpha).
// Device-resident: the vector update carries the r.r partials, the second
// SpMV of the S-apply carries the w.r partials, and the stopping test and
// the scalar recurrences run in the one-workgroup final reductions, which
// clear a system's live flag when it stops; every later kernel skips it.
// The host only reads the live flags between chunks of iterations.
#include <algorithm>
#include <cmath>

#include "device.hpp"

namespace hykkt {

namespace {

// w = S v = Jh A^{-1} Jh^T v, or out = alpha*S v + beta*y_in; with dot_w, the
// partials of out.dot_w go to o.part
void apply_s(SchurOps& o, const double* v, double alpha, double beta, const double* y_in,
             double* out, const dev::Batch& b, const double* dot_w = nullptr)
{
  const double fl = 2.0 * o.nnz * b.count;
  if (o.prof) o.prof->begin(o.st);
  dev::spmv(o.n, o.jth_rp, o.jth_ci, nullptr, o.jth_v, 1.0, nullptr, v, 0.0, nullptr, o.t1,
            o.jth_blk, o.jth_nblk, b, o.st);
  if (o.prof)
    o.prof->end(PK_SPMV, o.st, fl, spmv_bytes(o.n, o.m, o.nnz, false, false, false) * b.count);
  o.chol->solve(o.t1, b);
  if (o.prof) o.prof->begin(o.st);
  dev::spmv(o.m, o.jh_rp, o.jh_ci, nullptr, o.jh_v, alpha, nullptr, o.t1, beta, y_in, out,
            o.jh_blk, o.jh_nblk, b, o.st, dot_w, dot_w ? o.part : nullptr);
  if (o.prof)
    o.prof->end(PK_SPMV, o.st, fl,
                spmv_bytes(o.m, o.n, o.nnz, false, y_in != nullptr, dot_w != nullptr) * b.count);
}

}  // namespace

int schur_partials(int m, int jh_nblk) { return std::max(dev::reduce_blocks(m), jh_nblk); }

std::vector<CgResult> schur_cg_run(SchurOps& o, const double* b, double tol, int itmax,
                                   const dev::Batch& bt, bool x0_zero, int chunk, int chunk_next)
{
  const int m = o.m, cnt = bt.count;
  const dev::Batch live = dev::masked(bt, o.live);
  const dev::CgFlags fl{o.live, o.iters, o.status};
  std::vector<CgResult> res(cnt);
  // every system taking part starts live; the others stay out
  if (bt.mask)
    HIPCHK(hipMemcpyAsync(o.live, bt.mask, sizeof(int) * cnt, hipMemcpyDeviceToDevice, o.st));
  else
    dev::fill_i(o.live, 1, cnt, o.st);
  // gam and gam1 of every system start at 0, so one that takes no part
  // reports ||r|| = 0, not a previous pass's value
  dev::fill(o.sc, 0.0, 2, dev::Batch{cnt, bt.vs, nullptr}, o.st);
  dev::fill_i(o.iters, 0, cnt, o.st);
  dev::fill_i(o.status, 0, cnt, o.st);
  // setup(): p = s = b (:89-100); r = b - S x0 (:104-122).  With x0 = 0 the
  // product is exactly zero and fma(-1, 0, b) = b, so r = b bit for bit and
  // the solve is skipped.
  dev::copy(o.p, b, m, live, o.st);
  dev::copy(o.s, b, m, live, o.st);
  if (o.tq) {
    dev::fill(o.tq, 0.0, o.n, live, o.st);
    dev::fill(o.tz, 0.0, o.n, live, o.st);
  }
  if (x0_zero) {
    dev::copy(o.r, b, m, live, o.st);
  } else {
    apply_s(o, o.x, -1.0, 1.0, b, o.r, live);
    if (o.tz) dev::copy(o.tz, o.t1, o.n, live, o.st);   // A^{-1} Jh^T x0
  }
  // gam = r.r; r.r == 0: x0 already solves the system (the reference would
  // divide 0/0 here) and the system stops with 0 iterations
  dev::dot_partial(m, o.r, o.r, o.part, live, o.st);
  dev::dot_final(dev::reduce_blocks(m), o.part, o.sc, 0, true, false, -1, tol, itmax, fl, live,
                 o.st);
  // w = S r, delta = w.r, alpha = gam/delta (:124-144)
  apply_s(o, o.r, 1.0, 0.0, nullptr, o.w, live, o.r);
  dev::dot_final(o.jh_nblk, o.part, o.sc, 2, false, true, -1, tol, itmax, fl, live, o.st);
  // host copies: pinned when the caller gave memory for them
  std::vector<double> own(o.pinned ? 0 : 4 * static_cast<size_t>(cnt));
  double* hostmem = o.pinned ? o.pinned : own.data();
  int* lv = reinterpret_cast<int*>(hostmem);   // cnt ints; the results reuse the memory
  int next_check = std::max(1, chunk);
  for (int i = 0; i < itmax; i++) {
    dev::cg_update_dot(m, o.sc, o.p, o.s, o.x, o.r, o.w, o.part, live, o.st);
    if (o.tq) dev::cg_track(o.n, o.sc, o.tq, o.tz, o.t1, live, o.st);
    dev::dot_final(dev::reduce_blocks(m), o.part, o.sc, 1, true, false, i, tol, itmax, fl, live,
                   o.st);
    apply_s(o, o.r, 1.0, 0.0, nullptr, o.w, live, o.r);
    dev::dot_final(o.jh_nblk, o.part, o.sc, 2, false, true, i, tol, itmax, fl, live, o.st);
    if (i + 1 == next_check && i + 1 < itmax) {
      HIPCHK(hipMemcpyAsync(lv, o.live, sizeof(int) * cnt, hipMemcpyDeviceToHost, o.st));
      HIPCHK(hipStreamSynchronize(o.st));
      if (std::none_of(lv, lv + cnt, [](int v) { return v != 0; })) break;
      next_check += std::max(1, chunk_next);
    }
  }
  // per system: iterations, status and ||r|| at exit (gam1, or gam for a
  // system that stopped at the start), packed on the device, one copy
  dev::cg_pack(o.gath, o.sc, o.iters, o.status, dev::Batch{cnt, bt.vs, nullptr}, o.st);
  HIPCHK(hipMemcpyAsync(hostmem, o.gath, sizeof(double) * 4 * cnt, hipMemcpyDeviceToHost, o.st));
  if (o.defer && o.pinned) return res;
  HIPCHK(hipStreamSynchronize(o.st));
  cg_unpack(hostmem, res);
  return res;
}

void cg_unpack(const double* pinned, std::vector<CgResult>& res)
{
  for (size_t k = 0; k < res.size(); k++) {
    const double* h = pinned + 4 * k;
    res[k].iters = static_cast<int>(h[2]);
    res[k].status = static_cast<int>(h[3]);
    res[k].final_norm = std::sqrt(res[k].iters == 0 ? h[0] : h[1]);
  }
}

}  // namespace hykkt
