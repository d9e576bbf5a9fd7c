// capi.cpp -- extern "C" boundary of the library (declared in include/hykkt.h).
// Exceptions never cross it: each entry point returns an error code and keeps
// the message for hykkt_last_error().
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hykkt.h"
#include "device.hpp"
#include "mmio.hpp"
#include "solver.hpp"

struct hykkt_solver_s {
  hykkt::Solver impl;
  explicit hykkt_solver_s(double g, int cap = 1) : impl(g, cap) {}
};

struct hykkt_chol_s {
  hipStream_t st = nullptr;
  std::unique_ptr<hykkt::CholEngine> eng;
  hykkt::DevBuf<double> vals, x;
  std::vector<int> post;
  int n = 0, nnz = 0;
  ~hykkt_chol_s()
  {
    eng.reset();
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f)
{
  try {
    g_err.clear();
    return f();
  } catch (const hykkt::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host out of memory";
    return HYKKT_E_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return HYKKT_E_STATE;
  }
}

#define NEED(p) HYKKT_REQUIRE((p) != nullptr, HYKKT_E_ARG, "null handle or pointer: " #p)

}  // namespace

extern "C" {

const char* hykkt_version(void) { return "hykkt-mi355x 0.1.0 (gfx950)"; }
const char* hykkt_last_error(void) { return g_err.c_str(); }

int hykkt_default_options(hykkt_options* o)
{
  return guard([&] {
    NEED(o);
    o->ruiz_its = 2;
    o->norm_tol = 1e-2;
    o->pivot_tol = 1e-12;
    o->cg_tol = 1e-12;
    o->cg_itmax = 100;
    o->delta = 0.0;
    o->inputs_on_device = 0;
    o->cg_check_every = 0;
    o->verbose = 0;
    o->refine_max = 3;
    o->refine_tol = 9e-9;   // the 1e-8 accuracy target less a 10% margin
    o->autotune = 1;
    o->sweep_share = 1;
    return HYKKT_OK;
  });
}

int hykkt_create(double gamma, hykkt_solver* out)
{
  return guard([&] {
    NEED(out);
    *out = new hykkt_solver_s(gamma);
    return HYKKT_OK;
  });
}

int hykkt_destroy(hykkt_solver s)
{
  return guard([&] {
    delete s;
    return HYKKT_OK;
  });
}

#define SETBLK(name, id)                                                                   \
  int name(hykkt_solver s, const int* rp, const int* ci, const double* v, int n, int m,   \
           int nnz)                                                                        \
  {                                                                                        \
    return guard([&] {                                                                     \
      NEED(s);                                                                             \
      s->impl.set_block(id, rp, ci, v, n, m, nnz);                                         \
      return HYKKT_OK;                                                                     \
    });                                                                                    \
  }
SETBLK(hykkt_set_H, hykkt::BLK_H)
SETBLK(hykkt_set_Ds, hykkt::BLK_DS)
SETBLK(hykkt_set_Jc, hykkt::BLK_JC)
SETBLK(hykkt_set_Jd, hykkt::BLK_JD)

#define SETVEC(name, id)                                   \
  int name(hykkt_solver s, const double* v, int n)         \
  {                                                        \
    return guard([&] {                                     \
      NEED(s);                                             \
      s->impl.set_vec(id, v, n);                           \
      return HYKKT_OK;                                     \
    });                                                    \
  }
SETVEC(hykkt_set_rx, hykkt::VEC_RX)
SETVEC(hykkt_set_rs, hykkt::VEC_RS)
SETVEC(hykkt_set_ry, hykkt::VEC_RY)
SETVEC(hykkt_set_ryd, hykkt::VEC_RYD)
SETVEC(hykkt_set_x_host, hykkt::VEC_X)
SETVEC(hykkt_set_s_host, hykkt::VEC_S)
SETVEC(hykkt_set_y_host, hykkt::VEC_Y)
SETVEC(hykkt_set_yd_host, hykkt::VEC_YD)

int hykkt_set_gamma(hykkt_solver s, double gamma)
{
  return guard([&] {
    NEED(s);
    s->impl.set_gamma(gamma);
    return HYKKT_OK;
  });
}

int hykkt_read_matrix_files(hykkt_solver s, const char* h, const char* ds, const char* jc,
                            const char* jd, const char* rx, const char* rs, const char* ry,
                            const char* ryd, int skip_lines)
{
  return guard([&] {
    NEED(s);
    const char* f[8] = {h, ds, jc, jd, rx, rs, ry, ryd};
    for (const char* p : f) NEED(p);
    s->impl.read_files(f, skip_lines);
    return HYKKT_OK;
  });
}

int hykkt_execute(hykkt_solver s)
{
  return guard([&] {
    NEED(s);
    return s->impl.execute();
  });
}

int hykkt_profile_enable(hykkt_solver s, int on)
{
  return guard([&] {
    NEED(s);
    s->impl.prof.enable(on != 0);
    return HYKKT_OK;
  });
}

int hykkt_profile_get(hykkt_solver s, hykkt_kernel_time* out, int capacity, int* count)
{
  return guard([&] {
    NEED(s);
    NEED(count);
    HYKKT_REQUIRE(capacity >= 0 && (out || capacity == 0), HYKKT_E_ARG, "bad output buffer");
    // the kernels launched since profiling was enabled
    int n = 0;
    for (int k = 0; k < hykkt::PK_COUNT; k++) {
      const auto& t = s->impl.prof.total(k);
      if (t.launches == 0) continue;
      if (n < capacity) {
        std::memset(out[n].name, 0, sizeof(out[n].name));
        std::strncpy(out[n].name, hykkt::Profiler::name(k), sizeof(out[n].name) - 1);
        out[n].launches = t.launches;
        out[n].ms = t.ms;
        out[n].flops = t.flops;
        out[n].bytes = t.bytes;
      }
      n++;
    }
    *count = n;
    return HYKKT_OK;
  });
}

int hykkt_set_options(hykkt_solver s, const hykkt_options* o)
{
  return guard([&] {
    NEED(s);
    NEED(o);
    HYKKT_REQUIRE(o->ruiz_its >= 0 && o->cg_itmax >= 0 && o->cg_check_every >= 0 &&
                      o->refine_max >= 0,
                  HYKKT_E_ARG, "negative iteration count");
    HYKKT_REQUIRE(o->sweep_share >= 1, HYKKT_E_ARG, "sweep_share must be at least 1");
    s->impl.opts = *o;
    return HYKKT_OK;
  });
}

int hykkt_get_tuning(hykkt_solver s, hykkt_tuning* t)
{
  return guard([&] {
    NEED(s);
    NEED(t);
    *t = s->impl.tuning();
    return HYKKT_OK;
  });
}

int hykkt_get_options(hykkt_solver s, hykkt_options* o)
{
  return guard([&] {
    NEED(s);
    NEED(o);
    *o = s->impl.opts;
    return HYKKT_OK;
  });
}

int hykkt_get_stats(hykkt_solver s, hykkt_stats* st)
{
  return guard([&] {
    NEED(s);
    NEED(st);
    *st = s->impl.stats;
    return HYKKT_OK;
  });
}

int hykkt_get_solution(hykkt_solver s, double* dx, double* ds, double* dyc, double* dyd)
{
  return guard([&] {
    NEED(s);
    s->impl.get_solution(dx, ds, dyc, dyd);
    return HYKKT_OK;
  });
}

int hykkt_get_solution_device(hykkt_solver s, const double** dx, const double** ds,
                              const double** dyc, const double** dyd)
{
  return guard([&] {
    NEED(s);
    s->impl.get_solution_device(0, dx, ds, dyc, dyd);
    return HYKKT_OK;
  });
}

// ------------------------------------------------------------------ batched form
int hykkt_create_batch(double gamma, int capacity, hykkt_solver* out)
{
  return guard([&] {
    NEED(out);
    HYKKT_REQUIRE(capacity >= 1, HYKKT_E_ARG, "batch capacity must be at least 1");
    *out = new hykkt_solver_s(gamma, capacity);
    return HYKKT_OK;
  });
}

int hykkt_batch_capacity(hykkt_solver s, int* capacity)
{
  return guard([&] {
    NEED(s);
    NEED(capacity);
    *capacity = s->impl.capacity();
    return HYKKT_OK;
  });
}

int hykkt_bind_system(hykkt_solver s, int k, double gamma, const double* h_vals,
                      const double* ds_vals, const double* jc_vals, const double* jd_vals,
                      const double* rx, const double* rs, const double* ry, const double* ryd,
                      const double* y0)
{
  return guard([&] {
    NEED(s);
    const double* v[hykkt::IN_COUNT] = {h_vals, ds_vals, jc_vals, jd_vals, rx, rs, ry, ryd, y0};
    s->impl.bind_system(k, gamma, v);
    return HYKKT_OK;
  });
}

int hykkt_execute_batch(hykkt_solver s, int count, int* status)
{
  return guard([&] {
    NEED(s);
    NEED(status);
    const std::vector<int> rc = s->impl.execute_batch(count);
    std::copy(rc.begin(), rc.end(), status);
    return HYKKT_OK;
  });
}

int hykkt_get_stats_k(hykkt_solver s, int k, hykkt_stats* st)
{
  return guard([&] {
    NEED(s);
    NEED(st);
    *st = s->impl.stats_of(k);
    return HYKKT_OK;
  });
}

int hykkt_get_solution_k(hykkt_solver s, int k, double* dx, double* ds, double* dyc, double* dyd)
{
  return guard([&] {
    NEED(s);
    s->impl.get_solution(k, dx, ds, dyc, dyd);
    return HYKKT_OK;
  });
}

int hykkt_get_solution_device_k(hykkt_solver s, int k, const double** dx, const double** ds,
                                const double** dyc, const double** dyd)
{
  return guard([&] {
    NEED(s);
    s->impl.get_solution_device(k, dx, ds, dyc, dyd);
    return HYKKT_OK;
  });
}

int hykkt_get_permutation(hykkt_solver s, int* perm, int n)
{
  return guard([&] {
    NEED(s);
    NEED(perm);
    s->impl.get_perm(perm, n);
    return HYKKT_OK;
  });
}

int hykkt_set_permutation(hykkt_solver s, const int* perm, int n)
{
  return guard([&] {
    NEED(s);
    NEED(perm);
    s->impl.set_perm(perm, n);
    return HYKKT_OK;
  });
}

long long hykkt_debug_copy(hykkt_solver s, int what, void* out, long long cap)
{
  long long r = 0;
  int rc = guard([&] {
    NEED(s);
    NEED(out);
    r = s->impl.debug_copy(what, out, cap);
    return HYKKT_OK;
  });
  return rc == HYKKT_OK ? r : rc;
}

int hykkt_device_malloc(void** ptr, unsigned long long bytes)
{
  return guard([&] {
    NEED(ptr);
    *ptr = nullptr;
    if (bytes) HIPCHK(hipMalloc(ptr, bytes));
    return HYKKT_OK;
  });
}

int hykkt_device_free(void* ptr)
{
  return guard([&] {
    if (ptr) HIPCHK(hipFree(ptr));
    return HYKKT_OK;
  });
}

int hykkt_memcpy_htod(void* dst, const void* src, unsigned long long bytes)
{
  return guard([&] {
    if (bytes) HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return HYKKT_OK;
  });
}

int hykkt_mm_read_csr(const char* path, int skip_lines, int symmetric, int* n, int* m, int* nnz,
                      int* rowptr, int* colidx, double* vals)
{
  return guard([&] {
    NEED(path);
    NEED(n);
    NEED(m);
    NEED(nnz);
    hykkt::CsrHost c = hykkt::read_mm_csr(path, skip_lines, symmetric != 0);
    *n = c.n;
    *m = c.m;
    *nnz = c.nnz();
    if (rowptr) std::copy(c.rp.begin(), c.rp.end(), rowptr);
    if (colidx) std::copy(c.ci.begin(), c.ci.end(), colidx);
    if (vals) std::copy(c.v.begin(), c.v.end(), vals);
    return HYKKT_OK;
  });
}

int hykkt_mm_read_rhs(const char* path, int* n, double* vals)
{
  return guard([&] {
    NEED(path);
    NEED(n);
    std::vector<double> v = hykkt::read_rhs(path);
    *n = static_cast<int>(v.size());
    if (vals) std::copy(v.begin(), v.end(), vals);
    return HYKKT_OK;
  });
}

int hykkt_device_copy_f64(double* dst, const double* src, long long count)
{
  return guard([&] {
    HYKKT_REQUIRE(count >= 0, HYKKT_E_ARG, "negative count");
    if (count) {
      NEED(dst);
      NEED(src);
      hykkt::dev::copy(dst, src, count, hykkt::dev::Batch{}, nullptr);
    }
    return HYKKT_OK;
  });
}

int hykkt_memcpy_dtoh(void* dst, const void* src, unsigned long long bytes)
{
  return guard([&] {
    if (bytes) HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return HYKKT_OK;
  });
}

int hykkt_device_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes)
{
  return guard([&] {
    HYKKT_REQUIRE(free_bytes && total_bytes, HYKKT_E_ARG, "null output");
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return HYKKT_OK;
  });
}

int hykkt_device_synchronize(void)
{
  return guard([&] {
    HIPCHK(hipDeviceSynchronize());
    return HYKKT_OK;
  });
}

// ------------------------------------------------------------------ components
int hykkt_ruiz_scale(int its, int n, int m2, const int* h_rp, const int* h_ci, double* h_v,
                     int h_nnz, const int* jt_rp, const int* jt_ci, double* jt_v,
                     const int* j_rp, const int* j_ci, double* j_v, int j_nnz, double* rhs1,
                     double* rhs2, double* max_d)
{
  return guard([&] {
    HYKKT_REQUIRE(n >= 0 && m2 >= 0 && h_nnz >= 0 && j_nnz >= 0 && its >= 0, HYKKT_E_ARG,
                  "negative size");
    NEED(h_rp);
    NEED(jt_rp);
    NEED(j_rp);
    NEED(rhs1);
    NEED(rhs2);
    NEED(max_d);
    using hykkt::DevBuf;
    DevBuf<int> hrp, hci, jtrp, jtci, jrp, jci;
    DevBuf<double> hv, jtv, jv, r1, r2, md, sc;
    auto upi = [](DevBuf<int>& b, const int* p, int cnt) {
      std::vector<int> v(p, p + cnt);
      if (v.empty()) v.push_back(0);
      b.upload(v);
    };
    auto upd = [](DevBuf<double>& b, const double* p, int cnt) {
      std::vector<double> v(p, p + cnt);
      if (v.empty()) v.push_back(0.0);
      b.upload(v);
    };
    upi(hrp, h_rp, n + 1);
    upi(hci, h_ci, h_nnz);
    upd(hv, h_v, h_nnz);
    upi(jtrp, jt_rp, n + 1);
    upi(jtci, jt_ci, j_nnz);
    upd(jtv, jt_v, j_nnz);
    upi(jrp, j_rp, m2 + 1);
    upi(jci, j_ci, j_nnz);
    upd(jv, j_v, j_nnz);
    upd(r1, rhs1, n);
    upd(r2, rhs2, m2);
    md.alloc(std::max(1, n + m2));
    sc.alloc(std::max(1, n + m2));
    hipStream_t st = nullptr;
    const hykkt::dev::Batch one;
    if (its == 0) hykkt::dev::fill(md.p, 1.0, n + m2, one, st);
    for (int it = 0; it < its; it++) {
      hykkt::dev::ruiz_rowmax(n, n + m2, hrp.p, hv.p, jtrp.p, nullptr, jtv.p, jrp.p, jv.p, sc.p,
                              nullptr, one, st);
      hykkt::dev::ruiz_scale(n, n + m2, hrp.p, hci.p, hv.p, jtrp.p, jtci.p, jtv.p, jrp.p, jci.p,
                             jv.p, sc.p, r1.p, r2.p, md.p, it == 0, one, st);
    }
    HIPCHK(hipDeviceSynchronize());
    auto down = [](double* dst, const DevBuf<double>& b, int cnt) {
      if (cnt) HIPCHK(hipMemcpy(dst, b.p, cnt * sizeof(double), hipMemcpyDeviceToHost));
    };
    down(h_v, hv, h_nnz);
    down(jt_v, jtv, j_nnz);
    down(j_v, jv, j_nnz);
    down(rhs1, r1, n);
    down(rhs2, r2, m2);
    down(max_d, md, n + m2);
    return HYKKT_OK;
  });
}

int hykkt_ruiz_scale_noalloc(int its, int n, int m2, const int* h_rp, const int* h_ci,
                             double* h_v, int h_nnz, const int* jt_rp, const int* jt_ci,
                             double* jt_v, const int* j_rp, const int* j_ci, double* j_v,
                             int j_nnz, double* rhs1, double* rhs2, double* max_d)
{
  return guard([&] {
    HYKKT_REQUIRE(n >= 0 && m2 >= 0 && h_nnz >= 0 && j_nnz >= 0 && its >= 0, HYKKT_E_ARG,
                  "negative size");
    NEED(h_rp);
    NEED(jt_rp);
    NEED(j_rp);
    NEED(rhs1);
    NEED(rhs2);
    NEED(max_d);
    // every array the kernels touch, carved from the device scratch in
    // 256-byte slots: (host source or null, bytes)
    struct Part {
      const void* src;
      int64_t bytes;
      void* dev;
    };
    Part parts[] = {{h_rp, 4ll * (n + 1)},   {h_ci, 4ll * h_nnz},   {h_v, 8ll * h_nnz},
                    {jt_rp, 4ll * (n + 1)},  {jt_ci, 4ll * j_nnz},  {jt_v, 8ll * j_nnz},
                    {j_rp, 4ll * (m2 + 1)},  {j_ci, 4ll * j_nnz},   {j_v, 8ll * j_nnz},
                    {rhs1, 8ll * n},         {rhs2, 8ll * m2},      {nullptr, 8ll * (n + m2)},
                    {nullptr, 8ll * (n + m2)}};
    int64_t need = 0;
    for (Part& p : parts) need += (std::max<int64_t>(p.bytes, 1) + 255) / 256 * 256;
    const int64_t cap = static_cast<int64_t>(hykkt::dev::noalloc_scratch_bytes());
    HYKKT_REQUIRE(need <= cap, HYKKT_E_ARG,
                  "hykkt_ruiz_scale_noalloc: the system needs " + std::to_string(need) +
                      " bytes of device scratch; the scratch holds " + std::to_string(cap));
    char* base = static_cast<char*>(hykkt::dev::noalloc_scratch());
    int64_t at = 0;
    for (Part& p : parts) {
      p.dev = base + at;
      at += (std::max<int64_t>(p.bytes, 1) + 255) / 256 * 256;
      if (p.src && p.bytes > 0)
        HIPCHK(hipMemcpy(p.dev, p.src, static_cast<size_t>(p.bytes), hipMemcpyHostToDevice));
    }
    auto I = [&](int k) { return static_cast<int*>(parts[k].dev); };
    auto D = [&](int k) { return static_cast<double*>(parts[k].dev); };
    hipStream_t st = nullptr;
    const hykkt::dev::Batch one;
    if (its == 0) hykkt::dev::fill(D(11), 1.0, n + m2, one, st);
    for (int it = 0; it < its; it++) {
      hykkt::dev::ruiz_rowmax(n, n + m2, I(0), D(2), I(3), nullptr, D(5), I(6), D(8), D(12),
                              nullptr, one, st);
      hykkt::dev::ruiz_scale(n, n + m2, I(0), I(1), D(2), I(3), I(4), D(5), I(6), I(7), D(8),
                             D(12), D(9), D(10), D(11), it == 0, one, st);
    }
    HIPCHK(hipStreamSynchronize(st));
    auto down = [](double* dst, const double* src, int64_t cnt) {
      if (cnt > 0)
        HIPCHK(hipMemcpy(dst, src, static_cast<size_t>(cnt) * sizeof(double),
                         hipMemcpyDeviceToHost));
    };
    down(h_v, D(2), h_nnz);
    down(jt_v, D(5), j_nnz);
    down(j_v, D(8), j_nnz);
    down(rhs1, D(9), n);
    down(rhs2, D(10), m2);
    down(max_d, D(11), n + m2);
    return HYKKT_OK;
  });
}

int hykkt_perm_reverse(int n, const int* perm, int* rev)
{
  return guard([&] {
    NEED(perm);
    NEED(rev);
    std::vector<int> p(perm, perm + n), r;
    hykkt::perm_reverse(p, r);
    std::memcpy(rev, r.data(), sizeof(int) * n);
    return HYKKT_OK;
  });
}

static hykkt::Pattern make_pattern(int n, int m, const int* rows, const int* cols)
{
  hykkt::Pattern a;
  a.n = n;
  a.m = m;
  a.rp.assign(rows, rows + n + 1);
  a.ci.assign(cols, cols + rows[n]);
  return a;
}

int hykkt_perm_map_rc(int n, const int* rows, const int* cols, const int* perm, const int* rev,
                      int* prows, int* pcols, int* pmap)
{
  return guard([&] {
    NEED(rows);
    NEED(perm);
    NEED(rev);
    hykkt::Pattern a = make_pattern(n, n, rows, cols), pa;
    std::vector<int> p(perm, perm + n), r(rev, rev + n), map;
    hykkt::perm_map_rc(a, p, r, pa, map);
    std::memcpy(prows, pa.rp.data(), sizeof(int) * (n + 1));
    std::memcpy(pcols, pa.ci.data(), sizeof(int) * pa.ci.size());
    std::memcpy(pmap, map.data(), sizeof(int) * map.size());
    return HYKKT_OK;
  });
}

int hykkt_perm_map_c(int n, const int* rows, const int* cols, const int* rev, int* pcols, int* pmap)
{
  return guard([&] {
    NEED(rows);
    NEED(rev);
    int mcols = 0;
    for (int q = 0; q < rows[n]; q++) mcols = std::max(mcols, cols[q] + 1);
    hykkt::Pattern a = make_pattern(n, mcols, rows, cols), pa;
    std::vector<int> r(rev, rev + mcols), map;
    hykkt::perm_map_c(a, r, pa, map);
    std::memcpy(pcols, pa.ci.data(), sizeof(int) * pa.ci.size());
    std::memcpy(pmap, map.data(), sizeof(int) * map.size());
    return HYKKT_OK;
  });
}

int hykkt_perm_map_r(int n, const int* rows, const int* cols, const int* perm, int* prows,
                     int* pcols, int* pmap)
{
  return guard([&] {
    NEED(rows);
    NEED(perm);
    hykkt::Pattern a = make_pattern(n, 0, rows, cols), pa;
    std::vector<int> p(perm, perm + n), map;
    hykkt::perm_map_r(a, p, pa, map);
    std::memcpy(prows, pa.rp.data(), sizeof(int) * (n + 1));
    std::memcpy(pcols, pa.ci.data(), sizeof(int) * pa.ci.size());
    std::memcpy(pmap, map.data(), sizeof(int) * map.size());
    return HYKKT_OK;
  });
}

int hykkt_amd_order(int n, const int* rp, const int* ci, int* perm)
{
  return guard([&] {
    NEED(rp);
    NEED(perm);
    hykkt::Pattern a = make_pattern(n, n, rp, ci);
    std::vector<int> p = hykkt::amd_order(a);
    std::memcpy(perm, p.data(), sizeof(int) * n);
    return HYKKT_OK;
  });
}

int hykkt_chol_analyze_host(int n, const int* rp, const int* ci, long long* nnz_l,
                            double* flops, int* ns, int* nlev, int* max_front)
{
  return guard([&] {
    NEED(rp);
    hykkt::Pattern a = make_pattern(n, n, rp, ci);
    hykkt::CholPlan p = hykkt::chol_analyze(a, hykkt::CholOptions{});
    if (nnz_l) *nnz_l = p.nnz_l_true;
    if (flops) *flops = p.flops;
    if (ns) *ns = p.ns;
    if (nlev) *nlev = static_cast<int>(p.level_ptr.size()) - 1;
    if (max_front) *max_front = p.max_front;
    return HYKKT_OK;
  });
}

int hykkt_chol_create(int n, int nnz, const int* rp, const int* ci, hykkt_chol* out)
{
  return guard([&] {
    NEED(rp);
    NEED(out);
    HYKKT_REQUIRE(rp[n] == nnz, HYKKT_E_ARG, "rowptr[n] != nnz");
    auto c = std::make_unique<hykkt_chol_s>();
    HIPCHK(hipStreamCreate(&c->st));
    hykkt::Pattern a = make_pattern(n, n, rp, ci);
    c->eng = std::make_unique<hykkt::CholEngine>(a, hykkt::CholOptions{}, c->st);
    c->eng->own(1);
    c->vals.alloc(std::max(1, nnz));
    c->x.alloc(std::max(1, n));
    c->post = c->eng->plan().post;
    c->n = n;
    c->nnz = nnz;
    *out = c.release();
    return HYKKT_OK;
  });
}

int hykkt_chol_factor(hykkt_chol c, const double* vals, double tol, int* singularity)
{
  return guard([&] {
    NEED(c);
    NEED(vals);
    if (c->nnz)
      HIPCHK(hipMemcpyAsync(c->vals.p, vals, c->nnz * sizeof(double), hipMemcpyHostToDevice, c->st));
    const hykkt::dev::Batch one{1, c->eng->stride(), nullptr};
    c->eng->assemble(c->vals.p, one);
    c->eng->factor(tol, one);
    int piv = c->eng->pivot();
    if (singularity) *singularity = piv;
    return piv >= 0 ? HYKKT_E_PIVOT : HYKKT_OK;
  });
}

int hykkt_chol_solve(hykkt_chol c, const double* b, double* x)
{
  return guard([&] {
    NEED(c);
    NEED(b);
    NEED(x);
    const int n = c->n;
    std::vector<double> t(n);
    for (int k = 0; k < n; k++) t[k] = b[c->post[k]];
    HIPCHK(hipMemcpyAsync(c->x.p, t.data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
    c->eng->solve(c->x.p);
    HIPCHK(hipMemcpyAsync(t.data(), c->x.p, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    HYKKT_REQUIRE(!c->eng->solve_error(), HYKKT_E_STATE,
                  "triangular-solve sweep exceeded its wait bound (results invalid)");
    for (int k = 0; k < n; k++) x[c->post[k]] = t[k];
    return HYKKT_OK;
  });
}

int hykkt_chol_info(hykkt_chol c, long long* nnz_l, int* ns, int* nlev, double* flops)
{
  return guard([&] {
    NEED(c);
    const hykkt::CholPlan& p = c->eng->plan();
    if (nnz_l) *nnz_l = p.nnz_l_true;
    if (ns) *ns = p.ns;
    if (nlev) *nlev = static_cast<int>(p.level_ptr.size()) - 1;
    if (flops) *flops = p.flops;
    return HYKKT_OK;
  });
}

int hykkt_chol_destroy(hykkt_chol c)
{
  return guard([&] {
    delete c;
    return HYKKT_OK;
  });
}

int hykkt_schur_cg(hykkt_chol c, int m, int n, const int* j_rp, const int* j_ci,
                   const double* j_v, double* x, const double* b, double tol, int itmax,
                   int* iters, double* final_norm)
{
  return guard([&] {
    NEED(c);
    NEED(j_rp);
    NEED(x);
    NEED(b);
    HYKKT_REQUIRE(n == c->n, HYKKT_E_ARG, "J must have as many columns as the factor");
    using hykkt::DevBuf;
    hykkt::Pattern J = make_pattern(m, n, j_rp, j_ci), Jh;
    std::vector<int> ipos(n), jmap;
    for (int k = 0; k < n; k++) ipos[c->post[k]] = k;
    hykkt::perm_map_c(J, ipos, Jh, jmap);
    hykkt::TransposeMap jt = hykkt::transpose_map(Jh);
    std::vector<double> jhv(jmap.size()), jthv(jt.map.size());
    for (size_t q = 0; q < jmap.size(); q++) jhv[q] = j_v[jmap[q]];
    for (size_t q = 0; q < jt.map.size(); q++) jthv[q] = jhv[jt.map[q]];
    // row blocks of the LDS-staged SpMVs: whole rows, at most 256 of them and
    // at most spmv_lds_entries() entries per block (a longer row alone)
    auto blocks = [](const std::vector<int>& rp, int rows) {
      const int E = hykkt::dev::spmv_lds_entries();
      std::vector<int> bl{0};
      for (int r = 0; r < rows;) {
        int r1 = r + 1;
        if (rp[r + 1] - rp[r] <= E)
          while (r1 < rows && r1 - r < 256 && rp[r1 + 1] - rp[r] <= E) r1++;
        bl.push_back(r1);
        r = r1;
      }
      return bl;
    };
    const std::vector<int> jb = blocks(Jh.rp, m), jtb = blocks(jt.at.rp, n);
    DevBuf<int> jrp, jci, jtrp, jtci, jblk, jtblk, flg;
    DevBuf<double> jv, jtv, t1, xv, r, w, p, s, bb, sc, part, gath;
    auto upi = [](DevBuf<int>& d, const std::vector<int>& v) {
      if (v.empty()) d.upload(std::vector<int>{0});
      else d.upload(v);
    };
    auto upd = [](DevBuf<double>& d, const std::vector<double>& v) {
      if (v.empty()) d.upload(std::vector<double>{0.0});
      else d.upload(v);
    };
    upi(jrp, Jh.rp);
    upi(jci, Jh.ci);
    upd(jv, jhv);
    upi(jtrp, jt.at.rp);
    upi(jtci, jt.at.ci);
    upd(jtv, jthv);
    upi(jblk, jb);
    upi(jtblk, jtb);
    upd(xv, std::vector<double>(x, x + m));
    upd(bb, std::vector<double>(b, b + m));
    t1.alloc(std::max(1, n));
    for (DevBuf<double>* q : {&r, &w, &p, &s}) q->alloc(std::max(1, m));
    sc.alloc(8);
    part.alloc(hykkt::schur_partials(m, static_cast<int>(jb.size()) - 1));
    gath.alloc(4);
    flg.alloc(3);
    hykkt::SchurOps o;
    o.m = m;
    o.n = n;
    o.jh_rp = jrp.p;
    o.jh_ci = jci.p;
    o.jh_blk = jblk.p;
    o.jh_nblk = static_cast<int>(jb.size()) - 1;
    o.jh_v = jv.p;
    o.jth_rp = jtrp.p;
    o.jth_ci = jtci.p;
    o.jth_blk = jtblk.p;
    o.jth_nblk = static_cast<int>(jtb.size()) - 1;
    o.jth_v = jtv.p;
    o.chol = c->eng.get();
    o.st = c->st;
    o.t1 = t1.p;
    o.x = xv.p;
    o.r = r.p;
    o.w = w.p;
    o.p = p.p;
    o.s = s.p;
    o.sc = sc.p;
    o.part = part.p;
    o.live = flg.p;
    o.iters = flg.p + 1;
    o.status = flg.p + 2;
    o.gath = gath.p;
    // one system: its vectors are separate buffers, the factor the engine's
    const hykkt::dev::Batch one{1, c->eng->stride(), nullptr};
    const std::vector<hykkt::CgResult> res =
        hykkt::schur_cg_run(o, bb.p, tol, itmax, one, false, 1, 1);
    HIPCHK(hipMemcpyAsync(x, xv.p, m * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    HYKKT_REQUIRE(!c->eng->solve_error(), HYKKT_E_STATE,
                  "triangular-solve sweep exceeded its wait bound (results invalid)");
    if (iters) *iters = res[0].iters;
    if (final_norm) *final_norm = res[0].final_norm;
    return res[0].status;
  });
}

}  // extern "C"
