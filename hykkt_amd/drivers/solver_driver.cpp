// solver_driver.cpp -- the reference's API-path driver
// (examples/cuSolver_driver_solver.cpp:12-138) on this library: the same 18
// arguments (8 files of system 1, 8 files of system 2, skip_lines, gamma).
// System 1 is bound through the HykktSolver setters and solved; system 2 is
// then copied into the same host buffers (same sparsity pattern) and solved
// again without re-binding.  Exit status: 0 if both residual tests pass.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/HykktSolver.hpp"
#include "solution_dump.hpp"

namespace {

struct Csr {
  int n = 0, m = 0, nnz = 0;
  std::vector<int> rp, ci;
  std::vector<double> v;
};

Csr read_csr(const char* path, int skip, int symmetric)
{
  Csr c;
  if (hykkt_mm_read_csr(path, skip, symmetric, &c.n, &c.m, &c.nnz, nullptr, nullptr, nullptr)) {
    std::fprintf(stderr, "%s\n", hykkt_last_error());
    std::exit(EXIT_FAILURE);
  }
  c.rp.resize(c.n + 1);
  c.ci.resize(c.nnz > 0 ? c.nnz : 1);
  c.v.resize(c.nnz > 0 ? c.nnz : 1);
  hykkt_mm_read_csr(path, skip, symmetric, &c.n, &c.m, &c.nnz, c.rp.data(), c.ci.data(),
                    c.v.data());
  return c;
}

std::vector<double> read_vec(const char* path)
{
  int n = 0;
  if (hykkt_mm_read_rhs(path, &n, nullptr)) {
    std::fprintf(stderr, "%s\n", hykkt_last_error());
    std::exit(EXIT_FAILURE);
  }
  std::vector<double> v(n > 0 ? n : 1);
  hykkt_mm_read_rhs(path, &n, v.data());
  v.resize(n);
  return v;
}

struct System {
  Csr H, Ds, Jc, Jd;
  std::vector<double> rx, rs, ry, ryd;
};

System load(char** f, int skip)
{
  System s;
  s.H = read_csr(f[0], skip, 1);   // sym_coo_to_csr
  s.Ds = read_csr(f[1], skip, 0);
  s.Jc = read_csr(f[2], skip, 0);
  s.Jd = read_csr(f[3], skip, 0);
  s.rx = read_vec(f[4]);
  s.rs = read_vec(f[5]);
  s.ry = read_vec(f[6]);
  s.ryd = read_vec(f[7]);
  return s;
}

// copy_mmatrix / copy_vector of the reference driver: same sizes, overwrite
void overwrite(Csr& dst, const Csr& src)
{
  if (dst.n != src.n || dst.nnz != src.nnz) {
    std::fprintf(stderr, "system 2 does not share system 1's pattern sizes\n");
    std::exit(EXIT_FAILURE);
  }
  dst.rp = src.rp;   // same lengths: the vectors keep their storage
  dst.ci = src.ci;
  dst.v = src.v;
}

}  // namespace

int main(int argc, char* argv[])
{
  if (argc != 19) {
    std::printf("Incorrect number of inputs. Exiting ...\n");
    return -1;
  }
  std::printf("Starting ...\n");
  const int skip = std::atoi(argv[17]);
  const double gamma = std::atof(argv[18]);
  System a = load(argv + 1, skip);
  System b = load(argv + 9, skip);
  std::vector<double> x(a.H.n, 0.0), s(a.Ds.n, 0.0), y(a.Jc.n, 0.0), yd(a.Jd.n, 0.0);

  HykktSolver hs(gamma);
  hs.set_H(a.H.rp.data(), a.H.ci.data(), a.H.v.data(), a.H.n, a.H.m, a.H.nnz);
  hs.set_Ds(a.Ds.rp.data(), a.Ds.ci.data(), a.Ds.v.data(), a.Ds.n, a.Ds.m, a.Ds.nnz);
  hs.set_Jc(a.Jc.rp.data(), a.Jc.ci.data(), a.Jc.v.data(), a.Jc.n, a.Jc.m, a.Jc.nnz);
  hs.set_Jd(a.Jd.rp.data(), a.Jd.ci.data(), a.Jd.v.data(), a.Jd.n, a.Jd.m, a.Jd.nnz);
  hs.set_rx(a.rx.data(), a.H.n);
  hs.set_rs(a.rs.data(), a.Ds.n);
  hs.set_ry(a.ry.data(), a.Jc.n);
  hs.set_ryd(a.ryd.data(), a.Jd.n);
  hs.set_x_host(x.data(), a.H.n);
  hs.set_s_host(s.data(), a.Ds.n);
  hs.set_y_host(y.data(), a.Jc.n);
  hs.set_yd_host(yd.data(), a.Jd.n);
  int status = hs.execute();
  hykkt_stats st = hs.stats();
  std::printf("system 1: status %d resid2 %.3e resid4 %.3e cg_iters %d\n", status, st.resid2,
              st.resid4, st.cg_iters);
  if (!dump_solution(hs, "system1")) return EXIT_FAILURE;
  if (status == 1) return 1;

  overwrite(a.H, b.H);
  overwrite(a.Ds, b.Ds);
  overwrite(a.Jc, b.Jc);
  overwrite(a.Jd, b.Jd);
  a.rx = b.rx;
  a.rs = b.rs;
  a.ry = b.ry;
  a.ryd = b.ryd;
  status = hs.execute();
  st = hs.stats();
  std::printf("system 2: status %d resid2 %.3e resid4 %.3e cg_iters %d\n", status, st.resid2,
              st.resid4, st.cg_iters);
  if (!dump_solution(hs, "system2")) return EXIT_FAILURE;
  return status;
}
