// solver_emulator.cpp -- the reference's interior-point sequence harness
// (examples/cuSolver_driver_solver_emulator.cpp:50-101) on this library.
// Arguments: base_dir base_name seq_list skip_lines gamma, where seq_list is
// comma-separated ("09,10"); for each entry the eight files
//   block_{H,Ds,Jc,Jd}_matrix_<base_name>_<seq>.mtx
//   block_{rx,rs,ry,rd}_rhs_<base_name>_<seq>.mtx
// are read with read_matrix_files and solved by one solver instance, so a
// same-pattern sequence reuses the symbolic analysis.  Exit status 1 as soon
// as a residual test fails.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/HykktSolver.hpp"
#include "solution_dump.hpp"

namespace {

std::vector<std::string> split(const std::string& s, char delim)
{
  std::vector<std::string> out;
  size_t start = 0;
  for (;;) {
    size_t p = s.find(delim, start);
    out.push_back(s.substr(start, p == std::string::npos ? std::string::npos : p - start));
    if (p == std::string::npos) break;
    start = p + 1;
  }
  return out;
}

std::string file(const std::string& dir, const char* kind, const std::string& base,
                 const std::string& seq)
{
  return dir + "/" + kind + "_" + base + "_" + seq + ".mtx";
}

}  // namespace

int main(int argc, char* argv[])
{
  if (argc != 6) {
    std::printf("Incorrect number of inputs. Exiting ...\n");
    return -1;
  }
  std::printf("Starting ...\n");
  const std::string dir = argv[1], base = argv[2];
  const std::vector<std::string> seqs = split(argv[3], ',');
  const int skip = std::atoi(argv[4]);
  const double gamma = std::atof(argv[5]);
  HykktSolver hs(gamma);
  for (const std::string& q : seqs) {
    std::printf("Working on Sequence: %s\n", q.c_str());
    const std::string f[8] = {
        file(dir, "block_H_matrix", base, q),  file(dir, "block_Ds_matrix", base, q),
        file(dir, "block_Jc_matrix", base, q), file(dir, "block_Jd_matrix", base, q),
        file(dir, "block_rx_rhs", base, q),    file(dir, "block_rs_rhs", base, q),
        file(dir, "block_ry_rhs", base, q),    file(dir, "block_rd_rhs", base, q)};
    hs.read_matrix_files(f[0].c_str(), f[1].c_str(), f[2].c_str(), f[3].c_str(), f[4].c_str(),
                         f[5].c_str(), f[6].c_str(), f[7].c_str(), skip);
    const int status = hs.execute();
    const hykkt_stats st = hs.stats();
    std::printf("sequence %s: status %d resid2 %.3e resid4 %.3e cg_iters %d refine %d "
                "solve %.2f ms symbolic %.1f ms\n",
                q.c_str(), status, st.resid2, st.resid4, st.cg_iters, st.refine_steps,
                st.t_solve_ms, st.t_symbolic_ms);
    if (!dump_solution(hs, "sequence_" + q)) return EXIT_FAILURE;
    if (status == 1) return 1;
  }
  return 0;
}
