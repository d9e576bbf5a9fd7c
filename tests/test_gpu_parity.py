"""Parity of the HIP path (through the C ABI) with the CPU oracle and with the
reference drivers' known answers.  Needs an MI355X.

Tolerances: integer/index results and the IEEE-exact formation steps
(H~ and H_gamma SpGEMM values, Ruiz scaling, max_d) are compared bit for bit;
solutions are compared by the relative KKT residual (<= 1e-8 over all four
block rows) and by their relative difference from the oracle's solution, which
is bounded by a multiple of the system's conditioning (stated per test).
"""
import numpy as np
import pytest

import _oracle as O
import fixtures as F
import hykkt_amd as H
from hykkt_amd.generator import GridPattern, workload

pytestmark = pytest.mark.gpu


def to_csr(b):
    return O.Csr(b.n, b.m, b.rp, b.ci, b.v)


def bind(s, H_, Dsv, Jc, Jd, rx, rs, ry, ryd):
    d = len(Dsv)
    s.set_H(H_.rp, H_.ci, H_.v, H_.n, H_.m, H_.nnz)
    s.set_Ds(np.arange(d + 1), np.arange(d), Dsv, d, d, d)
    s.set_Jc(Jc.rp, Jc.ci, Jc.v, Jc.n, Jc.m, Jc.nnz)
    s.set_Jd(Jd.rp, Jd.ci, Jd.v, Jd.n, Jd.m, Jd.nnz)
    s.set_rx(rx, len(rx))
    s.set_rs(rs, len(rs))
    s.set_ry(ry, len(ry))
    s.set_ryd(ryd, len(ryd))


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def solve_both(H_, Dsv, Jc, Jd, rx, rs, ry, ryd, gamma):
    s = H.HykktSolver(gamma)
    bind(s, H_, Dsv, Jc, Jd, rx, rs, ry, ryd)
    status = s.execute()
    st = s.stats()
    mine = s.solution()
    ref = O.execute(H_, Dsv, Jc, Jd, rx, rs, ry, ryd, gamma, perm=s.permutation(),
                    intermediates=True)
    return s, status, st, mine, ref


# ------------------------------------------------------------------ components
def test_ruiz_kat_gpu():
    """examples/Ruiz_driver.cpp:145-182 on the GPU kernels, and bit-exact
    against the oracle restatement of matrix_vector_ops.cu:146-266."""
    n = 1024
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(2 * n - 1, np.int32)
    v = np.zeros(2 * n - 1)
    for i in range(n):
        if i:
            v[2 * i - 1], ci[2 * i - 1], rp[i] = i + 1, i - 1, 2 * i - 1
        ci[2 * i] = i
    rp[n] = 2 * n - 1
    A = O.Csr(n, n, rp, ci, v)
    A.v[A.ci == np.repeat(np.arange(n), np.diff(A.rp))] += 1.0
    At = O.transpose(A)
    Hh = O.Csr(n, n, np.arange(n + 1), np.arange(n), np.full(n, np.sqrt(n)))
    hv, tv, jv, r1, r2, md = H.ruiz_scale(2, (Hh.rp, Hh.ci, Hh.v), (At.rp, At.ci, At.v),
                                          (A.rp, A.ci, A.v), np.ones(n), np.ones(n))
    tol = 1e-8
    assert abs(hv[n // 2 - 1] - 0.062378167641326) < tol
    assert abs(jv[-1] - 0.005524271728020) < tol
    assert abs(tv[1] - 0.5) < tol
    assert abs(r1[n // 2 - 1] - 0.044151078568835) < tol
    assert abs(r2[n // 2 - 1] - 0.044194173824159) < tol
    assert abs(md[32] - 0.171498585142) < tol
    Ho, Ato, Ao, o1, o2, omd = O.ruiz(2, Hh, At, A, np.ones(n), np.ones(n))
    for a, b in ((hv, Ho.v), (tv, Ato.v), (jv, Ao.v), (r1, o1), (r2, o2), (md, omd)):
        assert np.array_equal(a, b)


def test_cholesky_kat_gpu(mats):
    """examples/cuSolver_driver_cholesky.cpp:32-75 (abs 3e-2), plus agreement
    with the oracle factor's solution to 1e-9 (cond(H) ~ 1e7)."""
    A = O.read_mm(mats("a200_chol/H_matrix_ACTIVSg200_AC_00.mtx"), 3, False)
    b = O.read_rhs(mats("a200_chol/rhs_chol_ACTIVSg200_AC_00.mtx"))
    ch = H.Cholesky(A.rp, A.ci)
    assert ch.factor(A.v, 1e-12) == -1
    x = ch.solve(b)
    assert abs(x[0] + 97.663191790329750) < 3e-2
    assert abs(x[6] + 1.389053208413921) < 3e-2
    oc = O.Chol(A)
    oc.factor()
    assert rel(x, oc.solve(b)) < 1e-9


def test_schur_cg_kat_gpu(mats):
    """examples/cuSolver_driver_schur_cg.cpp:130-154 (x0 = 0, abs 1e-6)."""
    J = O.read_mm(mats("a200_schur/JC_matrix_ACTIVSg200_AC_00.mtx"), 3, False)
    A = O.read_mm(mats("a200_schur/H_matrix_ACTIVSg200_AC_00.mtx"), 3, False)
    b = O.read_rhs(mats("a200_schur/CG_rhs_ACTIVSg200_AC_00.mtx"))
    ch = H.Cholesky(A.rp, A.ci)
    assert ch.factor(A.v) == -1
    x, rc, it, fn = ch.schur_cg(J.rp, J.ci, J.v, b)
    assert rc == 0 and fn < 1e-12
    assert abs(x[0] - 22.171865776354700) < 1e-6
    assert abs(x[6] + 4446.628667344612) < 1e-6


def dense_to_csr(A, lower_only=False):
    rp = [0]
    ci, v = [], []
    for i in range(A.shape[0]):
        nz = np.nonzero(A[i, : i + 1] if lower_only else A[i])[0]
        ci += nz.tolist()
        v += A[i, nz].tolist()
        rp.append(len(ci))
    return np.array(rp), np.array(ci), np.array(v)


@pytest.mark.parametrize("n", [1, 7, 128, 129, 300, 520])
def test_cholesky_dense_spd(n):
    """One dense supernode: 128-column steps with a partial last block, MFMA
    trailing updates and inverted diagonal blocks; solution vs numpy to 1e-11
    (cond <= ~1e3)."""
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T / n + np.eye(n)
    rp, ci, v = dense_to_csr(A)
    ch = H.Cholesky(rp, ci)
    assert ch.factor(v) == -1
    b = rng.standard_normal(n)
    x = ch.solve(b)
    assert rel(x, np.linalg.solve(A, b)) < 1e-11


def test_cholesky_block_arrow():
    """Three dense 200-blocks coupled through a dense 150-row border: leaves
    with update matrices of 150 rows (two MFMA tiles), extend-add into the
    border supernode, forward pulls and backward gathers across supernodes."""
    rng = np.random.default_rng(3)
    nb, nbord = 200, 150
    n = 3 * nb + nbord
    A = np.zeros((n, n))
    for k in range(3):
        Bk = rng.standard_normal((nb, nb))
        A[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb] = Bk @ Bk.T / nb + 4 * np.eye(nb)
        C = rng.standard_normal((nbord, nb)) * 0.05
        A[3 * nb:, k * nb:(k + 1) * nb] = C
        A[k * nb:(k + 1) * nb, 3 * nb:] = C.T
    Bb = rng.standard_normal((nbord, nbord))
    A[3 * nb:, 3 * nb:] = Bb @ Bb.T / nbord + 8 * np.eye(nbord)
    rp, ci, v = dense_to_csr(A)
    ch = H.Cholesky(rp, ci)
    assert ch.factor(v) == -1
    info = ch.info()
    assert info["n_supernodes"] >= 3
    for seed in range(3):
        b = np.random.default_rng(seed).standard_normal(n)
        assert rel(ch.solve(b), np.linalg.solve(A, b)) < 1e-11


def test_cholesky_pivot_failure_reported():
    """An indefinite matrix yields a pivot index, not a crash or an exit
    (CholeskyClass.cpp:83-89 only prints)."""
    n = 50
    A = np.diag(np.full(n, 2.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    A[20, 20] = -5.0
    rp = [0]
    ci, v = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci += nz.tolist()
        v += A[i, nz].tolist()
        rp.append(len(ci))
    ch = H.Cholesky(rp, ci)
    piv = ch.factor(v)
    assert piv >= 0


# ------------------------------------------------------------------ end to end
@pytest.mark.parametrize("kind,step", [(F.A200, "09"), (F.A200, "10"), (F.A200, "00"),
                                       (F.A2000, "09")])
def test_execute_fixture_vs_oracle(mats, kind, step):
    """CTest hybrid/solver systems through the API path.  Formation is bit-exact
    with the oracle; the solution reaches resid4 <= 1e-8 and agrees with the
    oracle's (same ordering) to 1e-6 relative."""
    s = F.load(mats, kind, step)
    solver, status, st, mine, ref = solve_both(s["H"], s["Ds"].v, s["Jc"], s["Jd"], s["rx"],
                                               s["rs"], s["ry"], s["ryd"], 1e4)
    assert status == 0 and st["pivot"] == -1 and st["cg_status"] == 0
    assert st["resid4"] <= 1e-8 and st["resid2"] <= 1e-8
    assert st["nnz_htil"] == ref["stats"]["nnz_htil"]
    assert st["nnz_hgam"] == ref["stats"]["nnz_hgam"]
    # bit-exact formation: H~ (before Ruiz), max_d, H_gamma (after Ruiz)
    assert np.array_equal(solver.debug(11, st["nnz_htil"]), ref["htil_v"])
    assert np.array_equal(solver.debug(4, st["n"] + st["m"]), ref["max_d"])
    # H_gamma as the factorization received it, read back from the factor storage
    got = solver.debug(10, st["nnz_hgam"])
    want = F.hgamma_in_factor_order(solver, st["nnz_hgam"], st["n"], ref["hgam_v"])
    assert F.same_entries(got, want)
    assert np.array_equal(solver.debug(12, st["n"]), ref["rx_hat"])
    x = np.concatenate(mine)
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert rel(x, xo) < 1e-6
    r2, r4 = O.kkt_residual(s["H"], s["Ds"].v, s["Jc"], s["Jd"], s["rx"], s["rs"], s["ry"],
                            s["ryd"], *mine)
    assert r4 <= 1e-8 and abs(r4 - st["resid4"]) <= 0.05 * r4 + 1e-14


def test_solver_value_reuse_sequence(mats):
    """examples/cuSolver_driver_solver.cpp:96-138: bind system 09, execute,
    overwrite the bound host buffers with system 10 in place, execute again."""
    a = F.load(mats, F.A200, "09")
    b = F.load(mats, F.A200, "10")
    s = H.HykktSolver(1e4)
    bind(s, a["H"], a["Ds"].v, a["Jc"], a["Jd"], a["rx"], a["rs"], a["ry"], a["ryd"])
    assert s.execute() == 0
    t_sym = s.stats()["t_symbolic_ms"]
    kept = s._keep
    for key, src in (("H", b["H"]), ("Jc", b["Jc"]), ("Jd", b["Jd"])):
        kept[key][2][:] = src.v
    kept["Ds"][2][:] = b["Ds"].v
    for key in ("rx", "rs", "ry", "ryd"):
        kept[key][:] = b[key]
    assert s.execute() == 0
    st = s.stats()
    assert st["n_solves"] == 2 and st["t_symbolic_ms"] == t_sym   # no re-analysis
    ref = O.execute(b["H"], b["Ds"].v, b["Jc"], b["Jd"], b["rx"], b["rs"], b["ry"], b["ryd"],
                    1e4, perm=s.permutation())
    x = np.concatenate(s.solution())
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert rel(x, xo) < 1e-6 and st["resid4"] <= 1e-8


def test_read_matrix_files_path(mats):
    """HykktSolver::read_matrix_files (HykktSolver.cpp:113-161) with the CTest
    hybrid1 arguments (a200 step 10, skip 2, gamma 1e4)."""
    fl = F.files(F.A200, "10")
    s = H.HykktSolver(1e4)
    s.read_matrix_files(*[mats(fl[k]) for k in ("H", "Ds", "Jc", "Jd", "rx", "rs", "ry", "ryd")],
                        fl["skip"])
    assert s.execute() == 0
    assert s.stats()["resid4"] <= 1e-8
    # the solution of the file path against the oracle's solve of the same files
    sy = F.load(mats, F.A200, "10")
    ref = O.execute(sy["H"], sy["Ds"].v, sy["Jc"], sy["Jd"], sy["rx"], sy["rs"], sy["ry"],
                    sy["ryd"], 1e4, perm=s.permutation())
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert rel(np.concatenate(s.solution()), xo) < 1e-6
    # a different Jd emptiness on the file path is refused (HykktSolver.cpp:140-144,243-250)
    fl0 = F.files(F.A200, "00")
    s.read_matrix_files(*[mats(fl0[k]) for k in ("H", "Ds", "Jc", "Jd", "rx", "rs", "ry", "ryd")],
                        fl0["skip"])
    assert s.execute() == 1


def test_generated_grid500_vs_oracle_and_dense():
    """The 500-bus workload: resid4 <= 1e-8, and the solution vs the oracle's."""
    g = workload("grid500")
    sy = g.system()
    Hc, Jc, Jd = to_csr(sy.H), to_csr(sy.Jc), to_csr(sy.Jd)
    solver, status, st, mine, ref = solve_both(Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd,
                                               sy.gamma)
    assert status == 0 and st["resid4"] <= 1e-8
    x = np.concatenate(mine)
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert rel(x, xo) < 1e-6


def test_small_system_vs_dense_lu():
    """Independent of the oracle: dense LU of the full 4x4 block KKT."""
    from test_oracle import dense_kkt
    g = GridPattern(25, 33, seed=11, gamma=1e4)
    sy = g.system()
    Hc, Jc, Jd = to_csr(sy.H), to_csr(sy.Jc), to_csr(sy.Jd)
    s = H.HykktSolver(sy.gamma)
    bind(s, Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd)
    assert s.execute() == 0
    K = dense_kkt(Hc, sy.Ds.v, Jc, Jd)
    rhs = np.concatenate([sy.rx, sy.rs, sy.ry, sy.ryd])
    sol = np.linalg.solve(K, rhs)
    assert rel(np.concatenate(s.solution()), sol) < 1e-8


def test_deterministic_repeat():
    """Two solves of the same system give bit-identical solutions."""
    g = GridPattern(120, 160, seed=5)
    sy = g.system()
    Hc, Jc, Jd = to_csr(sy.H), to_csr(sy.Jc), to_csr(sy.Jd)
    s = H.HykktSolver(sy.gamma)
    bind(s, Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd)
    assert s.execute() == 0
    a = np.concatenate(s.solution())
    assert s.execute() == 0
    b = np.concatenate(s.solution())
    assert np.array_equal(a, b)


def test_device_resident_inputs():
    """inputs_on_device: the same system from device-resident buffers gives the
    same bits as from host arrays."""
    g = GridPattern(80, 104, seed=2)
    sy = g.system()
    Hc, Jc, Jd = to_csr(sy.H), to_csr(sy.Jc), to_csr(sy.Jd)
    ref = H.HykktSolver(sy.gamma)
    bind(ref, Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd)
    assert ref.execute() == 0
    D = H.DeviceArray
    i32 = lambda a: D(np.asarray(a, np.int32))
    f64 = lambda a: D(np.asarray(a, np.float64))
    s = H.HykktSolver(sy.gamma)
    s.options(inputs_on_device=1)
    d = len(sy.Ds.v)
    s.set_H(i32(Hc.rp), i32(Hc.ci), f64(Hc.v), Hc.n, Hc.m, Hc.nnz)
    s.set_Ds(i32(np.arange(d + 1)), i32(np.arange(d)), f64(sy.Ds.v), d, d, d)
    s.set_Jc(i32(Jc.rp), i32(Jc.ci), f64(Jc.v), Jc.n, Jc.m, Jc.nnz)
    s.set_Jd(i32(Jd.rp), i32(Jd.ci), f64(Jd.v), Jd.n, Jd.m, Jd.nnz)
    for name, v in (("rx", sy.rx), ("rs", sy.rs), ("ry", sy.ry), ("ryd", sy.ryd)):
        getattr(s, "set_" + name)(f64(v), len(v))
    assert s.execute() == 0
    assert np.array_equal(np.concatenate(s.solution()), np.concatenate(ref.solution()))


def test_indefinite_hgamma_returns_failure():
    """gamma = 0 with an indefinite H: non-positive pivot -> execute() == 1 and
    the failing pivot is reported (no exit, no NaN-propagating crash)."""
    g = GridPattern(30, 40, seed=9, gamma=0.0)
    sy = g.system()
    Hc, Jc, Jd = to_csr(sy.H), to_csr(sy.Jc), to_csr(sy.Jd)
    Hc.v = -np.abs(Hc.v)
    s = H.HykktSolver(0.0)
    bind(s, Hc, sy.Ds.v * 0 + 1e-8, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd)
    assert s.execute() == 1
    assert s.stats()["pivot"] >= 0


def test_iterative_refinement_gpu():
    """Ds spread 1e-4..1e8, gamma 1e5: the single pass (refine_max=0, the
    reference) leaves resid4 above 1e-9, as the oracle's single pass does;
    with refinement resid4 <= 1e-10 and the solution matches the oracle's
    refined one."""
    g = GridPattern(50, 65, seed=3, gamma=1e5, ds_range=(1e-4, 1e8))
    sy = g.system()
    Hc, Jc, Jd = to_csr(sy.H), to_csr(sy.Jc), to_csr(sy.Jd)
    s = H.HykktSolver(sy.gamma)
    bind(s, Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd)
    s.options(refine_max=0)
    assert s.execute() == 0
    st0 = s.stats()
    assert st0["refine_steps"] == 0 and st0["resid4"] > 1e-9
    s.options(refine_max=3, refine_tol=1e-9)
    assert s.execute() == 0
    st = s.stats()
    assert st["refine_steps"] >= 1 and st["resid4"] <= 1e-10
    assert st["resid4_first"] == st0["resid4"]
    ref = O.execute(Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd, sy.gamma,
                    perm=s.permutation(), refine_tol=1e-9)
    x = np.concatenate(s.solution())
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert rel(x, xo) < 1e-8


def test_delta_hook_gpu():
    """add_diag(delta) on the device: the single pass with delta solves the
    KKT system whose H is H + delta diag(max_d)^-2 (the scaled diagonal
    shift in the caller's variables) and matches the oracle run with the same
    delta; with refinement the unshifted system's resid4 comes back below
    1e-9 for a small delta; delta = 0 again gives the unshifted result bit for
    bit."""
    from test_oracle import dense_kkt
    g = GridPattern(50, 65, seed=5, gamma=1e4, ds_range=(1e-4, 1e4))
    sy = g.system()
    Hc, Jc, Jd = to_csr(sy.H), to_csr(sy.Jc), to_csr(sy.Jd)
    s = H.HykktSolver(sy.gamma)
    bind(s, Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd)
    assert s.execute() == 0
    x_plain = np.concatenate(s.solution())
    rhs = np.concatenate([sy.rx, sy.rs, sy.ry, sy.ryd])
    K = dense_kkt(Hc, sy.Ds.v, Jc, Jd)
    delta = 1e-6
    s.options(delta=delta, refine_max=0)
    s.execute()
    st = s.stats()
    x = np.concatenate(s.solution())
    md = s.debug(4, st["n"] + st["m"])[:Hc.n]
    ref = O.execute(Hc, sy.Ds.v, Jc, Jd, sy.rx, sy.rs, sy.ry, sy.ryd, sy.gamma,
                    perm=s.permutation(), intermediates=True, refine_max=0, delta=delta)
    assert np.array_equal(md, ref["max_d"][:Hc.n])
    assert rel(x, np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])) < 1e-8
    Kd = K.copy()
    Kd[np.arange(Hc.n), np.arange(Hc.n)] += delta / md ** 2
    assert np.linalg.norm(Kd @ x - rhs) / np.linalg.norm(rhs) <= 1e-9
    assert st["resid4"] > 1e-6   # the unshifted system is not what was solved
    s.options(delta=1e-8, refine_max=3, refine_tol=1e-10)
    s.execute()
    st = s.stats()
    assert st["refine_steps"] >= 1 and st["resid4"] <= 1e-9
    s.options(delta=0.0, refine_tol=9e-9)   # the first execute's options again
    assert s.execute() == 0
    assert np.array_equal(np.concatenate(s.solution()), x_plain)
