"""Pins the CPU oracle (oracle/hykkt_oracle.c) before anything is compared to it.

Each check is independent of the oracle's own formulas: the reference drivers'
known-answer tests, the reference's own host sources (oracle/_ref), dense
brute force, and closed-form properties.
"""
import os

import numpy as np
import pytest

import _oracle as O
import fixtures as F
from hykkt_amd.generator import GridPattern

RNG = np.random.default_rng(1234)


def rand_csr(n, m, density, rng, sym=False, diag=0.0):
    A = (rng.random((n, m)) < density) * rng.standard_normal((n, m))
    if sym:
        A = np.triu(A) + np.triu(A, 1).T
    if diag:
        A = A + diag * np.eye(n)
    rp = [0]
    ci, v = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci += nz.tolist()
        v += A[i, nz].tolist()
        rp.append(len(ci))
    return O.Csr(n, m, rp, ci, v), A


# ---------------------------------------------------------------- KATs
def test_ruiz_kat():
    """examples/Ruiz_driver.cpp:22-49,54-182 (6 probed values, tol 1e-8)."""
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
    A.v[A.ci == np.repeat(np.arange(n), np.diff(A.rp))] += 1.0   # add_diag(1.0), Ruiz_driver.cpp:89
    H = O.Csr(n, n, np.arange(n + 1), np.arange(n), np.full(n, np.sqrt(n)))
    Hs, Ats, As, r1, r2, md = O.ruiz(2, H, O.transpose(A), A, np.ones(n), np.ones(n))
    tol = 1e-8
    assert abs(Hs.v[n // 2 - 1] - 0.062378167641326) < tol
    assert abs(As.v[-1] - 0.005524271728020) < tol
    assert abs(Ats.v[1] - 0.5) < tol
    assert abs(r1[n // 2 - 1] - 0.044151078568835) < tol
    assert abs(r2[n // 2 - 1] - 0.044194173824159) < tol
    assert abs(md[32] - 0.171498585142) < tol


def test_perm_kat():
    """examples/perm_driver.cpp:15-25 expected arrays (exact); the output of the
    column map is compared (the driver itself compares its input, :89)."""
    a_i = [0, 2, 5, 7, 9]
    a_j = [0, 2, 0, 1, 2, 1, 2, 1, 3]
    rev, mp = O.perm_maps(a_i, a_j, [2, 0, 3, 1])
    assert list(mp["rc"][0]) == [0, 2, 4, 6, 9]
    assert list(mp["rc"][1]) == [0, 3, 0, 1, 2, 3, 0, 1, 3]
    assert list(mp["r"][0]) == [0, 2, 4, 6, 9]
    assert list(mp["r"][1]) == [1, 2, 0, 2, 1, 3, 0, 1, 2]
    assert list(mp["c"][0]) == [0, 1, 0, 1, 3, 0, 3, 2, 3]
    assert list(rev) == [1, 3, 0, 2]


def test_perm_maps_closed_form():
    """P A P^T from the maps equals the dense permuted matrix."""
    A, D = rand_csr(40, 40, 0.15, RNG, sym=True, diag=1.0)
    perm = RNG.permutation(40).astype(np.int32)
    rev, mp = O.perm_maps(A.rp, A.ci, perm)
    prp, pci, pmap = mp["rc"]
    P = O.Csr(40, 40, prp, pci, A.v[pmap]).dense()
    assert np.array_equal(P, D[np.ix_(perm, perm)])
    pcols, pmap = mp["c"]
    assert np.array_equal(O.Csr(40, 40, A.rp, pcols, A.v[pmap]).dense(), D[:, perm])
    prp, pcols, pmap = mp["r"]
    assert np.array_equal(O.Csr(40, 40, prp, pcols, A.v[pmap]).dense(), D[perm, :])
    for i in range(40):   # sorted columns within permuted rows
        assert np.all(np.diff(mp["rc"][1][mp["rc"][0][i]:mp["rc"][0][i + 1]]) > 0)


@pytest.mark.parametrize("kind", [0, 1])
def test_cholesky_kat(mats, kind):
    """examples/cuSolver_driver_cholesky.cpp:32-75: lower-triangle file read as
    general, x[0], x[6] within 3e-2; plus the exact dense solution.  Both
    forms of the oracle factorization (up-looking, supernodal)."""
    H = O.read_mm(mats("a200_chol/H_matrix_ACTIVSg200_AC_00.mtx"), 3, False)
    b = O.read_rhs(mats("a200_chol/rhs_chol_ACTIVSg200_AC_00.mtx"))
    ch = O.Chol(H, kind)
    assert ch.factor(1e-12) == -1
    x = ch.solve(b)
    assert abs(x[0] + 97.663191790329750) < 3e-2
    assert abs(x[6] + 1.389053208413921) < 3e-2
    Dl = np.tril(H.dense())
    Hd = Dl + np.tril(Dl, -1).T
    xd = np.linalg.solve(Hd, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    assert np.linalg.norm(Hd @ x - b) / np.linalg.norm(b) < 1e-12


def _spd(n, dens, seed, shift=2.0):
    import scipy.sparse as sp
    A = sp.random(n, n, density=dens, random_state=seed, format="csr")
    A = (A + A.T + sp.eye(n) * (0.1 * n * dens * 10 + shift)).tocsr()
    A.sort_indices()
    return A, O.Csr(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                    A.data.astype(np.float64))


@pytest.mark.parametrize("n,dens,seed", [(1, 1.0, 0), (7, 0.4, 1), (60, 0.08, 2),
                                         (400, 0.015, 3), (4000, 0.001, 4)])
def test_cholesky_supernodal_matches_uplooking(n, dens, seed):
    """The supernodal form factors the same matrix in another order: same
    nnz(L) and flop count (pattern only), solutions equal to rounding, and a
    small residual against the matrix itself (scipy)."""
    A, C = _spd(n, dens, seed)
    up, snd = O.Chol(C, 0), O.Chol(C, 1)
    assert up.factor() == -1 and snd.factor() == -1
    assert up.nnz == snd.nnz and up.flops == snd.flops
    b = np.random.default_rng(seed).standard_normal(n)
    x0, x1 = up.solve(b), snd.solve(b)
    assert np.linalg.norm(x1 - x0) <= 1e-12 * np.linalg.norm(x0)
    assert np.linalg.norm(A @ x1 - b) <= 1e-12 * np.linalg.norm(b)


def test_cholesky_supernodal_dense_and_arrow():
    """Dense SPD (one supernode) and a block-arrow matrix (several supernodes
    feeding a dense border) against numpy's dense Cholesky solve."""
    rng = np.random.default_rng(5)
    M = rng.standard_normal((150, 150))
    D = M @ M.T + 150 * np.eye(150)
    blocks = [D]
    nb, bs, border = 3, 90, 40
    n = nb * bs + border
    Aa = np.zeros((n, n))
    for k in range(nb):
        Q = rng.standard_normal((bs, bs))
        Aa[k * bs:(k + 1) * bs, k * bs:(k + 1) * bs] = Q @ Q.T + bs * np.eye(bs)
        Bk = rng.standard_normal((border, bs))
        Aa[nb * bs:, k * bs:(k + 1) * bs] = Bk
        Aa[k * bs:(k + 1) * bs, nb * bs:] = Bk.T
    Aa[nb * bs:, nb * bs:] = 10 * n * np.eye(border)
    blocks.append(Aa)
    for Dm in blocks:
        import scipy.sparse as sp
        m = Dm.shape[0]
        nzr, nzc = np.nonzero(Dm)
        S = sp.csr_matrix((Dm[nzr, nzc], (nzr, nzc)), shape=(m, m))
        S.sort_indices()
        C = O.Csr(m, m, S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data)
        ch = O.Chol(C, 1)
        assert ch.factor() == -1
        b = rng.standard_normal(m)
        x = ch.solve(b)
        xd = np.linalg.solve(Dm, b)
        assert np.linalg.norm(x - xd) <= 1e-11 * np.linalg.norm(xd)


@pytest.mark.parametrize("kind", [0, 1])
def test_cholesky_pivot_failure_index(kind):
    """One negative diagonal entry: both forms report that row."""
    n = 30
    import scipy.sparse as sp
    A = sp.diags([np.full(n - 1, 0.1), np.full(n, 4.0), np.full(n - 1, 0.1)], [-1, 0, 1]).tolil()
    A[17, 17] = -1.0
    A = A.tocsr()
    A.sort_indices()
    C = O.Csr(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    assert O.Chol(C, kind).factor() == 17


def test_schur_cg_kat(mats):
    """examples/cuSolver_driver_schur_cg.cpp:130-154 (x0 = 0, tol 1e-6), and the
    dense direct solve of S = J H^-1 J^T."""
    J = O.read_mm(mats("a200_schur/JC_matrix_ACTIVSg200_AC_00.mtx"), 3, False)
    H = O.read_mm(mats("a200_schur/H_matrix_ACTIVSg200_AC_00.mtx"), 3, False)
    b = O.read_rhs(mats("a200_schur/CG_rhs_ACTIVSg200_AC_00.mtx"))
    ch = O.Chol(H)
    assert ch.factor(1e-12) == -1
    x, st, it, fn = O.schur_cg(J, O.transpose(J), ch, b)
    assert st == 0 and fn < 1e-12
    assert abs(x[0] - 22.171865776354700) < 1e-6
    assert abs(x[6] + 4446.628667344612) < 1e-6
    Dl = np.tril(H.dense())
    Hd = Dl + np.tril(Dl, -1).T
    Jd = J.dense()
    S = Jd @ np.linalg.solve(Hd, Jd.T)
    xd = np.linalg.solve(S, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8


# ---------------------------------------------------------------- ingest
@pytest.mark.parametrize("rel,skip,sym,nnz", [
    ("a200/block_H_matrix_ACTIVSg200_AC_09.mtx", 2, True, 4632),
    ("a2000/block_H_matrix_ACTIVSg2000_AC_09.mtx", 3, True, 51137),
    ("a200/block_J_matrix_ACTIVSg200_AC_10.mtx", 2, False, 6784),
    ("a200/block_Jd_matrix_ACTIVSg200_AC_00.mtx", 2, False, 0),
    ("a200_chol/H_matrix_ACTIVSg200_AC_00.mtx", 3, False, 11304),
])
def test_ingest_counts_and_reference(mats, rel, skip, sym, nnz):
    A = O.read_mm(mats(rel), skip, sym)
    assert A.nnz == nnz
    for i in range(A.n):
        assert np.all(np.diff(A.ci[A.rp[i]:A.rp[i + 1]]) > 0)
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes as C
    n, m = C.c_int(), C.c_int()
    rp = np.zeros(A.n + 2, np.int32)
    ci = np.zeros(max(nnz, 1) + 1, np.int32)
    cv = np.zeros(max(nnz, 1) + 1)
    k = R.ref_read_csr(mats(rel).encode(), skip, int(sym), C.byref(n), C.byref(m), len(ci),
                       rp, ci, cv)
    assert k == nnz and n.value == A.n and m.value == A.m
    assert np.array_equal(rp[:A.n + 1], A.rp)
    assert np.array_equal(ci[:nnz], A.ci) and np.array_equal(cv[:nnz], A.v)


def test_read_rhs_reference(mats):
    rel = "a2000/block_rx_ACTIVSg2000_AC_09.mtx"
    x = O.read_rhs(mats(rel))
    assert len(x) == 25910
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    y = np.zeros(len(x))
    R.ref_read_rhs(mats(rel).encode(), y)
    assert np.array_equal(x, y)


# ---------------------------------------------------------------- kernels
def test_transpose_and_spmv_dense():
    A, D = rand_csr(37, 53, 0.1, RNG)
    At = O.transpose(A)
    assert np.array_equal(At.dense(), D.T)
    x = RNG.standard_normal(53)
    y = RNG.standard_normal(37)
    np.testing.assert_allclose(O.spmv(A, x, 2.0, -0.5, y), 2.0 * D @ x - 0.5 * y, rtol=1e-13,
                               atol=1e-13)


def test_spgemm_dense_brute_force():
    """Pattern = structural union (no dropping); values = D + alpha*A*B."""
    A, Ad = rand_csr(30, 20, 0.15, RNG)
    B, Bd = rand_csr(20, 30, 0.15, RNG)
    Dm, Dd = rand_csr(30, 30, 0.1, RNG)
    E = O.spgemm(A, B, Dm, alpha_p=3.5)
    struct = ((Ad != 0).astype(int) @ (Bd != 0).astype(int) > 0) | (Dd != 0)
    Ed = np.zeros((30, 30), bool)
    for i in range(30):
        Ed[i, E.ci[E.rp[i]:E.rp[i + 1]]] = True
    assert np.array_equal(Ed, struct)
    np.testing.assert_allclose(E.dense(), Dd + 3.5 * Ad @ Bd, rtol=1e-13, atol=1e-13)
    # exact cancellation must keep the entry (structural, not numerical)
    A2 = O.Csr(1, 2, [0, 2], [0, 1], [1.0, 1.0])
    B2 = O.Csr(2, 1, [0, 1, 2], [0, 0], [1.0, -1.0])
    Z = O.Csr(1, 1, [0, 0], [], [])
    E2 = O.spgemm(A2, B2, Z)
    assert E2.nnz == 1 and E2.v[0] == 0.0


# ---------------------------------------------------------------- end to end
def dense_kkt(H, Dsv, Jc, Jd):
    n, m, d = H.n, Jc.n, len(Dsv)
    K = np.zeros((n + d + m + d,) * 2)
    Hd, Jcd = H.dense(), Jc.dense()
    Jdd = Jd.dense() if Jd.nnz else np.zeros((d, n))
    o1, o2, o3 = n, n + d, n + d + m
    K[:n, :n] = Hd
    K[:n, o2:o3] = Jcd.T
    K[:n, o3:] = Jdd.T
    K[o1:o2, o1:o2] = np.diag(Dsv)
    K[o1:o2, o3:] = -np.eye(d)
    K[o2:o3, :n] = Jcd
    K[o3:, :n] = Jdd
    K[o3:, o1:o2] = -np.eye(d)
    return K


def test_execute_matches_dense_kkt_lu():
    """Small generated system: oracle solution vs dense LU of the full 4x4 KKT."""
    g = GridPattern(20, 26, seed=3, gamma=1e4, ds_range=(1e-4, 1e4))
    s = g.system()
    c = lambda b: O.Csr(b.n, b.m, b.rp, b.ci, b.v)
    H, Jc, Jd = c(s.H), c(s.Jc), c(s.Jd)
    out = O.execute(H, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma)
    K = dense_kkt(H, s.Ds.v, Jc, Jd)
    rhs = np.concatenate([s.rx, s.rs, s.ry, s.ryd])
    sol = np.linalg.solve(K, rhs)
    mine = np.concatenate([out["dx"], out["ds"], out["dyc"], out["dyd"]])
    assert out["status"] == 0
    assert np.linalg.norm(mine - sol) / np.linalg.norm(sol) < 1e-8
    assert out["stats"]["resid4"] < 1e-8
    assert np.linalg.norm(K @ mine - rhs) / np.linalg.norm(rhs) < 1e-8


@pytest.mark.parametrize("with_jd", [True, False])
def test_kkt_residual_vs_dense(with_jd):
    """orc_kkt_residual (HykktSolver.cpp:846-930 plus block rows 2 and 4)
    against numpy's ||K x - rhs|| on the dense 4x4 block KKT matrix.  x is
    the oracle's solution perturbed in one block at a time, so every block
    row carries a residual far above rounding; a dropped row, a wrong sign
    or a wrong denominator changes a value by O(1).  resid2 is rows 1 and 3
    (the reference's form), resid4 all four rows, both over ||rhs||."""
    g = GridPattern(20, 26, seed=4, gamma=1e4, ds_range=(1e-3, 1e3))
    s = g.system()
    c = lambda b: O.Csr(b.n, b.m, b.rp, b.ci, b.v)
    H, Jc, Jd = c(s.H), c(s.Jc), c(s.Jd)
    if not with_jd:
        Jd = O.Csr(Jd.n, Jd.m, np.zeros(Jd.n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    n, m, d = H.n, Jc.n, len(s.Ds.v)
    out = O.execute(H, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma)
    K = dense_kkt(H, s.Ds.v, Jc, Jd)
    rhs = np.concatenate([s.rx, s.rs, s.ry, s.ryd])
    x0 = np.concatenate([out["dx"], out["ds"], out["dyc"], out["dyd"]])
    off = np.cumsum([0, n, d, m, d])
    rng = np.random.default_rng(1)
    for blk in range(4):
        x = x0.copy()
        x[off[blk]:off[blk + 1]] += 1e-3 * rng.standard_normal(off[blk + 1] - off[blk])
        res = K @ x - rhs
        rows = [res[off[k]:off[k + 1]] for k in range(4)]
        nb = np.linalg.norm(rhs)
        want4 = np.linalg.norm(res) / nb
        want2 = np.sqrt(rows[0] @ rows[0] + rows[2] @ rows[2]) / nb
        got2, got4 = O.kkt_residual(H, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd,
                                    *(x[off[k]:off[k + 1]] for k in range(4)))
        # rows the perturbation does not reach keep the solution's own
        # residual (~1e-10 here), which the two evaluation orders round
        # differently: 1e-13 absolute covers that, far below the O(1e-3)
        # values of the perturbed rows
        assert want4 > 1e-6, blk   # the perturbation dominates rounding
        assert abs(got4 - want4) <= 1e-12 * want4 + 1e-13, (blk, got4, want4)
        assert abs(got2 - want2) <= 1e-12 * want2 + 1e-13, (blk, got2, want2)
        # which rows the perturbed block reaches (column structure of K)
        touched = [np.linalg.norm(r) > 1e-8 for r in rows]
        expect = {0: [True, False, True, with_jd], 1: [False, True, False, True],
                  2: [True, False, False, False], 3: [with_jd, True, False, False]}[blk]
        assert touched == expect, (blk, touched)


@pytest.mark.parametrize("kind,step,nnz_htil,nnz_hgam", [
    (F.A200, "09", 5612, 21318), (F.A200, "10", 5612, 21318), (F.A200, "00", 2278, 20338),
    (F.A2000, "09", 63961, 282614)])
def test_execute_fixtures(mats, kind, step, nnz_htil, nnz_hgam):
    """CTest hybrid/solver cases (examples/CMakeLists.txt:22-86): residual < 1e-2
    there; here <= 1e-8 over all four block rows.  Pattern counts are the
    structural SpGEMM counts (no value dropping)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    s = F.load(mats, kind, step)
    H, Jc = s["H"], s["Jc"]
    P = sp.csr_matrix((np.abs(H.v), H.ci, H.rp), shape=(H.n, H.n))
    J = sp.csr_matrix((np.abs(Jc.v), Jc.ci, Jc.rp), shape=(Jc.n, Jc.m))
    perm = reverse_cuthill_mckee((P + J.T @ J).tocsr(), symmetric_mode=True).astype(np.int32)
    out = O.execute(H, s["Ds"].v, Jc, s["Jd"], s["rx"], s["rs"], s["ry"], s["ryd"], 1e4,
                    perm=perm)
    st = out["stats"]
    assert st["nnz_htil"] == nnz_htil and st["nnz_hgam"] == nnz_hgam
    assert out["status"] == 0 and st["pivot"] == -1 and st["cg_status"] == 0
    assert st["resid4"] <= 1e-8
    # the supernodal form of the factorization gives the same solve
    sn = O.execute(H, s["Ds"].v, Jc, s["Jd"], s["rx"], s["rs"], s["ry"], s["ryd"], 1e4,
                   perm=perm, chol_kind=1)
    assert sn["status"] == 0 and sn["stats"]["resid4"] <= 1e-8
    for k in ("dx", "ds", "dyc", "dyd"):
        assert np.linalg.norm(sn[k] - out[k]) <= 1e-7 * max(np.linalg.norm(out[k]), 1e-300)


def test_execute_iterative_refinement():
    """Ds spread 1e-4..1e8 with gamma 1e5: the reference's single pass leaves
    resid4 ~ 1e-7; refinement on the 4x4 system brings it below 1e-10 and the
    solution onto the dense LU solution.  refine_max=0 is the single pass."""
    g = GridPattern(50, 65, seed=3, gamma=1e5, ds_range=(1e-4, 1e8))
    s = g.system()
    c = lambda b: O.Csr(b.n, b.m, b.rp, b.ci, b.v)
    H, Jc, Jd = c(s.H), c(s.Jc), c(s.Jd)
    one = O.execute(H, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma, refine_max=0)
    ref = O.execute(H, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma, refine_tol=1e-9)
    assert one["stats"]["refine_steps"] == 0
    assert one["stats"]["resid4"] == one["stats"]["resid4_first"] > 1e-9
    assert ref["stats"]["resid4_first"] == one["stats"]["resid4"]
    assert ref["stats"]["refine_steps"] >= 1 and ref["stats"]["resid4"] <= 1e-10
    K = dense_kkt(H, s.Ds.v, Jc, Jd)
    rhs = np.concatenate([s.rx, s.rs, s.ry, s.ryd])
    sol = np.linalg.solve(K, rhs)
    mine = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert np.linalg.norm(K @ mine - rhs) / np.linalg.norm(rhs) <= 1e-10
    single = np.concatenate([one["dx"], one["ds"], one["dyc"], one["dyd"]])
    assert np.linalg.norm(mine - sol) < np.linalg.norm(single - sol)


def test_delta_hook_solves_shifted_system():
    """add_diag(delta) adds delta to the diagonal of the Ruiz-scaled, permuted
    H_gamma.  In the caller's variables x = D x^ (D = diag(max_d[:n])) that is
    H + delta D^-2, so the single pass solves the dense 4x4 KKT with that H to
    the usual accuracy, while its residual against the unshifted system grows
    with delta; refinement against the unshifted system recovers it for a
    small delta.  delta = 0 is bit for bit the default."""
    g = GridPattern(50, 65, seed=5, gamma=1e4, ds_range=(1e-4, 1e4))
    s = g.system()
    c = lambda b: O.Csr(b.n, b.m, b.rp, b.ci, b.v)
    H, Jc, Jd = c(s.H), c(s.Jc), c(s.Jd)
    args = (H, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma)
    base = O.execute(*args)
    zero = O.execute(*args, delta=0.0)
    assert np.array_equal(base["dx"], zero["dx"]) and np.array_equal(base["dyc"], zero["dyc"])
    rhs = np.concatenate([s.rx, s.rs, s.ry, s.ryd])
    for delta in (1e-8, 1e-4):
        one = O.execute(*args, delta=delta, refine_max=0, intermediates=True)
        md = one["max_d"][:H.n]
        K = dense_kkt(H, s.Ds.v, Jc, Jd)
        Kd = K.copy()
        Kd[np.arange(H.n), np.arange(H.n)] += delta / md ** 2
        x = np.concatenate([one["dx"], one["ds"], one["dyc"], one["dyd"]])
        shifted = np.linalg.norm(Kd @ x - rhs) / np.linalg.norm(rhs)
        plain = np.linalg.norm(K @ x - rhs) / np.linalg.norm(rhs)
        assert shifted <= 1e-9 and plain > 1e3 * shifted, (delta, shifted, plain)
        assert np.linalg.norm(x - np.linalg.solve(Kd, rhs)) <= 1e-6 * np.linalg.norm(x)
    refined = O.execute(*args, delta=1e-8, refine_tol=1e-10)
    assert refined["stats"]["refine_steps"] >= 1 and refined["stats"]["resid4"] <= 1e-9
