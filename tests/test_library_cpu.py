"""CPU-only checks of libhykkt.so: it loads, exports every symbol include/hykkt.h
declares, and its host-side symbolic layer (reference permutation maps, AMD
ordering, supernodal Cholesky analysis) agrees with the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

import _oracle as O
import fixtures as F
import hykkt_amd as H
from hykkt_amd.generator import GridPattern, workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RNG = np.random.default_rng(7)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "hykkt.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hykkt_\w+)\s*\(", txt)))


def test_library_loads_and_exports_header_symbols():
    H.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (hykkt_\w+)", out))
    declared = header_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert set(H.exported_symbols()) <= exported


def test_perm_kat_product():
    """examples/perm_driver.cpp:15-25 through the product's C ABI."""
    mp = H.perm_maps([0, 2, 5, 7, 9], [0, 2, 0, 1, 2, 1, 2, 1, 3], [2, 0, 3, 1])
    assert list(mp["rc"][0]) == [0, 2, 4, 6, 9]
    assert list(mp["rc"][1]) == [0, 3, 0, 1, 2, 3, 0, 1, 3]
    assert list(mp["r"][1]) == [1, 2, 0, 2, 1, 3, 0, 1, 2]
    assert list(mp["c"][0]) == [0, 1, 0, 1, 3, 0, 3, 2, 3]


def sym_pattern(n, density, rng):
    A = sp.random(n, n, density=density, random_state=np.random.RandomState(rng.integers(1 << 30)))
    A = ((A + A.T) != 0).astype(float) + sp.eye(n)
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A


def test_perm_maps_bit_exact_vs_oracle():
    A = sym_pattern(400, 0.01, RNG)
    perm = RNG.permutation(400).astype(np.int32)
    mine = H.perm_maps(A.indptr, A.indices, perm)
    rev, ref = O.perm_maps(A.indptr, A.indices, perm)
    assert np.array_equal(mine["rev"], rev)
    for k in ("rc", "r", "c"):
        for a, b in zip(mine[k], ref[k]):
            assert np.array_equal(a, b), k
    J = sp.csr_matrix(sp.random(150, 400, density=0.02, random_state=3))
    J.sort_indices()
    mine = H.perm_maps(J.indptr, J.indices, perm, square=False)
    _, ref = O.perm_maps(J.indptr, J.indices, perm, nrows=150)
    assert all(np.array_equal(a, b) for a, b in zip(mine["c"], ref["c"]))


def hgamma_pattern(g):
    Hm = sp.csr_matrix((np.ones(g.H.nnz), g.H.ci, g.H.rp), shape=(g.n, g.n))
    J = sp.csr_matrix((np.ones(g.Jc.nnz), g.Jc.ci, g.Jc.rp), shape=(g.m, g.n))
    D = sp.csr_matrix((np.ones(g.Jd.nnz), g.Jd.ci, g.Jd.rp), shape=(g.d, g.n))
    P = (Hm + J.T @ J + D.T @ D).tocsr()
    P.sort_indices()
    return P


def permuted(P, perm):
    mp = H.perm_maps(P.indptr, P.indices, perm)
    return mp["rc"][0], mp["rc"][1]


@pytest.mark.parametrize("nbus", [60, 500])
def test_chol_analysis_matches_oracle_counts(nbus):
    """Column counts / flops of the supernodal analysis == the oracle's
    elimination-tree reach count (brute force over every row)."""
    g = GridPattern(nbus, int(1.4 * nbus), seed=nbus)
    P = hgamma_pattern(g)
    for perm in (H.amd_order(P.indptr, P.indices), np.arange(P.shape[0], dtype=np.int32),
                 RNG.permutation(P.shape[0]).astype(np.int32)):
        rp, ci = permuted(P, perm)
        info = H.chol_analyze_host(rp, ci)
        ch = O.Chol(O.Csr(P.shape[0], P.shape[0], rp, ci, np.ones(len(ci))))
        assert info["nnz_l"] == ch.nnz
        assert info["flops"] == ch.flops


def test_amd_valid_deterministic_and_fill_reducing():
    g = workload("grid500")
    P = hgamma_pattern(g)
    n = P.shape[0]
    p1 = H.amd_order(P.indptr, P.indices)
    p2 = H.amd_order(P.indptr, P.indices)
    assert np.array_equal(np.sort(p1), np.arange(n))
    assert np.array_equal(p1, p2)
    nnz_amd = H.chol_analyze_host(*permuted(P, p1))["nnz_l"]
    rcm = reverse_cuthill_mckee(P, symmetric_mode=True).astype(np.int32)
    nnz_rcm = H.chol_analyze_host(*permuted(P, rcm))["nnz_l"]
    nnz_nat = H.chol_analyze_host(P.indptr, P.indices)["nnz_l"]
    assert nnz_amd < 0.6 * nnz_rcm and nnz_amd < 0.5 * nnz_nat, (nnz_amd, nnz_rcm, nnz_nat)


def test_amd_vs_superlu_mmd(mats):
    """Fill of our AMD within 25% of SciPy SuperLU's MMD(A^T+A) on the a2000
    fixture's H_gamma pattern (an independent minimum-degree implementation)."""
    s = F.load(mats, F.A2000, "09")
    Hm = sp.csr_matrix((np.ones(s["H"].nnz), s["H"].ci, s["H"].rp), shape=(s["H"].n,) * 2)
    J = sp.csr_matrix((np.ones(s["Jc"].nnz), s["Jc"].ci, s["Jc"].rp), shape=(s["Jc"].n, s["Jc"].m))
    Dm = sp.csr_matrix((np.ones(s["Jd"].nnz), s["Jd"].ci, s["Jd"].rp), shape=(s["Jd"].n, s["Jd"].m))
    P = (Hm + J.T @ J + Dm.T @ Dm).tocsr()
    P.sort_indices()
    assert P.nnz == 282614
    ours = H.chol_analyze_host(*permuted(P, H.amd_order(P.indptr, P.indices)))["nnz_l"]
    A = (P + 10 * P.shape[0] * sp.eye(P.shape[0])).tocsc()
    lu = sp.linalg.splu(A, permc_spec="MMD_AT_PLUS_A",
                        options=dict(SymmetricMode=True), diag_pivot_thresh=0.0)
    mmd = H.chol_analyze_host(*permuted(P, lu.perm_c.astype(np.int32)))["nnz_l"]
    assert ours <= 1.25 * mmd, (ours, mmd)


def test_errors_are_codes_not_exits():
    """A bad call returns a negative code and a message (the reference exits)."""
    rp = np.array([0, 1], np.int32)
    r = H.lib().hykkt_amd_order(1, rp.ctypes.data, None, None)
    assert r == -1 and b"null" in H.lib().hykkt_last_error()


def _fixture_files():
    """Every fixture file of the reference (src/mats) with the skip count its
    CTest case or driver uses and whether it is read as a symmetric block."""
    out = []
    base = os.path.join(ROOT, "tests", "golden", "mats")
    for d in sorted(os.listdir(base)):
        skip = 2 if d == "a200" else 3
        for f in sorted(os.listdir(os.path.join(base, d))):
            rel = f"{d}/{f[:-3]}"
            if "rhs" in f or f.startswith(("block_rx", "block_rs", "block_ry")):
                out.append((rel, 0, None))
            else:
                out.append((rel, skip, False))
                if f.startswith("block_H_matrix"):   # read_matrix_files: H is symmetric
                    out.append((rel, skip, True))
    return out


@pytest.mark.parametrize("rel,skip,sym", _fixture_files())
def test_product_ingest_matches_oracle_and_reference(mats, rel, skip, sym):
    """hykkt_mm_read_csr / hykkt_mm_read_rhs (the product's readers behind
    read_matrix_files and the drivers) against the oracle and, where it is
    built, the reference's own input_functions.cpp, bit for bit."""
    import ctypes as C
    path = mats(rel)
    R = O.ref_lib()
    if sym is None:
        mine = H.mm_read_rhs(path)
        ref = O.read_rhs(path)
        assert np.array_equal(mine, ref) and len(mine) > 0
        if R is not None:
            y = np.zeros(len(ref))
            R.ref_read_rhs(path.encode(), y)
            assert np.array_equal(mine, y)
        return
    n, m, rp, ci, v = H.mm_read_csr(path, skip, sym)
    A = O.read_mm(path, skip, sym)
    assert (n, m) == (A.n, A.m)
    assert np.array_equal(rp, A.rp) and np.array_equal(ci, A.ci) and np.array_equal(v, A.v)
    if R is not None:
        nn, mm = C.c_int(), C.c_int()
        rrp = np.zeros(n + 2, np.int32)
        rci = np.zeros(len(ci) + 1, np.int32)
        rcv = np.zeros(len(ci) + 1)
        k = R.ref_read_csr(path.encode(), skip, int(sym), C.byref(nn), C.byref(mm), len(rci),
                           rrp, rci, rcv)
        assert k == len(ci) and (nn.value, mm.value) == (n, m)
        assert np.array_equal(rrp[:n + 1], rp)
        assert np.array_equal(rci[:k], ci) and np.array_equal(rcv[:k], v)
