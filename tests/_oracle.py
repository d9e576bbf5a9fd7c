"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the reference's
own host sources built under oracle/_ref (libhykkt_ref.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libhykkt_ref.so")

_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None
_ref = None


def _build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            _build_oracle()
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.orc_transpose.argtypes = [C.c_int, C.c_int, _i, _i, _d, _i, _i, _d]
        L.orc_spmv.argtypes = [C.c_int, _i, _i, _d, C.c_double, _d, C.c_double, _d]
        L.orc_spgemm_nnz.argtypes = [C.c_int, C.c_int, _i, _i, _i, _i, _i, _i]
        L.orc_spgemm_pattern.argtypes = [C.c_int, C.c_int, _i, _i, _i, _i, _i, _i, _i, _i]
        L.orc_spgemm_values.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        _i, _i, _d, _i, _i, _d, _i, _i, _d, _i, _i, _d]
        L.orc_ruiz.argtypes = [C.c_int, C.c_int, C.c_int, _i, _i, _d, _i, _i, _d, _i, _i, _d,
                               _d, _d, _d]
        L.orc_perm_rev.argtypes = [C.c_int, _i, _i]
        L.orc_map_rc.argtypes = [C.c_int, _i, _i, _i, _i, _i, _i, _i]
        L.orc_map_c.argtypes = [C.c_int, _i, _i, _i, _i, _i]
        L.orc_map_r.argtypes = [C.c_int, _i, _i, _i, _i, _i, _i]
        L.orc_chol_create.argtypes = [C.c_int, _i, _i]
        L.orc_chol_create.restype = vp
        L.orc_chol_create_kind.argtypes = [C.c_int, _i, _i, C.c_int]
        L.orc_chol_create_kind.restype = vp
        L.orc_chol_factor.argtypes = [vp, _d, C.c_double]
        L.orc_chol_solve.argtypes = [vp, _d, _d]
        L.orc_chol_nnz.argtypes = [vp]
        L.orc_chol_nnz.restype = C.c_longlong
        L.orc_chol_flops.argtypes = [vp]
        L.orc_chol_flops.restype = C.c_double
        L.orc_chol_free.argtypes = [vp]
        L.orc_schur_cg.argtypes = [C.c_int, C.c_int, _i, _i, _d, _i, _i, _d, vp, _d, _d,
                                   C.c_double, C.c_int, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double)]
        L.orc_mm_header.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_mm_read_coo.argtypes = [C.c_char_p, C.c_int, C.c_int, _i, _i, _d]
        L.orc_sym_coo_nnz.argtypes = [C.c_int, C.c_int, _i, _i]
        L.orc_sym_coo_to_csr.argtypes = [C.c_int, C.c_int, _i, _i, _d, _i, _i, _d]
        L.orc_coo_to_csr.argtypes = [C.c_int, C.c_int, _i, _i, _d, _i, _i, _d]
        L.orc_read_rhs.argtypes = [C.c_char_p, C.c_int, _d]
        L.orc_execute.argtypes = [C.POINTER(OrcOpts),
                                  C.c_int, _i, _i, _d, C.c_int, _d,
                                  C.c_int, _i, _i, _d, vp, vp, vp, C.c_int,
                                  _d, _d, _d, _d, vp, vp,
                                  _d, _d, _d, _d, C.POINTER(OrcStats),
                                  C.POINTER(OrcIntermediates)]
        L.orc_kkt_residual.argtypes = [C.c_int, _i, _i, _d, C.c_int, _d, C.c_int, _i, _i, _d,
                                       vp, vp, vp, C.c_int, _d, _d, _d, _d, _d, _d, _d, _d,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


class OrcOpts(C.Structure):
    _fields_ = [("gamma", C.c_double), ("ruiz_its", C.c_int), ("cg_tol", C.c_double),
                ("cg_itmax", C.c_int), ("pivot_tol", C.c_double), ("norm_tol", C.c_double),
                ("chol_kind", C.c_int), ("refine_max", C.c_int), ("refine_tol", C.c_double),
                ("delta", C.c_double)]


class OrcStats(C.Structure):
    _fields_ = [("resid2", C.c_double), ("resid4", C.c_double), ("cg_iters", C.c_int),
                ("cg_status", C.c_int), ("pivot", C.c_int), ("nnz_htil", C.c_int),
                ("nnz_hgam", C.c_int), ("nnz_l", C.c_longlong), ("resid4_first", C.c_double),
                ("refine_steps", C.c_int), ("t_symbolic", C.c_double), ("t_numeric", C.c_double)]


class OrcIntermediates(C.Structure):
    _fields_ = [("htil_v", C.c_void_p), ("hgam_v", C.c_void_p), ("max_d", C.c_void_p),
                ("rx_hat", C.c_void_p), ("schur_rhs", C.c_void_p)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- CSR helpers
class Csr:
    """Plain CSR triple with int32 indices and float64 values."""

    def __init__(self, n, m, rp, ci, v):
        self.n, self.m = int(n), int(m)
        self.rp = np.ascontiguousarray(rp, dtype=np.int32)
        self.ci = np.ascontiguousarray(ci, dtype=np.int32)
        self.v = np.ascontiguousarray(v, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.rp[-1])

    def dense(self):
        A = np.zeros((self.n, self.m))
        for i in range(self.n):
            for p in range(self.rp[i], self.rp[i + 1]):
                A[i, self.ci[p]] += self.v[p]
        return A

    def copy(self):
        return Csr(self.n, self.m, self.rp.copy(), self.ci.copy(), self.v.copy())


def read_mm(path, skip, symmetric):
    """Oracle ingest (input_functions.cpp semantics)."""
    L = lib()
    n, m, nnz = C.c_int(), C.c_int(), C.c_int()
    assert L.orc_mm_header(path.encode(), skip, C.byref(n), C.byref(m), C.byref(nnz)) == 0
    r = np.zeros(max(nnz.value, 1), np.int32)
    c = np.zeros(max(nnz.value, 1), np.int32)
    v = np.zeros(max(nnz.value, 1), np.float64)
    k = L.orc_mm_read_coo(path.encode(), skip, nnz.value, r, c, v)
    assert k == nnz.value, (k, nnz.value)
    if symmetric:
        t = L.orc_sym_coo_nnz(n.value, k, r, c)
    else:
        t = k
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(max(t, 1), np.int32)
    cv = np.zeros(max(t, 1), np.float64)
    if symmetric:
        L.orc_sym_coo_to_csr(n.value, k, r, c, v, rp, ci, cv)
    else:
        L.orc_coo_to_csr(n.value, k, r, c, v, rp, ci, cv)
    return Csr(n.value, m.value, rp, ci[:t], cv[:t])


def read_rhs(path, cap=10_000_000):
    out = np.zeros(cap, np.float64)
    k = lib().orc_read_rhs(path.encode(), cap, out)
    assert k >= 0
    return out[:k].copy()


def transpose(A: Csr) -> Csr:
    trp = np.zeros(A.m + 1, np.int32)
    tci = np.zeros(max(A.nnz, 1), np.int32)
    tv = np.zeros(max(A.nnz, 1), np.float64)
    lib().orc_transpose(A.n, A.m, A.rp, A.ci, A.v, trp, tci, tv)
    return Csr(A.m, A.n, trp, tci[:A.nnz], tv[:A.nnz])


def spmv(A: Csr, x, alpha=1.0, beta=0.0, y=None):
    y = np.zeros(A.n) if y is None else np.array(y, dtype=np.float64)
    lib().orc_spmv(A.n, A.rp, A.ci, A.v, alpha, np.ascontiguousarray(x, np.float64), beta, y)
    return y


def spgemm(A: Csr, B: Csr, D: Csr, alpha_p=1.0, alpha_s=1.0, beta_s=1.0) -> Csr:
    L = lib()
    n, ncols = A.n, B.m
    nnz = L.orc_spgemm_nnz(n, ncols, A.rp, A.ci, B.rp, B.ci, D.rp, D.ci)
    erp = np.zeros(n + 1, np.int32)
    eci = np.zeros(max(nnz, 1), np.int32)
    L.orc_spgemm_pattern(n, ncols, A.rp, A.ci, B.rp, B.ci, D.rp, D.ci, erp, eci)
    ev = np.zeros(max(nnz, 1), np.float64)
    L.orc_spgemm_values(n, ncols, alpha_p, alpha_s, beta_s, A.rp, A.ci, A.v, B.rp, B.ci, B.v,
                        D.rp, D.ci, D.v, erp, eci, ev)
    return Csr(n, ncols, erp, eci[:nnz], ev[:nnz])


def ruiz(its, H: Csr, Jt: Csr, J: Csr, rhs1, rhs2):
    """Returns scaled copies (H, Jt, J, rhs1, rhs2, max_d)."""
    H, Jt, J = H.copy(), Jt.copy(), J.copy()
    r1 = np.array(rhs1, dtype=np.float64)
    r2 = np.array(rhs2, dtype=np.float64)
    n, totn = H.n, H.n + J.n
    md = np.zeros(totn)
    lib().orc_ruiz(its, n, totn, H.rp, H.ci, H.v, Jt.rp, Jt.ci, Jt.v, J.rp, J.ci, J.v, r1, r2,
                   md)
    return H, Jt, J, r1, r2, md


def perm_maps(rows, cols, perm, nrows=None):
    """(rev, maps).  maps["c"] = (pcols, pmap) for a matrix of `nrows` rows whose
    columns are permuted; for a square (len(perm) rows) matrix also
    maps["rc"] = (prows, pcols, pmap) and maps["r"] = (prows, pcols, pmap)."""
    L = lib()
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    perm = np.ascontiguousarray(perm, np.int32)
    n = len(perm)
    nr = len(rows) - 1 if nrows is None else nrows
    assert len(rows) == nr + 1
    nnz = int(rows[-1])
    rev = np.zeros(n, np.int32)
    L.orc_perm_rev(n, perm, rev)
    out = {}
    pcols = np.zeros(max(nnz, 1), np.int32)
    pmap = np.zeros(max(nnz, 1), np.int32)
    L.orc_map_c(nr, rows, cols, rev, pcols, pmap)
    out["c"] = (pcols[:nnz], pmap[:nnz])
    if nr == n:
        prows = np.zeros(n + 1, np.int32)
        pcols = np.zeros(max(nnz, 1), np.int32)
        pmap = np.zeros(max(nnz, 1), np.int32)
        L.orc_map_rc(n, rows, cols, perm, rev, prows, pcols, pmap)
        out["rc"] = (prows, pcols[:nnz], pmap[:nnz])
        prows = np.zeros(n + 1, np.int32)
        pcols = np.zeros(max(nnz, 1), np.int32)
        pmap = np.zeros(max(nnz, 1), np.int32)
        L.orc_map_r(n, rows, cols, perm, prows, pcols, pmap)
        out["r"] = (prows, pcols[:nnz], pmap[:nnz])
    return rev, out


class Chol:
    """Oracle sparse Cholesky of the lower triangle of a CSR matrix."""

    def __init__(self, A: Csr, kind=0):
        """kind 0: up-looking; kind 1: supernodal left-looking (OpenMP)."""
        self.A = A
        self.h = lib().orc_chol_create_kind(A.n, A.rp, A.ci, kind)

    def factor(self, tol=1e-12, vals=None):
        return lib().orc_chol_factor(self.h, self.A.v if vals is None else vals, tol)

    def solve(self, b):
        x = np.zeros(self.A.n)
        lib().orc_chol_solve(self.h, np.ascontiguousarray(b, np.float64), x)
        return x

    @property
    def nnz(self):
        return lib().orc_chol_nnz(self.h)

    @property
    def flops(self):
        return lib().orc_chol_flops(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_chol_free(self.h)
            self.h = None


def schur_cg(J: Csr, Jt: Csr, chol: Chol, b, x0=None, tol=1e-12, itmax=100):
    x = np.zeros(J.n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int()
    fn = C.c_double()
    st = lib().orc_schur_cg(J.n, J.m, J.rp, J.ci, J.v, Jt.rp, Jt.ci, Jt.v, chol.h, x,
                            np.ascontiguousarray(b, np.float64), tol, itmax, C.byref(it),
                            C.byref(fn))
    return x, st, it.value, fn.value


def threads():
    """OpenMP threads the oracle's parallel loops use."""
    return int(lib().orc_threads())


def execute(H: Csr, Ds_vals, Jc: Csr, Jd: Csr, rx, rs, ry, ryd, gamma, perm=None, y0=None,
            ruiz_its=2, cg_tol=1e-12, cg_itmax=100, pivot_tol=1e-12, norm_tol=1e-2,
            intermediates=False, chol_kind=0, refine_max=3, refine_tol=9e-9, delta=0.0):
    """Oracle execute(); returns dict with dx, ds, dyc, dyd, stats (+ intermediates).
    chol_kind 1 selects the supernodal Cholesky (same factor to rounding);
    refine_max 0 stops after the reference's single pass; delta is added to
    the scaled, permuted H_gamma's diagonal before factorization."""
    L = lib()
    n, m, d = H.n, Jc.n, len(Ds_vals)
    o = OrcOpts(gamma, ruiz_its, cg_tol, cg_itmax, pivot_tol, norm_tol, chol_kind, refine_max,
                refine_tol, delta)
    st = OrcStats()
    dx, dsv, dyc, dyd = np.zeros(n), np.zeros(d), np.zeros(m), np.zeros(d)
    pm = None if perm is None else np.ascontiguousarray(perm, np.int32)
    y0a = None if y0 is None else np.ascontiguousarray(y0, np.float64)
    im = None
    bufs = {}
    if intermediates:
        # Upper bounds for the pattern sizes: ask the oracle for them first.
        bufs = {"max_d": np.zeros(n + m), "rx_hat": np.zeros(n), "schur_rhs": np.zeros(m)}
        hmax = H.nnz + 64 * n + Jd.nnz * 16 + 1
        bufs["htil_v"] = np.zeros(hmax)
        bufs["hgam_v"] = np.zeros(hmax + 64 * Jc.nnz)
        im = OrcIntermediates(*(_ptr(bufs[k]) for k in
                                ("htil_v", "hgam_v", "max_d", "rx_hat", "schur_rhs")))
    nnz_jd = Jd.nnz if Jd is not None else 0
    rc = L.orc_execute(C.byref(o), n, H.rp, H.ci, H.v, d, np.ascontiguousarray(Ds_vals, np.float64),
                       m, Jc.rp, Jc.ci, Jc.v,
                       _ptr(Jd.rp), _ptr(Jd.ci), _ptr(Jd.v), nnz_jd,
                       np.ascontiguousarray(rx, np.float64), np.ascontiguousarray(rs, np.float64),
                       np.ascontiguousarray(ry, np.float64), np.ascontiguousarray(ryd, np.float64),
                       _ptr(y0a), _ptr(pm), dx, dsv, dyc, dyd, C.byref(st),
                       C.byref(im) if im is not None else None)
    out = {"status": rc, "dx": dx, "ds": dsv, "dyc": dyc, "dyd": dyd,
           "stats": {f: getattr(st, f) for f, _ in OrcStats._fields_}}
    if intermediates:
        out["htil_v"] = bufs["htil_v"][: st.nnz_htil].copy()
        out["hgam_v"] = bufs["hgam_v"][: st.nnz_hgam].copy()
        for k in ("max_d", "rx_hat", "schur_rhs"):
            out[k] = bufs[k]
    return out


def kkt_residual(H, Ds_vals, Jc, Jd, rx, rs, ry, ryd, dx, dsv, dyc, dyd):
    r2, r4 = C.c_double(), C.c_double()
    f = lambda a: np.ascontiguousarray(a, np.float64)
    lib().orc_kkt_residual(H.n, H.rp, H.ci, H.v, len(Ds_vals), f(Ds_vals), Jc.n, Jc.rp, Jc.ci,
                           Jc.v, _ptr(Jd.rp), _ptr(Jd.ci), _ptr(Jd.v), Jd.nnz,
                           f(rx), f(rs), f(ry), f(ryd), f(dx), f(dsv), f(dyc), f(dyd),
                           C.byref(r2), C.byref(r4))
    return r2.value, r4.value


# ------------------------------------------------------- reference host code
def ref_lib():
    """The reference's own input_functions/MMatrix host code, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        R = C.CDLL(REF_SO)
        R.ref_read_csr.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.c_int, _i, _i, _d]
        R.ref_read_rhs.argtypes = [C.c_char_p, _d]
        _ref = R
    return _ref
