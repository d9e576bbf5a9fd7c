"""MI355X-native HyKKT: the KKT solve of ORNL/hykkt on gfx950.

The compute path is libhykkt.so (HIP kernels + C ABI declared in
include/hykkt.h).  This module binds it with ctypes and mirrors the
reference's HykktSolver interface (src/HykktSolver.hpp:17-137):

    s = HykktSolver(gamma)
    s.set_H(rowptrs, colidx, vals, n, m, nnz); s.set_Ds(...); s.set_Jc(...); s.set_Jd(...)
    s.set_rx(v, n); s.set_rs(...); s.set_ry(...); s.set_ryd(...)
    status = s.execute()          # 0 ok, 1 residual/pattern failure (as the reference)

Arrays passed to the setters are borrowed, as in the reference: they must stay
alive and may be overwritten in place between execute() calls.  There is no
CPU fallback: importing succeeds without the library, but every call that
needs it raises if libhykkt.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhykkt.so")
_lib = None

HYKKT_OK = 0
HYKKT_RESIDUAL_FAIL = 1
ERRORS = {-1: "E_ARG", -2: "E_STATE", -3: "E_HIP", -4: "E_PATTERN", -5: "E_PIVOT",
          -6: "E_IO", -7: "E_NOMEM"}


class HykktError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("ruiz_its", C.c_int), ("norm_tol", C.c_double), ("pivot_tol", C.c_double),
                ("cg_tol", C.c_double), ("cg_itmax", C.c_int), ("delta", C.c_double),
                ("inputs_on_device", C.c_int), ("cg_check_every", C.c_int),
                ("verbose", C.c_int), ("refine_max", C.c_int), ("refine_tol", C.c_double),
                ("autotune", C.c_int), ("sweep_share", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("resid2", C.c_double), ("resid4", C.c_double), ("cg_final_norm", C.c_double),
                ("cg_iters", C.c_int), ("cg_status", C.c_int), ("pivot", C.c_int),
                ("n", C.c_int), ("m", C.c_int), ("d", C.c_int), ("nnz_htil", C.c_int),
                ("nnz_hgam", C.c_int), ("nnz_l", C.c_longlong), ("n_supernodes", C.c_int),
                ("n_levels", C.c_int), ("max_front", C.c_int), ("factor_flops", C.c_double),
                ("t_symbolic_ms", C.c_double), ("t_solve_ms", C.c_double),
                ("t_factor_ms", C.c_double), ("t_cg_ms", C.c_double), ("n_solves", C.c_int),
                ("resid4_first", C.c_double), ("refine_steps", C.c_int),
                ("t_refine_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("short_max", C.c_int), ("group_max", C.c_int), ("pair_min", C.c_int),
                ("count", C.c_int), ("candidates", C.c_int), ("factor_ms", C.c_double),
                ("default_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_longlong), ("ms", C.c_double),
                ("flops", C.c_double), ("bytes", C.c_double)]


_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

_SIGS = {
    "hykkt_version": ([], C.c_char_p),
    "hykkt_last_error": ([], C.c_char_p),
    "hykkt_create": ([C.c_double, C.POINTER(_vp)], C.c_int),
    "hykkt_destroy": ([_vp], C.c_int),
    "hykkt_set_H": ([_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "hykkt_set_Ds": ([_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "hykkt_set_Jc": ([_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "hykkt_set_Jd": ([_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "hykkt_set_rx": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_rs": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_ry": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_ryd": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_x_host": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_s_host": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_y_host": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_yd_host": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_gamma": ([_vp, C.c_double], C.c_int),
    "hykkt_read_matrix_files": ([_vp] + [C.c_char_p] * 8 + [C.c_int], C.c_int),
    "hykkt_execute": ([_vp], C.c_int),
    "hykkt_default_options": ([C.POINTER(Options)], C.c_int),
    "hykkt_set_options": ([_vp, C.POINTER(Options)], C.c_int),
    "hykkt_get_options": ([_vp, C.POINTER(Options)], C.c_int),
    "hykkt_get_stats": ([_vp, C.POINTER(Stats)], C.c_int),
    "hykkt_get_tuning": ([_vp, C.POINTER(Tuning)], C.c_int),
    "hykkt_get_solution": ([_vp, _vp, _vp, _vp, _vp], C.c_int),
    "hykkt_get_solution_device": ([_vp, _vp, _vp, _vp, _vp], C.c_int),
    "hykkt_create_batch": ([C.c_double, C.c_int, C.POINTER(_vp)], C.c_int),
    "hykkt_batch_capacity": ([_vp, _ip], C.c_int),
    "hykkt_bind_system": ([_vp, C.c_int, C.c_double] + [_vp] * 9, C.c_int),
    "hykkt_execute_batch": ([_vp, C.c_int, _vp], C.c_int),
    "hykkt_get_stats_k": ([_vp, C.c_int, C.POINTER(Stats)], C.c_int),
    "hykkt_get_solution_k": ([_vp, C.c_int, _vp, _vp, _vp, _vp], C.c_int),
    "hykkt_get_solution_device_k": ([_vp, C.c_int, _vp, _vp, _vp, _vp], C.c_int),
    "hykkt_get_permutation": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_set_permutation": ([_vp, _vp, C.c_int], C.c_int),
    "hykkt_debug_copy": ([_vp, C.c_int, _vp, C.c_longlong], C.c_longlong),
    "hykkt_device_malloc": ([C.POINTER(_vp), C.c_ulonglong], C.c_int),
    "hykkt_device_free": ([_vp], C.c_int),
    "hykkt_memcpy_htod": ([_vp, _vp, C.c_ulonglong], C.c_int),
    "hykkt_memcpy_dtoh": ([_vp, _vp, C.c_ulonglong], C.c_int),
    "hykkt_device_synchronize": ([], C.c_int),
    "hykkt_device_copy_f64": ([_vp, _vp, C.c_longlong], C.c_int),
    "hykkt_device_mem_info": ([C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)], C.c_int),
    "hykkt_mm_read_csr": ([C.c_char_p, C.c_int, C.c_int, _ip, _ip, _ip, _vp, _vp, _vp], C.c_int),
    "hykkt_mm_read_rhs": ([C.c_char_p, _ip, _vp], C.c_int),
    "hykkt_profile_enable": ([_vp, C.c_int], C.c_int),
    "hykkt_profile_get": ([_vp, C.POINTER(KernelTime), C.c_int, _ip], C.c_int),
    "hykkt_ruiz_scale": ([C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp,
                          _vp, _vp, _vp, C.c_int, _vp, _vp, _vp], C.c_int),
    "hykkt_ruiz_scale_noalloc": ([C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp,
                                  _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp], C.c_int),
    "hykkt_perm_reverse": ([C.c_int, _vp, _vp], C.c_int),
    "hykkt_perm_map_rc": ([C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "hykkt_perm_map_c": ([C.c_int, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "hykkt_perm_map_r": ([C.c_int, _vp, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "hykkt_amd_order": ([C.c_int, _vp, _vp, _vp], C.c_int),
    "hykkt_chol_analyze_host": ([C.c_int, _vp, _vp, C.POINTER(C.c_longlong), _dp, _ip, _ip,
                                 _ip], C.c_int),
    "hykkt_chol_create": ([C.c_int, C.c_int, _vp, _vp, C.POINTER(_vp)], C.c_int),
    "hykkt_chol_factor": ([_vp, _vp, C.c_double, _ip], C.c_int),
    "hykkt_chol_solve": ([_vp, _vp, _vp], C.c_int),
    "hykkt_chol_info": ([_vp, C.POINTER(C.c_longlong), _ip, _ip, _dp], C.c_int),
    "hykkt_chol_destroy": ([_vp], C.c_int),
    "hykkt_schur_cg": ([_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int,
                        _ip, _dp], C.c_int),
}


def lib():
    """The loaded libhykkt.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hykkt_amd.build`"
                               " (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def _check(rc, allow=(HYKKT_OK,)):
    if rc in allow:
        return rc
    raise HykktError(rc, lib().hykkt_last_error().decode())


def _p(a):
    """Pointer to a numpy array's data, a raw integer address, or None."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class HykktSolver:
    """Mirror of the reference's HykktSolver (src/HykktSolver.hpp:17-137).

    Setter arguments are numpy arrays (host) or, with
    ``options(inputs_on_device=1)``, DeviceArray objects / raw device addresses.
    The solver keeps references to the arrays it was given (they are
    borrowed, HykktSolver.cpp:932-959)."""

    def __init__(self, gamma: float):
        self._h = C.c_void_p()
        _check(lib().hykkt_create(float(gamma), C.byref(self._h)))
        self._keep = {}
        self._on_dev = False   # options().inputs_on_device, kept by options()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hykkt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- matrix blocks -------------------------------------------------------
    def _set_block(self, fn, key, rowptrs, colidx, vals, n, m, nnz):
        if not self._on_dev:
            rowptrs, colidx, vals = _i32(rowptrs), _i32(colidx), _f64(vals)
        self._keep[key] = (rowptrs, colidx, vals)
        _check(fn(self._h, _p(rowptrs), _p(colidx), _p(vals), int(n), int(m), int(nnz)))

    def set_H(self, rowptrs, colidx, vals, n, m, nnz):
        self._set_block(lib().hykkt_set_H, "H", rowptrs, colidx, vals, n, m, nnz)

    def set_Ds(self, rowptrs, colidx, vals, n, m, nnz):
        self._set_block(lib().hykkt_set_Ds, "Ds", rowptrs, colidx, vals, n, m, nnz)

    def set_Jc(self, rowptrs, colidx, vals, n, m, nnz):
        self._set_block(lib().hykkt_set_Jc, "Jc", rowptrs, colidx, vals, n, m, nnz)

    def set_Jd(self, rowptrs, colidx, vals, n, m, nnz):
        self._set_block(lib().hykkt_set_Jd, "Jd", rowptrs, colidx, vals, n, m, nnz)

    # -- vectors -------------------------------------------------------------
    def _set_vec(self, fn, key, vals, n):
        if not self._on_dev and not isinstance(vals, int):
            vals = _f64(vals)
        self._keep[key] = vals
        _check(fn(self._h, _p(vals), int(n)))

    def set_rx(self, vals, n):
        self._set_vec(lib().hykkt_set_rx, "rx", vals, n)

    def set_rs(self, vals, n):
        self._set_vec(lib().hykkt_set_rs, "rs", vals, n)

    def set_ry(self, vals, n):
        self._set_vec(lib().hykkt_set_ry, "ry", vals, n)

    def set_ryd(self, vals, n):
        self._set_vec(lib().hykkt_set_ryd, "ryd", vals, n)

    def set_x_host(self, vals, n):
        self._set_vec(lib().hykkt_set_x_host, "x", vals, n)

    def set_s_host(self, vals, n):
        self._set_vec(lib().hykkt_set_s_host, "s", vals, n)

    def set_y_host(self, vals, n):
        self._set_vec(lib().hykkt_set_y_host, "y", vals, n)

    def set_yd_host(self, vals, n):
        self._set_vec(lib().hykkt_set_yd_host, "yd", vals, n)

    def set_gamma(self, gamma):
        _check(lib().hykkt_set_gamma(self._h, float(gamma)))

    def read_matrix_files(self, h_file, ds_file, jc_file, jd_file, rx_file, rs_file, ry_file,
                          ryd_file, skip_lines):
        files = [f.encode() for f in (h_file, ds_file, jc_file, jd_file, rx_file, rs_file,
                                      ry_file, ryd_file)]
        _check(lib().hykkt_read_matrix_files(self._h, *files, int(skip_lines)))

    def execute(self) -> int:
        return _check(lib().hykkt_execute(self._h), allow=(HYKKT_OK, HYKKT_RESIDUAL_FAIL))

    # -- additions -----------------------------------------------------------
    def options(self, **kw) -> Options:
        o = Options()
        _check(lib().hykkt_get_options(self._h, C.byref(o)))
        if kw:
            for k, v in kw.items():
                setattr(o, k, v)
            _check(lib().hykkt_set_options(self._h, C.byref(o)))
        self._on_dev = bool(o.inputs_on_device)
        return o

    def stats(self) -> dict:
        st = Stats()
        _check(lib().hykkt_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def tuning(self) -> dict:
        """The launch shapes the autotuner chose for the current pattern."""
        t = Tuning()
        _check(lib().hykkt_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def profile(self, on=True):
        """Start (and clear) or stop per-kernel timing with HIP events."""
        _check(lib().hykkt_profile_enable(self._h, int(bool(on))))

    def kernel_times(self) -> dict:
        """{kernel: {launches, ms, flops, bytes}} summed since profile(True)."""
        buf = (KernelTime * 16)()
        cnt = C.c_int()
        _check(lib().hykkt_profile_get(self._h, buf, 16, C.byref(cnt)))
        return {buf[k].name.decode(): {"launches": buf[k].launches, "ms": buf[k].ms,
                                       "flops": buf[k].flops, "bytes": buf[k].bytes}
                for k in range(min(cnt.value, 16))}

    def solution(self):
        st = self.stats()
        n, m, d = st["n"], st["m"], st["d"]
        dx, ds, dyc, dyd = np.zeros(n), np.zeros(d), np.zeros(m), np.zeros(d)
        _check(lib().hykkt_get_solution(self._h, _p(dx), _p(ds), _p(dyc), _p(dyd)))
        return dx, ds, dyc, dyd

    def solution_device(self):
        ptrs = [C.c_void_p() for _ in range(4)]
        _check(lib().hykkt_get_solution_device(self._h, *[C.byref(p) for p in ptrs]))
        return [p.value for p in ptrs]

    def permutation(self):
        n = self.stats()["n"]
        p = np.zeros(n, np.int32)
        _check(lib().hykkt_get_permutation(self._h, _p(p), n))
        return p

    def set_permutation(self, perm):
        perm = _i32(perm)
        _check(lib().hykkt_set_permutation(self._h, _p(perm), len(perm)))

    def debug(self, what, count, dtype=np.float64):
        out = np.zeros(max(int(count), 1), dtype=dtype)
        r = lib().hykkt_debug_copy(self._h, int(what), _p(out), len(out))
        if r < 0:
            raise HykktError(int(r), lib().hykkt_last_error().decode())
        return out[:r]


class BatchSolver(HykktSolver):
    """Batched form (no counterpart in the reference): up to `capacity`
    systems sharing one sparsity pattern, solved together, each with exactly
    the arithmetic it gets alone.  The pattern comes from the set_H / set_Ds /
    set_Jc / set_Jd setters (their values are not used); system k's values
    from bind_system(k, ...)."""

    def __init__(self, gamma: float, capacity: int):
        self._h = C.c_void_p()
        _check(lib().hykkt_create_batch(float(gamma), int(capacity), C.byref(self._h)))
        self._keep = {}
        self._on_dev = False
        self.capacity = int(capacity)

    def bind_system(self, k, gamma, H, Ds, Jc, Jd, rx, rs, ry, ryd, y0=None):
        """Value arrays of system k (numpy arrays, or DeviceArray / device
        addresses with inputs_on_device)."""
        arrs = [H, Ds, Jc, Jd, rx, rs, ry, ryd, y0]
        if not self._on_dev:
            arrs = [None if a is None else _f64(a) for a in arrs]
        self._keep[("sys", int(k))] = arrs
        _check(lib().hykkt_bind_system(self._h, int(k), float(gamma), *[_p(a) for a in arrs]))

    def execute_batch(self, count) -> list:
        st = np.zeros(int(count), np.int32)
        _check(lib().hykkt_execute_batch(self._h, int(count), _p(st)))
        return [int(v) for v in st]

    def stats_k(self, k) -> dict:
        st = Stats()
        _check(lib().hykkt_get_stats_k(self._h, int(k), C.byref(st)))
        return st.as_dict()

    def solution_k(self, k):
        st = self.stats_k(k)
        n, m, d = st["n"], st["m"], st["d"]
        dx, ds, dyc, dyd = np.zeros(n), np.zeros(d), np.zeros(m), np.zeros(d)
        _check(lib().hykkt_get_solution_k(self._h, int(k), _p(dx), _p(ds), _p(dyc), _p(dyd)))
        return dx, ds, dyc, dyd

    def solution_device_k(self, k):
        ptrs = [C.c_void_p() for _ in range(4)]
        _check(lib().hykkt_get_solution_device_k(self._h, int(k), *[C.byref(p) for p in ptrs]))
        return [p.value for p in ptrs]


class DeviceArray:
    """A numpy array copied into device memory (HBM) through the library's own
    HIP runtime; `.ptr` is the device address to pass with inputs_on_device."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.dtype, self.shape, self.nbytes = a.dtype, a.shape, a.nbytes
        p = C.c_void_p()
        _check(lib().hykkt_device_malloc(C.byref(p), max(a.nbytes, 8)))
        self.ptr = p.value
        _check(lib().hykkt_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().hykkt_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().hykkt_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize():
    _check(lib().hykkt_device_synchronize())


def device_mem_info():
    """(free, total) device memory in bytes."""
    f, t = C.c_ulonglong(), C.c_ulonglong()
    _check(lib().hykkt_device_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def device_copy(dst: "DeviceArray", src: "DeviceArray"):
    """dst <- src (float64) with the library's plain streaming copy kernel."""
    n = src.nbytes // 8
    assert dst.nbytes >= src.nbytes and src.dtype == np.float64
    _check(lib().hykkt_device_copy_f64(dst.ptr, src.ptr, n))


def mm_read_csr(path, skip_lines, symmetric):
    """input_functions.cpp ingest: (n, m, rowptr, colidx, vals) of a
    MatrixMarket file (symmetric: lower triangle expanded to both)."""
    n, m, nnz = C.c_int(), C.c_int(), C.c_int()
    p = str(path).encode()
    _check(lib().hykkt_mm_read_csr(p, int(skip_lines), int(bool(symmetric)), C.byref(n),
                                   C.byref(m), C.byref(nnz), None, None, None))
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    v = np.zeros(max(nnz.value, 1))
    _check(lib().hykkt_mm_read_csr(p, int(skip_lines), int(bool(symmetric)), C.byref(n),
                                   C.byref(m), C.byref(nnz), _p(rp), _p(ci), _p(v)))
    return n.value, m.value, rp, ci[:nnz.value], v[:nnz.value]


def mm_read_rhs(path):
    """input_functions.cpp read_rhs."""
    n = C.c_int()
    p = str(path).encode()
    _check(lib().hykkt_mm_read_rhs(p, C.byref(n), None))
    v = np.zeros(max(n.value, 1))
    _check(lib().hykkt_mm_read_rhs(p, C.byref(n), _p(v)))
    return v[:n.value]


# ---------------------------------------------------------------- components
def ruiz_scale(its, H, Jt, J, rhs1, rhs2, noalloc=False):
    """RuizClass (RuizClass.cpp:54-98) on the GPU.  H, Jt, J are (rp, ci, v)
    triples (copied); returns scaled copies (Hv, Jtv, Jv, rhs1, rhs2, max_d).
    noalloc: through hykkt_ruiz_scale_noalloc (no device allocation; the
    library's fixed device scratch)."""
    hrp, hci, hv = _i32(H[0]), _i32(H[1]), _f64(H[2]).copy()
    trp, tci, tv = _i32(Jt[0]), _i32(Jt[1]), _f64(Jt[2]).copy()
    jrp, jci, jv = _i32(J[0]), _i32(J[1]), _f64(J[2]).copy()
    r1, r2 = _f64(rhs1).copy(), _f64(rhs2).copy()
    n, m2 = len(hrp) - 1, len(jrp) - 1
    md = np.zeros(n + m2)
    fn = lib().hykkt_ruiz_scale_noalloc if noalloc else lib().hykkt_ruiz_scale
    _check(fn(int(its), n, m2, _p(hrp), _p(hci), _p(hv), len(hv), _p(trp), _p(tci), _p(tv),
              _p(jrp), _p(jci), _p(jv), len(jv), _p(r1), _p(r2), _p(md)))
    return hv, tv, jv, r1, r2, md


def amd_order(rp, ci):
    rp, ci = _i32(rp), _i32(ci)
    n = len(rp) - 1
    perm = np.zeros(n, np.int32)
    _check(lib().hykkt_amd_order(n, _p(rp), _p(ci), _p(perm)))
    return perm


def perm_maps(rows, cols, perm, square=True):
    """Reference maps (permcheck.cu:99-195): dict with "rev", "c" and, for a
    square pattern, "rc" and "r"."""
    rows, cols, perm = _i32(rows), _i32(cols), _i32(perm)
    n, nr, nnz = len(perm), len(rows) - 1, int(rows[-1])
    rev = np.zeros(n, np.int32)
    _check(lib().hykkt_perm_reverse(n, _p(perm), _p(rev)))
    out = {"rev": rev}
    pc, pm = np.zeros(max(nnz, 1), np.int32), np.zeros(max(nnz, 1), np.int32)
    _check(lib().hykkt_perm_map_c(nr, _p(rows), _p(cols), _p(rev), _p(pc), _p(pm)))
    out["c"] = (pc[:nnz], pm[:nnz])
    if square:
        pr, pc, pm = (np.zeros(n + 1, np.int32), np.zeros(max(nnz, 1), np.int32),
                      np.zeros(max(nnz, 1), np.int32))
        _check(lib().hykkt_perm_map_rc(n, _p(rows), _p(cols), _p(perm), _p(rev), _p(pr), _p(pc),
                                       _p(pm)))
        out["rc"] = (pr, pc[:nnz], pm[:nnz])
        pr, pc, pm = (np.zeros(n + 1, np.int32), np.zeros(max(nnz, 1), np.int32),
                      np.zeros(max(nnz, 1), np.int32))
        _check(lib().hykkt_perm_map_r(n, _p(rows), _p(cols), _p(perm), _p(pr), _p(pc), _p(pm)))
        out["r"] = (pr, pc[:nnz], pm[:nnz])
    return out


def chol_analyze_host(rp, ci):
    rp, ci = _i32(rp), _i32(ci)
    nnz, fl = C.c_longlong(), C.c_double()
    ns, nl, mf = C.c_int(), C.c_int(), C.c_int()
    _check(lib().hykkt_chol_analyze_host(len(rp) - 1, _p(rp), _p(ci), C.byref(nnz), C.byref(fl),
                                         C.byref(ns), C.byref(nl), C.byref(mf)))
    return {"nnz_l": nnz.value, "flops": fl.value, "n_supernodes": ns.value,
            "n_levels": nl.value, "max_front": mf.value}


class Cholesky:
    """CholeskyClass (CholeskyClass.cpp:9-101): lower triangle of a CSR matrix."""

    def __init__(self, rp, ci):
        self.rp, self.ci = _i32(rp), _i32(ci)
        self.n = len(self.rp) - 1
        self._h = C.c_void_p()
        _check(lib().hykkt_chol_create(self.n, int(self.rp[-1]), _p(self.rp), _p(self.ci),
                                       C.byref(self._h)))

    def factor(self, vals, tol=1e-12):
        vals = _f64(vals)
        sing = C.c_int(-1)
        rc = lib().hykkt_chol_factor(self._h, _p(vals), float(tol), C.byref(sing))
        _check(rc, allow=(HYKKT_OK, -5))
        return sing.value

    def solve(self, b):
        b = _f64(b)
        x = np.zeros(self.n)
        _check(lib().hykkt_chol_solve(self._h, _p(b), _p(x)))
        return x

    def info(self):
        nnz, fl = C.c_longlong(), C.c_double()
        ns, nl = C.c_int(), C.c_int()
        _check(lib().hykkt_chol_info(self._h, C.byref(nnz), C.byref(ns), C.byref(nl),
                                     C.byref(fl)))
        return {"nnz_l": nnz.value, "n_supernodes": ns.value, "n_levels": nl.value,
                "flops": fl.value}

    def schur_cg(self, J_rp, J_ci, J_v, b, x0=None, tol=1e-12, itmax=100):
        """SchurComplementConjugateGradient on S = J A^{-1} J^T."""
        J_rp, J_ci, J_v, b = _i32(J_rp), _i32(J_ci), _f64(J_v), _f64(b)
        m = len(J_rp) - 1
        x = np.zeros(m) if x0 is None else _f64(x0).copy()
        it, fn = C.c_int(), C.c_double()
        rc = lib().hykkt_schur_cg(self._h, m, self.n, _p(J_rp), _p(J_ci), _p(J_v), _p(x), _p(b),
                                  float(tol), int(itmax), C.byref(it), C.byref(fn))
        _check(rc, allow=(0, 1))
        return x, rc, it.value, fn.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hykkt_chol_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
