"""Failure paths of the solve (the cases the reference either lets run into
NaN or ignores), on the CPU where the behaviour is host logic or the oracle's,
and through the C ABI on the GPU where it is the product's.

- A Jc row with no nonzero value: the reference's Ruiz sweep scales it by
  1/sqrt(0) (matrix_vector_ops.cu:258-264) and the solve turns to NaN; the
  product reports an argument error naming the row.
- CG reaching itmax: the reference ignores it (HykktSolver.cpp:406,
  SchurComplementConjugateGradient.cpp:189-192); the product reports it in
  its statistics.
- A served pattern whose slot does not fit the HBM budget is refused, and the
  stream is served without it.
- Switching the option inputs_on_device between solves of one analysed
  pattern gives the same bits either way.
"""
import numpy as np
import pytest

import _oracle as O
import hykkt_amd as H
from hykkt_amd.generator import GridPattern


def small_system(seed=3, nbus=30, nbr=39, gamma=1e4):
    g = GridPattern(nbus, nbr, seed=seed, gamma=gamma)
    s = g.system()
    c = lambda b: O.Csr(b.n, b.m, b.rp, b.ci, b.v)
    return s, c(s.H), c(s.Jc), c(s.Jd)


def with_zero_jc_row(Jc, row, structural):
    """Jc with row `row` emptied (structural) or its values set to zero."""
    J = Jc.copy()
    if not structural:
        J.v[J.rp[row]:J.rp[row + 1]] = 0.0
        return J
    keep = np.ones(J.nnz, bool)
    keep[J.rp[row]:J.rp[row + 1]] = False
    cnt = np.diff(J.rp)
    cnt[row] = 0
    return O.Csr(J.n, J.m, np.concatenate([[0], np.cumsum(cnt)]), J.ci[keep], J.v[keep])


# ---------------------------------------------------------------- CPU
@pytest.mark.parametrize("structural", [False, True])
def test_oracle_zero_jc_row_is_the_reference_nan(structural):
    """The oracle restates the reference: the zero row's scale is infinite and
    the solve does not pass its residual test (status 1, non-finite residual)."""
    s, Hc, Jc, Jd = small_system()
    J0 = with_zero_jc_row(Jc, 7, structural)
    out = O.execute(Hc, s.Ds.v, J0, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma, refine_max=0)
    assert out["status"] == 1
    assert not np.isfinite(out["stats"]["resid2"])


def test_oracle_cg_itmax_reported():
    """cg_itmax = 1 on a system that needs more: the oracle's CG stops with
    status 1 after one iteration, and execute() still returns the residual
    test's verdict (the reference ignores the CG status)."""
    s, Hc, Jc, Jd = small_system(gamma=1e2)
    full = O.execute(Hc, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma, refine_max=0)
    assert full["stats"]["cg_status"] == 0 and full["stats"]["cg_iters"] > 1
    out = O.execute(Hc, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma, refine_max=0,
                    cg_itmax=1)
    assert out["stats"]["cg_status"] == 1 and out["stats"]["cg_iters"] == 1
    assert out["stats"]["resid2"] > full["stats"]["resid2"]


class _FakeSolver:
    """Stands in for a pattern's solvers in the admission logic: the batch
    slot's first batch 'takes' the bytes its pattern asks for."""
    mem = {"free": 100e9}

    def __init__(self, gamma, capacity=1):
        self.need, self.taken = 0, 0

    def options(self, **kw):
        return None

    def execute_batch(self, count):
        self.taken = self.need
        _FakeSolver.mem["free"] -= self.need
        return [0] * count

    def execute(self):
        return 0

    def stats(self):
        return {"resid2": 1e-12, "resid4": 1e-12, "pivot": -1}

    def close(self):
        _FakeSolver.mem["free"] += self.taken


class _FakePattern:
    def __init__(self, pid, nbus, need):
        self.pid, self.nbus, self.need = pid, nbus, need
        self.pool = [0, 1]
        self.g = type("G", (), {"gamma": 1e5})()

    def set_pattern(self, s, k=0):
        s.need = self.need

    def bind_system(self, b, slot, k):
        pass


def test_server_refuses_a_slot_over_the_budget(monkeypatch):
    """Admission by HBM budget (largest pattern first): the slot that would
    push the measured footprint over the budget is refused with its size, the
    others are served."""
    from hykkt_amd import served as S
    _FakeSolver.mem["free"] = 100e9
    monkeypatch.setattr(S.H, "HykktSolver", _FakeSolver)
    monkeypatch.setattr(S.H, "BatchSolver", _FakeSolver)
    monkeypatch.setattr(S.H, "device_mem_info", lambda: (_FakeSolver.mem["free"], 288e9))
    monkeypatch.setattr(S.H, "device_synchronize", lambda: None)
    pats = [_FakePattern(0, 1000, 2e9), _FakePattern(1, 20000, 30e9), _FakePattern(2, 5000, 9e9)]
    srv = S.Server(pats, hbm_budget=35e9)
    assert srv.served == [0, 1]                 # 20k (30 GB) then 5k (9 GB) is over: refused
    assert [r["pattern"] for r in srv.refused] == [2]
    assert srv.refused[0]["bytes"] == 9e9
    assert srv.hbm_used == 32e9
    assert _FakeSolver.mem["free"] == 100e9 - 32e9   # the refused slot was released


# ---------------------------------------------------------------- GPU
def bind(s, Hc, Dsv, Jc, Jd, sy):
    d = len(Dsv)
    s.set_H(Hc.rp, Hc.ci, Hc.v, Hc.n, Hc.m, Hc.nnz)
    s.set_Ds(np.arange(d + 1), np.arange(d), Dsv, d, d, d)
    s.set_Jc(Jc.rp, Jc.ci, Jc.v, Jc.n, Jc.m, Jc.nnz)
    s.set_Jd(Jd.rp, Jd.ci, Jd.v, Jd.n, Jd.m, Jd.nnz)
    for name in ("rx", "rs", "ry", "ryd"):
        v = getattr(sy, name)
        getattr(s, "set_" + name)(v, len(v))


@pytest.mark.gpu
@pytest.mark.parametrize("structural", [False, True])
def test_zero_jc_row_is_an_argument_error(structural):
    s, Hc, Jc, Jd = small_system()
    J0 = with_zero_jc_row(Jc, 7, structural)
    solver = H.HykktSolver(s.gamma)
    bind(solver, Hc, s.Ds.v, J0, Jd, s)
    with pytest.raises(H.HykktError) as e:
        solver.execute()
    assert e.value.code == -1 and "row 7 of Jc has no nonzero value" in str(e.value)
    # the same solver solves the intact system afterwards (values only change
    # for the value-zeroed case; the structural case is a new pattern)
    bind(solver, Hc, s.Ds.v, Jc, Jd, s)
    assert solver.execute() == 0 and solver.stats()["resid4"] <= 1e-8


@pytest.mark.gpu
def test_cg_itmax_reported_gpu():
    s, Hc, Jc, Jd = small_system(gamma=1e2)
    solver = H.HykktSolver(s.gamma)
    bind(solver, Hc, s.Ds.v, Jc, Jd, s)
    solver.options(refine_max=0)
    solver.execute()
    full = solver.stats()
    assert full["cg_status"] == 0 and full["cg_iters"] > 1
    solver.options(cg_itmax=1)
    rc = solver.execute()
    st = solver.stats()
    assert st["cg_status"] == 1 and st["cg_iters"] == 1
    assert rc == (0 if st["resid2"] < 1e-2 else 1)   # the residual test decides, as in the reference
    ref = O.execute(Hc, s.Ds.v, Jc, Jd, s.rx, s.rs, s.ry, s.ryd, s.gamma, refine_max=0,
                    cg_itmax=1, perm=solver.permutation())
    assert ref["stats"]["cg_status"] == 1
    x = np.concatenate(solver.solution())
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.gpu
def test_inputs_on_device_switch_both_ways():
    """A solver analysed with host inputs, then given device inputs, then host
    inputs again (no new analysis): every solve returns 0 with the same bits."""
    s, Hc, Jc, Jd = small_system(seed=8)
    solver = H.HykktSolver(s.gamma)
    bind(solver, Hc, s.Ds.v, Jc, Jd, s)
    assert solver.execute() == 0
    want = np.concatenate(solver.solution())
    t_sym = solver.stats()["t_symbolic_ms"]
    D = H.DeviceArray
    i32 = lambda a: D(np.asarray(a, np.int32))
    f64 = lambda a: D(np.asarray(a, np.float64))
    d = len(s.Ds.v)
    keep = []
    for first_dev in (True, False):
        solver.options(inputs_on_device=int(first_dev))
        if first_dev:
            args = [(solver.set_H, Hc), (solver.set_Jc, Jc), (solver.set_Jd, Jd)]
            for f, b in args:
                arrs = (i32(b.rp), i32(b.ci), f64(b.v))
                keep.append(arrs)
                f(*arrs, b.n, b.m, b.nnz)
            arrs = (i32(np.arange(d + 1)), i32(np.arange(d)), f64(s.Ds.v))
            keep.append(arrs)
            solver.set_Ds(*arrs, d, d, d)
            for name in ("rx", "rs", "ry", "ryd"):
                v = f64(getattr(s, name))
                keep.append(v)
                getattr(solver, "set_" + name)(v, len(getattr(s, name)))
        else:
            bind(solver, Hc, s.Ds.v, Jc, Jd, s)
        assert solver.execute() == 0
        assert np.array_equal(np.concatenate(solver.solution()), want)
        assert solver.stats()["t_symbolic_ms"] == t_sym


@pytest.mark.gpu
def test_server_refusal_on_the_device():
    """Two patterns of the same size and a budget of one and a half of the
    first one's measured footprint: the second is refused, the first is
    served, and a request for each is served and rejected respectively."""
    from hykkt_amd.served import Request, Server, ServedPattern
    pats = [ServedPattern(0, 300, seed=21), ServedPattern(1, 300, seed=22)]
    try:
        probe = Server(pats[:1], hbm_budget=64e9)
        one = probe.hbm_used
        probe.close()
        assert one > 0
        srv = Server(pats, hbm_budget=one * 1.5)
        assert srv.served == [0] and [r["pattern"] for r in srv.refused] == [1]
        assert srv.refused[0]["bytes"] > 0.5 * one and srv.hbm_used <= one * 1.5
        res, rej, pend, _ = srv.run([Request(0, 0, 0, 0.0), Request(1, 1, 0, 0.0)])
        assert rej == 1 and pend == 0 and len(res) == 1 and res[0].rc == 0
        srv.close()
    finally:
        for p in pats:
            p.free()
