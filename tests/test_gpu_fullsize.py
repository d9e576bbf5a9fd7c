"""Parity at the 10,000-bus size (n=40,000, m=20,000, d=26,000; gamma=1e5)
against the CPU oracle, whose supernodal Cholesky finishes one solve in a few
seconds on the host cores.  Needs an MI355X.

Per step: formation (H~ before Ruiz, max_d, H_gamma after Ruiz, r^x) is bit
for bit the oracle's; the GPU solution reaches resid4 <= 1e-8; and its
relative difference from the oracle's solution (same ordering, another
factorization order) is bounded by 1e-6.
"""
import numpy as np
import pytest

import _oracle as O
import fixtures as F
import hykkt_amd as H
from hykkt_amd.generator import ds_spread_schedule, workload

pytestmark = pytest.mark.gpu


def _csr(b):
    return O.Csr(b.n, b.m, b.rp, b.ci, b.v)


@pytest.fixture(scope="module")
def seq():
    g = workload("grid10k_seq20")
    s = H.HykktSolver(g.gamma)
    return g, s


def _bind(s, sy):
    d = len(sy.Ds.v)
    s.set_H(sy.H.rp, sy.H.ci, sy.H.v, sy.H.n, sy.H.m, sy.H.nnz)
    s.set_Ds(np.arange(d + 1), np.arange(d), sy.Ds.v, d, d, d)
    s.set_Jc(sy.Jc.rp, sy.Jc.ci, sy.Jc.v, sy.Jc.n, sy.Jc.m, sy.Jc.nnz)
    s.set_Jd(sy.Jd.rp, sy.Jd.ci, sy.Jd.v, sy.Jd.n, sy.Jd.m, sy.Jd.nnz)
    for name in ("rx", "rs", "ry", "ryd"):
        v = getattr(sy, name)
        getattr(s, "set_" + name)(v, len(v))


@pytest.mark.parametrize("step", [0, 19])
def test_grid10k_step_vs_oracle(seq, step):
    g, s = seq
    sy = g.system(step, ds_range=ds_spread_schedule(step, 20))
    _bind(s, sy)
    status = s.execute()
    st = s.stats()
    print(f"grid10k step {step}: status={status} resid4={st['resid4']:.3e} "
          f"cg={st['cg_iters']} factor={st['t_factor_ms']:.1f}ms", flush=True)
    assert status == 0 and st["pivot"] == -1 and st["cg_status"] == 0
    assert st["resid4"] <= 1e-8
    ref = O.execute(_csr(sy.H), sy.Ds.v, _csr(sy.Jc), _csr(sy.Jd), sy.rx, sy.rs, sy.ry, sy.ryd,
                    g.gamma, perm=s.permutation(), intermediates=True, chol_kind=1)
    assert ref["status"] == 0
    assert st["nnz_htil"] == ref["stats"]["nnz_htil"] == 576000
    assert st["nnz_hgam"] == ref["stats"]["nnz_hgam"]
    assert np.array_equal(s.debug(11, st["nnz_htil"]), ref["htil_v"])
    assert np.array_equal(s.debug(4, st["n"] + st["m"]), ref["max_d"])
    # H_gamma as the factorization received it, read back from the factor storage
    want = F.hgamma_in_factor_order(s, st["nnz_hgam"], st["n"], ref["hgam_v"])
    assert F.same_entries(s.debug(10, st["nnz_hgam"]), want)
    assert np.array_equal(s.debug(12, st["n"]), ref["rx_hat"])
    mine = s.solution()
    x = np.concatenate(mine)
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"grid10k step {step}: rel. difference from the oracle {err:.3e} "
          f"(oracle resid4 {ref['stats']['resid4']:.3e})", flush=True)
    assert err < 1e-6
    r2, r4 = O.kkt_residual(_csr(sy.H), sy.Ds.v, _csr(sy.Jc), _csr(sy.Jd), sy.rx, sy.rs, sy.ry,
                            sy.ryd, *mine)
    assert r4 <= 1e-8 and abs(r4 - st["resid4"]) <= 0.05 * r4 + 1e-14


def test_kernel_timing_accounting(seq):
    """Per-kernel event timing: every instrumented kernel ran, the two solve
    sweeps ran equally often, and the algorithmic flops charged to the
    factorization kernels add up to the factor's flop count (the stored
    supernodes hold a few amalgamation zeros, so slightly more)."""
    g, s = seq
    sy = g.system(1, ds_range=ds_spread_schedule(1, 20))
    _bind(s, sy)
    s.profile(True)
    assert s.execute() == 0
    kt = s.kernel_times()
    s.profile(False)
    st = s.stats()
    for k, v in kt.items():
        assert v["launches"] > 0 and v["ms"] > 0, k
    # small-supernode forest, top of the tree
    for fwd, bwd in (("k_fwd_forest", "k_solve_bwd<256, false>"),
                     ("k_solve_fwd<1024, true>", "k_solve_bwd<1024, true>")):
        assert kt[fwd]["launches"] == kt[bwd]["launches"] >= 3
    # the Cholesky's flops (the inverses formed after a level for the solves are extra)
    fl = sum(kt[k]["flops"] for k in ("k_dense_nt<128>", "k_dense_nt<32>", "k_potrf_inv",
                                      "k_trsm<128>", "k_trsm<32>") if k in kt)
    assert 0.999 <= fl / st["factor_flops"] <= 1.2, fl / st["factor_flops"]
    # formation: one H~ and one H_gamma value launch; the SpMVs: J_d^T once
    # per pass, J_c^T and J_c around the CG and twice per Schur apply
    assert kt["k_spgemm_values"]["launches"] == kt["k_spgemm_scatter"]["launches"] == 1
    passes = 1 + st["refine_steps"]
    assert kt["k_spmv"]["launches"] >= passes * 4 + 2 * st["cg_iters"]
    assert kt["k_spgemm_values"]["bytes"] > 12 * st["nnz_htil"]
    assert kt["k_spgemm_scatter"]["bytes"] > 20 * st["nnz_hgam"] / 2
    # event-timed kernels fit inside the solve's own wall span: each kernel's
    # sum, and the total, which may count the look-ahead stream's updates
    # twice over the time they overlap the main stream
    for k, v in kt.items():
        assert v["ms"] <= st["t_solve_ms"] * 1.01, k
    assert sum(v["ms"] for v in kt.values()) <= 2.0 * st["t_solve_ms"]


def test_grouped_updates_bit_identical(seq, monkeypatch):
    """Grouped trailing updates (consecutive block steps sharing one update
    with K = G x 128) give every factor entry the same FMAs in the same order
    as one step at a time, so the solution is bit-identical; the grouping is
    in effect on this system (fewer dense launches)."""
    g, _ = seq
    sy = g.system(19, ds_range=ds_spread_schedule(19, 20))
    out, launches = {}, {}
    for gmax in ("1", "default"):
        # the variable fixes the launch shapes (no autotuning)
        monkeypatch.setenv("HYKKT_GROUP_MAX", "1" if gmax == "1" else "8")
        s = H.HykktSolver(g.gamma)
        try:
            _bind(s, sy)
            s.profile(True)
            assert s.execute() == 0 and s.stats()["resid4"] <= 1e-8
            kt = s.kernel_times()
            launches[gmax] = sum(kt[k]["launches"] for k in ("k_dense_nt<128>", "k_dense_nt<32>") if k in kt)
            s.profile(False)
            out[gmax] = [np.copy(a) for a in s.solution()]
        finally:
            s.close()
    print(f"grouped updates: k_dense_nt launches {launches['1']} one step at a time, "
          f"{launches['default']} grouped", flush=True)
    assert launches["default"] < launches["1"]
    for a, b in zip(out["1"], out["default"]):
        assert np.array_equal(a, b)


def test_autotuned_shapes_bit_identical(seq, monkeypatch):
    """The autotuner times the factorization under its candidate launch
    shapes at the first solve of the pattern and keeps the fastest; the
    solution has the same bits as with the default shapes, and the tuning
    runs are left out of the kernel profile and the stats' times."""
    g, _ = seq
    monkeypatch.delenv("HYKKT_GROUP_MAX", raising=False)
    monkeypatch.delenv("HYKKT_PAIR_MIN", raising=False)
    sy = g.system(19, ds_range=ds_spread_schedule(19, 20))
    out, kt, st, tun = {}, {}, {}, {}
    for at in (0, 1):
        s = H.HykktSolver(g.gamma)
        try:
            s.options(autotune=at)
            _bind(s, sy)
            s.profile(True)
            assert s.execute() == 0
            kt[at] = s.kernel_times()
            st[at] = s.stats()
            tun[at] = s.tuning()
            s.profile(False)
            out[at] = [np.copy(a) for a in s.solution()]
        finally:
            s.close()
    t = tun[1]
    print(f"autotune: {t}; factor {st[0]['t_factor_ms']:.2f} ms default, "
          f"{st[1]['t_factor_ms']:.2f} ms tuned", flush=True)
    assert tun[0]["candidates"] == 0
    assert t["candidates"] == 7 and t["count"] == 1
    assert 0 < t["factor_ms"] <= t["default_ms"]
    # one factorization's launches in the profile, not eight
    assert kt[1]["k_potrf_inv"]["launches"] == kt[0]["k_potrf_inv"]["launches"]
    assert st[1]["t_solve_ms"] < 3 * st[0]["t_solve_ms"]
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
