"""Served form on the GPU: requests of several resident patterns run
concurrently on their solver slots give the same results as one at a time
(plain), every result reaches resid4 <= 1e-8, one small request agrees with
the CPU oracle, and a short Poisson stream is served within the HBM budget.
Needs an MI355X."""
import numpy as np
import pytest

import _oracle as O
import hykkt_amd as H
from hykkt_amd.served import (Request, Server, ServedPattern, poisson_requests, summarize)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served():
    pats = [ServedPattern(0, 500, seed=11), ServedPattern(1, 1000, seed=12),
            ServedPattern(2, 500, seed=13), ServedPattern(3, 2000, seed=14)]
    srv = Server(pats, hbm_budget=64e9, keep_solutions=True)
    yield pats, srv
    srv.close()
    for p in pats:
        p.free()


def test_results_same_alone_and_in_the_mix(served):
    pats, srv = served
    assert srv.served == [0, 1, 2, 3] and srv.hbm_used > 0
    reqs = [Request(i, i % 4, (i // 4) % 2, 0.0) for i in range(12)]
    alone, rej0, pend0, _ = srv.run(reqs, plain=True)
    mixed, rej1, pend1, _ = srv.run(reqs)
    assert rej0 == rej1 == pend0 == pend1 == 0
    alone = {r.rid: r for r in alone}
    mixed = {r.rid: r for r in mixed}
    assert sorted(alone) == sorted(mixed) == list(range(12))
    for rid in range(12):
        a, m = alone[rid], mixed[rid]
        assert a.rc == m.rc == 0 and a.resid4 <= 1e-8 and m.resid4 <= 1e-8, (rid, a, m)
        for xa, xm in zip(a.solution, m.solution):
            assert np.array_equal(xa, xm), rid   # bit for bit: own buffers, deterministic kernels


def test_served_result_matches_oracle(served):
    pats, srv = served
    p, k = pats[0], 1
    res, _, _, _ = srv.run([Request(0, 0, k, 0.0)])
    sy = p.host[k]
    c = lambda b: O.Csr(b.n, b.m, b.rp, b.ci, b.v)
    ref = O.execute(c(sy.H), sy.Ds.v, c(sy.Jc), c(sy.Jd), sy.rx, sy.rs, sy.ry, sy.ryd, sy.gamma,
                    perm=srv.slots[0].permutation())
    mine = np.concatenate(res[0].solution)
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    err = np.linalg.norm(mine - xo) / np.linalg.norm(xo)
    print(f"served request vs oracle: rel. difference {err:.3e}", flush=True)
    assert err < 1e-8


def test_poisson_stream(served):
    pats, srv = served
    reqs = poisson_requests(pats, rate=200.0, seconds=1.0, seed=3)
    out = summarize(*srv.run(reqs))
    print("served stream:", out, flush=True)
    assert out["rejected"] == 0 and out["pending"] == 0
    assert out["within_tol"] == out["completed"] == len(reqs)
