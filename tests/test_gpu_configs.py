"""The batched and served workloads at their stated sizes (needs an MI355X).

Batched: 256 systems, 32 patterns of 200..5000 buses x 8 value draws, a gamma
per system.  Every system that does not fail at a non-positive pivot comes
out bit for bit the same batched and one at a time through the drop-in path,
within resid4 <= 1e-8; one system of every size decile agrees with the CPU
oracle.

Served: 16 resident patterns of 1k/5k/20k buses (6:3:1) with pools of 20
fresh value sets, a Poisson stream at 2,000 requests/s over a short window.
Every result is within resid4 <= 1e-8; a sample of requests gives the same
bits alone (drop-in path) and in the mix; one request of each size agrees
with the CPU oracle.
"""
import numpy as np
import pytest

import _oracle as O
from hykkt_amd.batched import make_batch, make_server, run_batch, summarize_batch
from hykkt_amd.served import Request, Server, make_patterns, poisson_requests, summarize

pytestmark = pytest.mark.gpu


def _csr(b):
    return O.Csr(b.n, b.m, b.rp, b.ci, b.v)


def _vs_oracle(sy, perm, sol):
    ref = O.execute(_csr(sy.H), sy.Ds.v, _csr(sy.Jc), _csr(sy.Jd), sy.rx, sy.rs, sy.ry, sy.ryd,
                    sy.gamma, perm=perm, chol_kind=1)
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    return float(np.linalg.norm(np.concatenate(sol) - xo) / np.linalg.norm(xo))


def test_batched_256_systems():
    pats = make_batch(seed=3)
    assert len(pats) == 32 and sum(len(p.pool) for p in pats) == 256
    srv = make_server(pats, hbm_budget=200e9, keep_solutions=True)
    try:
        assert srv.served == [p.pid for p in pats], srv.refused
        many, wall_b = run_batch(srv, pats)
        one, wall_p = run_batch(srv, pats, plain=True)
        sb, sp = summarize_batch(many, wall_b, pats), summarize_batch(one, wall_p, pats)
        print(f"batched {sb['systems_per_s']:.0f} systems/s, plain {sp['systems_per_s']:.0f}; "
              f"failed {sb['failed']} (non-positive pivot {sb['non_positive_pivot']})", flush=True)
        byp = {(r.pattern, r.k): r for r in one}
        ok = 0
        for r in many:
            a = byp[(r.pattern, r.k)]
            assert r.rc == a.rc and r.pivot == a.pivot, (r, a)
            if r.pivot >= 0:
                assert r.rc == 1
                continue
            assert r.rc == 0 and r.resid4 <= 1e-8, r
            for xa, xb in zip(r.solution, a.solution):
                assert np.array_equal(xa, xb), (r.pattern, r.k)
            ok += 1
        assert ok >= 240 and sb["non_positive_pivot"] == sp["non_positive_pivot"]
        # one solved system per size decile against the CPU oracle
        dec = np.quantile([p.nbus for p in pats], np.linspace(0, 1, 11))
        done = set()
        for r in sorted(many, key=lambda x: (x.pattern, x.k)):
            p = srv.patterns[r.pattern]
            q = int(np.searchsorted(dec[1:-1], p.nbus, side="right"))
            if q in done or r.pivot >= 0:
                continue
            err = _vs_oracle(p.host[r.k], srv.slots[r.pattern].permutation(), r.solution)
            print(f"decile {q}: {p.nbus} buses, draw {r.k}: rel. difference {err:.2e}", flush=True)
            assert err < 1e-8
            done.add(q)
        assert len(done) >= 9
    finally:
        srv.close()
        for p in pats:
            p.free()


def test_served_16_patterns_short_window():
    pats = make_patterns(16, seed=100)
    assert sorted({p.nbus for p in pats}) == [1000, 5000, 20000]
    srv = Server(pats, hbm_budget=200e9, keep_solutions=True)
    try:
        assert srv.served == [p.pid for p in pats], srv.refused
        reqs = poisson_requests(pats, 2000.0, 1.5, seed=7, served=srv.served)
        res, rej, pend, wall = srv.run(reqs, deadline=6.0)
        out = summarize(res, rej, pend, wall, patterns=pats)
        print("served:", {k: out[k] for k in ("completed", "within_tol", "solves_per_s",
                                               "p50_ms", "p99_ms", "mean_batch")}, flush=True)
        assert out["failed"] == 0 and out["within_tol"] == out["completed"] > 0
        assert max(r.batch for r in res) > 1   # requests of a pattern shared batches
        # a sample, alone through the drop-in path: the same bits
        by_size = {}
        for r in sorted(res, key=lambda x: x.rid):
            by_size.setdefault(srv.patterns[r.pattern].nbus, []).append(r)
        sample = [rs[i] for rs in by_size.values() for i in (0, len(rs) // 2) if rs]
        alone, _, _, _ = srv.run([Request(r.rid, r.pattern, r.k, 0.0) for r in sample],
                                 plain=True)
        alone = {r.rid: r for r in alone}
        for r in sample:
            for xa, xb in zip(r.solution, alone[r.rid].solution):
                assert np.array_equal(xa, xb), (r.rid, r.batch)
        # one request of each size against the CPU oracle
        for nbus, rs in sorted(by_size.items()):
            r = rs[0]
            p = srv.patterns[r.pattern]
            err = _vs_oracle(p.host[r.k], srv.slots[r.pattern].permutation(), r.solution)
            print(f"{nbus} buses, request {r.rid} (batch of {r.batch}): rel. difference "
                  f"{err:.2e}", flush=True)
            assert err < 1e-8
    finally:
        srv.close()
        for p in pats:
            p.free()
