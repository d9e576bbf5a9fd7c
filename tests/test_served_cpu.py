"""Host logic of the served form (no GPU): the Poisson request stream keeps
the 6:3:1 size mix and its rate, and the summary counts what it should."""
from types import SimpleNamespace

import numpy as np

from hykkt_amd.served import Result, poisson_requests, summarize


def test_poisson_stream_mix_and_rate():
    pats = [SimpleNamespace(pid=i, nbus=n) for i, n in
            enumerate([1000] * 10 + [5000] * 5 + [20000])]
    reqs = poisson_requests(pats, rate=2000.0, seconds=5.0, seed=1)
    t = np.array([r.t_arrive for r in reqs])
    assert np.all(np.diff(t) > 0) and t[-1] < 5.0
    assert abs(len(reqs) / 5.0 - 2000.0) < 4 * np.sqrt(2000.0 / 5.0)   # Poisson count
    sizes = np.array([pats[r.pattern].nbus for r in reqs])
    for n, share in ((1000, 0.6), (5000, 0.3), (20000, 0.1)):
        assert abs(np.mean(sizes == n) - share) < 0.02, n
    assert set(r.k for r in reqs) == {0, 1}


def test_summary_counts():
    res = [Result(0, 0, 0.010, 0, 1e-10), Result(1, 0, 0.020, 0, 2e-8), Result(2, 1, 0.030, 1, 1e-9)]
    out = summarize(res, rejected=2, pending=3, wall=2.0)
    assert out["completed"] == 3 and out["within_tol"] == 1
    assert out["rejected"] == 2 and out["pending"] == 3 and out["solves_per_s"] == 0.5
    assert abs(out["p50_ms"] - 20.0) < 1e-9 and out["max_resid4"] == 2e-8
