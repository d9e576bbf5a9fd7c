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
    # a pattern's i-th request is its interior-point step i: pool entry i mod 20
    for pid in range(len(pats)):
        ks = [r.k for r in reqs if r.pattern == pid]
        assert ks == [i % 20 for i in range(len(ks))], pid


def test_summary_counts():
    res = [Result(0, 0, 0.010, 0, 1e-10), Result(1, 0, 0.020, 0, 2e-8), Result(2, 1, 0.030, 1, 1e-9)]
    out = summarize(res, rejected=2, pending=3, wall=2.0)
    assert out["completed"] == 3 and out["within_tol"] == 1
    assert out["rejected"] == 2 and out["pending"] == 3 and out["solves_per_s"] == 0.5
    assert abs(out["p50_ms"] - 20.0) < 1e-9 and out["max_resid4"] == 2e-8


class _FakeSolver:
    """Stands in for a pattern's drop-in solver: draw 1 of a pattern raises a
    library error, draw 2 meets a non-positive pivot, others solve."""

    def __init__(self):
        self.k = None

    def execute(self):
        import hykkt_amd as H
        if self.k == 1:
            raise H.HykktError(-5, "planted failure")
        return 1 if self.k == 2 else 0

    def stats(self):
        return {"resid4": float("nan") if self.k == 2 else 1e-10, "pivot": 7 if self.k == 2 else -1}


class _FakeBatch:
    """Stands in for a batch slot: per system, draw 1 ends with an error code
    of its own, draw 2 meets a non-positive pivot, others solve."""

    def __init__(self):
        self.ks = {}

    def execute_batch(self, count):
        return [(-5 if self.ks[j] == 1 else 1 if self.ks[j] == 2 else 0) for j in range(count)]

    def stats_k(self, j):
        k = self.ks[j]
        return {"resid4": float("nan") if k == 2 else 1e-10, "pivot": 7 if k == 2 else -1}


class _FakePattern:
    def __init__(self, pid, nbus):
        self.pid, self.nbus = pid, nbus

    def set_pattern(self, s, k=0):
        s.k = k

    def bind_system(self, b, slot, k):
        b.ks[slot] = k


def _fake_server(npat=3, cap=4):
    from hykkt_amd.served import Server
    srv = object.__new__(Server)
    srv.patterns = {i: _FakePattern(i, 1000) for i in range(npat)}
    srv.plain = {i: _FakeSolver() for i in range(npat)}
    srv.slots = {i: _FakeBatch() for i in range(npat)}
    srv.capacity = {i: cap for i in range(npat)}
    srv.served = list(range(npat))
    srv.keep, srv.queue_batches, srv.failures, srv.refused = False, 4, [], []
    return srv


class _BlockingSolver(_FakeSolver):
    """A solver whose solves wait for `gate`: requests pile up meanwhile."""

    def __init__(self, gate):
        super().__init__()
        self.gate = gate

    def execute(self):
        self.gate.wait(5.0)
        return super().execute()


class _BlockingBatch(_FakeBatch):
    def __init__(self, gate):
        super().__init__()
        self.gate = gate

    def execute_batch(self, count):
        self.gate.wait(5.0)
        return super().execute_batch(count)


def test_admission_bounds_the_queues():
    """A pattern admits queue_batches x its capacity waiting requests; the
    rest are rejected while its worker is busy, in both forms, so the queues
    (and an admitted request's wait) stay bounded under overload."""
    import threading
    from hykkt_amd.served import Request
    reqs = [Request(i, i % 3, 0, 0.0) for i in range(24)]   # 8 per pattern, all at once
    for plain in (True, False):
        gate = threading.Event()
        srv = _fake_server(cap=2)
        srv.queue_batches = 1   # 2 waiting requests per pattern
        srv.plain = {i: _BlockingSolver(gate) for i in range(3)}
        srv.slots = {i: _BlockingBatch(gate) for i in range(3)}
        timer = threading.Timer(0.5, gate.set)
        timer.start()
        res, rej, pend, wall = srv.run(reqs, plain=plain)
        timer.join()
        assert pend == 0 and len(res) + rej == 24
        # waiting <= 2 per pattern, plus what the busy workers had taken
        in_flight = 1 if plain else 3 * 2
        assert 24 - (6 + in_flight) <= rej <= 24 - 6, (plain, rej)


def test_workers_survive_failed_requests():
    """A request the library rejects (an error code) or that meets a
    non-positive pivot is a failed result; the slot goes on with the next
    request, in the served (batched) and the plain form alike."""
    from hykkt_amd.served import Request
    reqs = [Request(i, i % 3, i % 4, 0.0) for i in range(24)]
    for plain in (False, True):
        srv = _fake_server()
        res, rej, pend, wall = srv.run(reqs, plain=plain)
        assert rej == 0 and pend == 0 and len(res) == 24
        out = summarize(res, rej, pend, wall)
        assert out["failed"] == 12 and out["non_positive_pivot"] == 6 and out["within_tol"] == 12
        assert len(srv.failures) == 6
        if plain:
            assert all("planted" in f for f in srv.failures)
        else:
            assert all("status -5" in f for f in srv.failures)
            assert max(r.batch for r in res) > 1   # requests present together share a batch
        by = {r.rid: r for r in res}
        assert all(by[i].rc == (-5 if i % 4 == 1 else 1 if i % 4 == 2 else 0) for i in range(24))
