"""Served form: continuous batching of KKT solve requests over resident
sparsity patterns, beside the plain one-request-at-a-time path.

Device state.  Every pattern keeps its solver slots resident: a slot is one
HykktSolver (symbolic analysis, permutation, factor storage and workspace, and
its own HIP streams).  Slots are created up front, largest patterns first,
while the device memory they take (measured with hipMemGetInfo around each
slot's first solve) fits the HBM budget; a pattern that gets no slot is not
served.  The request values (perturbed H, Jc, Jd values, Ds, right-hand sides)
come from a pool per pattern that is resident in HBM, so the request path has
no host-to-device copies (inputs_on_device).

Serving.  Each slot has a host worker thread.  Requests wait in their
pattern's FIFO (bounded: arrivals beyond it are rejected and counted) and the
first free slot of the pattern takes the next one.  A request arriving while
others are on the GPU is admitted as soon as a slot of its pattern frees up,
and the work of different slots overlaps on the device: each slot's kernels
run on its own streams.  A solve's arithmetic does not depend on what else is
running (every slot has its own buffers and the kernels are deterministic),
so a result is the same alone and in the mix.

Plain.  One worker takes the requests one at a time in arrival order over all
patterns, on the same resident solvers: the drop-in path run back to back.

Non-positive pivots.  A system whose H_gamma is not positive definite (H~ not
positive definite on the null space of Jc) stops the factorization at a
non-positive pivot, as in the reference; its result carries status 1 and the
pivot, and it counts as failed.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

import hykkt_amd as H
from hykkt_amd.generator import GridPattern, ds_spread_schedule

# Bus counts of the served mix and their shares (6:3:1).
SERVED_SIZES = ((1_000, 0.6), (5_000, 0.3), (20_000, 0.1))


@dataclass
class Request:
    rid: int
    pattern: int
    k: int            # pool entry
    t_arrive: float


@dataclass
class Result:
    rid: int
    pattern: int
    latency_s: float
    rc: int
    resid4: float
    solution: tuple | None = None
    k: int = -1          # the pattern's value set
    pivot: int = -1      # first non-positive pivot of the factorization (-1: none)


class ServedPattern:
    """One resident sparsity pattern: its generator, index arrays and a pool
    of value sets in HBM (fresh 1+0.05 N(0,1) perturbations, a new rhs, Ds
    spread growing with the pool index up to 1e-4..1e8, or `ds_ranges[k]`;
    gamma `gamma`, or `gammas[k]`, for entry k)."""

    def __init__(self, pid: int, nbus: int, seed: int, pool: int = 2, gamma: float = 1e5,
                 ds_ranges=None, gammas=None):
        self.pid, self.nbus = pid, nbus
        self.gammas = [float(gamma)] * pool if gammas is None else [float(x) for x in gammas]
        self.g = GridPattern(nbus, int(1.3 * nbus), seed=seed, gamma=gamma)
        g = self.g
        i32 = lambda a: H.DeviceArray(np.asarray(a, np.int32))
        self.idx = dict(Hrp=i32(g.H.rp), Hci=i32(g.H.ci), Jcrp=i32(g.Jc.rp), Jcci=i32(g.Jc.ci),
                        Jdrp=i32(g.Jd.rp), Jdci=i32(g.Jd.ci), Dsrp=i32(np.arange(g.d + 1)),
                        Dsci=i32(np.arange(g.d)))
        self.pool = []
        self.host = []
        for k in range(pool):
            sy = g.system(k, ds_range=ds_spread_schedule(k, pool) if ds_ranges is None
                          else ds_ranges[k])
            sy.gamma = self.gammas[k]
            f = lambda a: H.DeviceArray(np.asarray(a, np.float64))
            self.pool.append(dict(H=f(sy.H.v), Jc=f(sy.Jc.v), Jd=f(sy.Jd.v), Ds=f(sy.Ds.v),
                                  rx=f(sy.rx), rs=f(sy.rs), ry=f(sy.ry), ryd=f(sy.ryd)))
            self.host.append(sy)

    def bind(self, s: "H.HykktSolver", k: int):
        g, v, ix = self.g, self.pool[k], self.idx
        s.set_gamma(self.gammas[k])
        s.set_H(ix["Hrp"], ix["Hci"], v["H"], g.n, g.n, g.H.nnz)
        s.set_Ds(ix["Dsrp"], ix["Dsci"], v["Ds"], g.d, g.d, g.d)
        s.set_Jc(ix["Jcrp"], ix["Jcci"], v["Jc"], g.m, g.n, g.Jc.nnz)
        s.set_Jd(ix["Jdrp"], ix["Jdci"], v["Jd"], g.d, g.n, g.Jd.nnz)
        s.set_rx(v["rx"], g.n)
        s.set_rs(v["rs"], g.d)
        s.set_ry(v["ry"], g.m)
        s.set_ryd(v["ryd"], g.d)

    def free(self):
        for a in self.idx.values():
            a.free()
        for v in self.pool:
            for a in v.values():
                a.free()


def make_patterns(npat: int = 16, seed: int = 0, pool: int = 2, sizes=SERVED_SIZES):
    """npat patterns with bus counts in the served mix's shares (largest
    remainder rounding), seeds seed, seed+1, ..."""
    counts = [int(np.floor(npat * p)) for _, p in sizes]
    rem = sorted(range(len(sizes)), key=lambda i: -(npat * sizes[i][1] - counts[i]))
    for i in rem[:npat - sum(counts)]:
        counts[i] += 1
    pats, pid = [], 0
    for (nbus, _), c in zip(sizes, counts):
        for _ in range(c):
            pats.append(ServedPattern(pid, nbus, seed + pid, pool=pool))
            pid += 1
    return pats


class Server:
    """Solver slots for `patterns` within `hbm_budget` bytes; see the module
    docstring.  slots_per_pattern: slots created per pattern when the budget
    allows (at least one each while it does)."""

    def __init__(self, patterns, hbm_budget: float, slots_per_pattern: int = 1,
                 queue_limit: int = 4096, keep_solutions: bool = False):
        self.patterns = {p.pid: p for p in patterns}
        self.keep = keep_solutions
        self.queue_limit = queue_limit
        self.slots = {p.pid: [] for p in patterns}
        self.failures = []   # messages of solves the library reported as failed
        self.refused = []    # slots not kept: failed first solve or over the budget
        self.hbm_used = 0.0
        order = sorted(patterns, key=lambda p: -p.nbus)
        for _ in range(slots_per_pattern):
            for p in order:
                free0, _ = H.device_mem_info()
                s = H.HykktSolver(p.g.gamma)
                s.options(inputs_on_device=1)
                # symbolic analysis + first solve: the slot's full footprint; a
                # value set with a non-positive pivot still sizes the slot, and
                # a slot the device cannot hold is refused
                p.bind(s, 0)
                try:
                    rc = s.execute()
                except H.HykktError as e:
                    self.refused.append({"pattern": p.pid, "nbus": p.nbus, "error": str(e)})
                    s.close()
                    continue
                if rc == H.HYKKT_RESIDUAL_FAIL and s.stats()["pivot"] >= 0:
                    rc = 0
                H.device_synchronize()
                free1, _ = H.device_mem_info()
                used = max(free0 - free1, 0)
                if rc != 0 or self.hbm_used + used > hbm_budget:
                    st = s.stats()
                    self.refused.append({"pattern": p.pid, "nbus": p.nbus, "rc": rc,
                                         "bytes": used, "resid2": st["resid2"],
                                         "resid4": st["resid4"], "pivot": st["pivot"]})
                    s.close()
                    continue
                self.hbm_used += used
                self.slots[p.pid].append(s)
        self.served = [p.pid for p in patterns if self.slots[p.pid]]
        # each pattern's solve time alone (second solve on its first slot,
        # pool entry 1): the unit for weighting completions by their work
        # (a solve that fails there keeps the pattern out of the weighting)
        self.solo_s = {}
        for pid in self.served:
            s = self.slots[pid][0]
            self.patterns[pid].bind(s, 1 % len(self.patterns[pid].pool))
            t = time.perf_counter()
            try:
                rc = s.execute()
            except H.HykktError as e:
                self.failures.append(f"solo timing of pattern {pid}: {e}")
                continue
            dt = time.perf_counter() - t
            if rc == 0:
                self.solo_s[pid] = dt

    def close(self):
        for ss in self.slots.values():
            for s in ss:
                s.close()
        self.slots = {}

    # ---- one run over a request list ----
    def _solve(self, s, req: Request) -> Result:
        """One request on slot `s`.  A solve the library reports as failed
        (an error code) is a result with that code and an infinite residual;
        the slot goes on with the next request."""
        p = self.patterns[req.pattern]
        try:
            p.bind(s, req.k)
            rc = s.execute()
        except H.HykktError as e:
            self.failures.append(str(e))
            return Result(req.rid, req.pattern, 0.0, e.code, float("inf"), None)
        st = s.stats()
        sol = tuple(np.copy(a) for a in s.solution()) if self.keep else None
        return Result(req.rid, req.pattern, 0.0, rc, st["resid4"], sol, req.k, st["pivot"])

    def run(self, requests, plain: bool = False, deadline: float | None = None):
        """Serve `requests` (arrival times relative to the start, sorted).
        plain: one worker, one request at a time in arrival order.  With a
        `deadline` (seconds), arrivals stop there and admitted requests not
        started by then stay pending.  Returns (results, rejected, pending,
        wall seconds)."""
        t0 = time.perf_counter()
        results, lock = [], threading.Lock()
        stop = threading.Event()
        rejected = [0]
        if plain:
            queues = {None: queue.Queue()}
            workers = [(None, [self.slots[pid][0] for pid in self.served])]
        else:
            queues = {pid: queue.Queue() for pid in self.served}
            workers = [(pid, [s]) for pid in self.served for s in self.slots[pid]]

        def worker(key, solvers):
            q = queues[key]
            while True:
                try:
                    req = q.get(timeout=0.05)
                except queue.Empty:
                    if stop.is_set():
                        return
                    continue
                if req is None:
                    return
                if deadline is not None and time.perf_counter() - t0 > deadline:
                    continue   # past the deadline: pending, not run
                s = solvers[0] if key is not None else \
                    self.slots[req.pattern][0]
                r = self._solve(s, req)
                r.latency_s = time.perf_counter() - t0 - req.t_arrive
                with lock:
                    results.append(r)

        threads = [threading.Thread(target=worker, args=w, daemon=True) for w in workers]
        for t in threads:
            t.start()
        arrived = 0
        for req in requests:
            dt = req.t_arrive - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            if deadline is not None and time.perf_counter() - t0 > deadline:
                break
            arrived += 1
            key = None if plain else req.pattern
            if key not in queues or queues[key].qsize() >= self.queue_limit:
                rejected[0] += 1
                continue
            queues[key].put(req)
        if deadline is not None:
            while time.perf_counter() - t0 < deadline and any(q.qsize() for q in queues.values()):
                time.sleep(0.01)
        for q in queues.values():
            for _ in range(len(threads)):
                q.put(None)
        stop.set()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        pending = arrived - len(results) - rejected[0]   # admitted, not finished
        return results, rejected[0], pending, wall


def poisson_requests(patterns, rate: float, seconds: float, seed: int = 0, pool: int = 2,
                     served=None):
    """Poisson arrivals at `rate` per second for `seconds`; each picks a
    pattern in the served mix's shares and a pool entry."""
    rng = np.random.default_rng(seed)
    ids = served if served is not None else [p.pid for p in patterns]
    sizes = np.array([patterns[i].nbus for i in ids])
    share = {n: p for n, p in SERVED_SIZES}
    w = np.array([share.get(int(n), 1.0) / max(np.sum(sizes == n), 1) for n in sizes])
    w /= w.sum()
    t, reqs, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        reqs.append(Request(rid, int(ids[rng.choice(len(ids), p=w)]), int(rng.integers(pool)), t))
        rid += 1
    return reqs


def summarize(results, rejected, pending, wall, tol=1e-8, patterns=None, solo_s=None):
    """Counts, throughput (results within `tol` per wall second) and latency
    quantiles; with `patterns`, also per bus count.  With `solo_s` (seconds
    of each pattern's solve run alone), also the solo seconds of the results
    within `tol` per wall second: a throughput that does not depend on which
    sizes were completed (1.0 for one solve at a time, back to back)."""
    lat = np.array([r.latency_s for r in results]) * 1e3
    ok = [r for r in results if r.rc == 0 and r.resid4 <= tol]
    out = {"completed": len(results), "within_tol": len(ok),
           "failed": sum(1 for r in results if r.rc != 0),
           "non_positive_pivot": sum(1 for r in results if r.pivot >= 0), "rejected": rejected,
           "pending": pending, "wall_s": wall,
           "solves_per_s": len(ok) / wall if wall > 0 else 0.0,
           "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
           "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
           "max_resid4": max((r.resid4 for r in results), default=None)}
    if solo_s is not None:
        out["solo_s_per_s"] = (sum(solo_s[r.pattern] for r in ok if r.pattern in solo_s) / wall
                               if wall > 0 else 0.0)
    if patterns is not None:
        by = {}
        for n in sorted({p.nbus for p in patterns}):
            sel = [r for r in results if patterns[r.pattern].nbus == n]
            ls = np.array([r.latency_s for r in sel]) * 1e3
            by[str(n)] = {"completed": len(sel),
                          "p50_ms": float(np.percentile(ls, 50)) if len(ls) else None,
                          "p99_ms": float(np.percentile(ls, 99)) if len(ls) else None}
        out["by_buses"] = by
    return out
