"""Served form: continuous batching of KKT solve requests over resident
sparsity patterns, beside the plain one-request-at-a-time path.

Device state.  Every admitted pattern keeps a batch slot resident: a
BatchSolver (symbolic analysis, permutation, and per-system storage for up to
`capacity` systems: factor panels, update matrices, workspace) with its own
HIP streams.  Slots are created up front, largest patterns first, while the
device memory they take (measured with hipMemGetInfo around each slot's first
batch) fits the HBM budget; a pattern that gets no slot is not served and its
requests are rejected.  Request values (perturbed H, Jc, Jd values, Ds,
right-hand sides) come from a pool per pattern resident in HBM, so the
request path copies nothing from the host.

Admission.  A pattern admits at most `queue_batches` batches' worth of
waiting requests (queue_batches x its slot capacity); arrivals beyond that
are rejected and counted, so under overload the queues, and the latency of
an admitted request, stay bounded.  The plain path admits by the same rule.

Serving.  Requests wait in their pattern's FIFO.  Each slot has a host worker: whenever the slot is
free it takes every request of its pattern that has arrived, up to the slot's
capacity, and solves them as one batch -- every kernel launch covers all of
them.  Requests that arrive while a batch is on the GPU join the next one, so
batches grow with the load (continuous batching), and the slots of different
patterns overlap on the device.  A system's arithmetic in a batch is exactly
its arithmetic alone, so a result is the same bits alone and in the mix.

Plain.  One worker takes the requests one at a time in arrival order over all
patterns, each through the drop-in HykktSolver.execute() of its pattern's
plain solver.

Non-positive pivots.  A system whose H_gamma is not positive definite (H~ not
positive definite on the null space of Jc) stops the factorization at a
non-positive pivot, as in the reference; its result carries status 1 and the
pivot, and it counts as failed.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass

import numpy as np

import hykkt_amd as H
from hykkt_amd.generator import GridPattern, ds_spread_schedule

# Bus counts of the served mix and their shares (6:3:1).
SERVED_SIZES = ((1_000, 0.6), (5_000, 0.3), (20_000, 0.1))
# The batch slots' sweeps each hold at most 1/SWEEP_SHARE of the device's
# workgroup slots, so the slots' latency-bound sweeps overlap; the plain
# solvers, one solve at a time, keep the whole device.
SWEEP_SHARE = 4
# Value sets per served pattern: request i of a pattern uses entry i mod POOL,
# whose Ds spread grows with the entry (the interior-point step index).
POOL = 20


def default_capacity(nbus: int) -> int:
    """Systems per batch slot: small patterns are latency-bound, so many of
    them share each launch; large ones fill the device on their own."""
    if nbus <= 2_000:
        return 16
    if nbus <= 8_000:
        return 8
    return 2


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
    batch: int = 1       # systems in the batch that solved it
    t_arrive: float = 0.0


class ServedPattern:
    """One resident sparsity pattern: its generator, index arrays and a pool
    of value sets in HBM (fresh 1+0.05 N(0,1) perturbations, a new rhs, Ds
    spread growing with the pool index up to 1e-4..1e8, or `ds_ranges[k]`;
    gamma `gamma`, or `gammas[k]`, for entry k)."""

    def __init__(self, pid: int, nbus: int, seed: int, pool: int = POOL, gamma: float = 1e5,
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

    def set_pattern(self, s: "H.HykktSolver", k: int = 0):
        """The pattern (and, for the drop-in path, the values of entry k)."""
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

    bind = set_pattern

    def bind_system(self, b: "H.BatchSolver", slot: int, k: int):
        """Pool entry k as system `slot` of batch solver b."""
        v = self.pool[k]
        b.bind_system(slot, self.gammas[k], v["H"], v["Ds"], v["Jc"], v["Jd"], v["rx"], v["rs"],
                      v["ry"], v["ryd"])

    def free(self):
        for a in self.idx.values():
            a.free()
        for v in self.pool:
            for a in v.values():
                a.free()


def make_patterns(npat: int = 16, seed: int = 0, pool: int = POOL, sizes=SERVED_SIZES):
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
    """A batch slot and a plain solver for each of `patterns` within
    `hbm_budget` bytes; see the module docstring.  capacity: systems per batch
    slot (a function of the bus count, default_capacity by default)."""

    def __init__(self, patterns, hbm_budget: float, capacity=None, queue_batches: int = 4,
                 keep_solutions: bool = False, sweep_share: int = SWEEP_SHARE):
        self.patterns = {p.pid: p for p in patterns}
        self.keep = keep_solutions
        self.queue_batches = queue_batches
        self.slots = {}      # pid -> BatchSolver
        self.plain = {}      # pid -> HykktSolver (the drop-in path)
        self.failures = []   # messages of solves the library reported as failed
        self.refused = []    # patterns not kept: failed first solve or over the budget
        self.hbm_used = 0.0
        cap_of = capacity if callable(capacity) else (
            (lambda nbus: capacity) if capacity else default_capacity)
        self.capacity = {}
        for p in sorted(patterns, key=lambda p: -p.nbus):
            cap = max(1, min(cap_of(p.nbus), len(p.pool)))
            free0, _ = H.device_mem_info()
            b = H.BatchSolver(p.g.gamma, cap)
            b.options(inputs_on_device=1, sweep_share=sweep_share)
            s = H.HykktSolver(p.g.gamma)
            s.options(inputs_on_device=1)
            p.set_pattern(b, 0)
            p.set_pattern(s, 0)
            # symbolic analysis + a first full batch and a first plain solve:
            # the slot's footprint; a value set with a non-positive pivot still
            # sizes the slot, and a slot the device cannot hold is refused
            try:
                for j in range(cap):
                    p.bind_system(b, j, j % len(p.pool))
                rcs = b.execute_batch(cap)
                rc = s.execute()
            except H.HykktError as e:
                self.refused.append({"pattern": p.pid, "nbus": p.nbus, "error": str(e)})
                b.close()
                s.close()
                continue
            H.device_synchronize()
            free1, _ = H.device_mem_info()
            used = max(free0 - free1, 0)
            bad = [r for j, r in enumerate(rcs) if r < 0]
            if bad or rc < 0 or self.hbm_used + used > hbm_budget:
                self.refused.append({"pattern": p.pid, "nbus": p.nbus, "rc": [rc] + rcs,
                                     "bytes": used})
                b.close()
                s.close()
                continue
            self.hbm_used += used
            self.slots[p.pid] = b
            self.plain[p.pid] = s
            self.capacity[p.pid] = cap
        self.served = [p.pid for p in patterns if p.pid in self.slots]
        # each pattern's solve time alone (a second plain solve, pool entry 1):
        # the unit for weighting completions by their work (a solve that fails
        # there keeps the pattern out of the weighting)
        self.solo_s = {}
        for pid in self.served:
            s = self.plain[pid]
            self.patterns[pid].set_pattern(s, 1 % len(self.patterns[pid].pool))
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
        for ss in (self.slots, self.plain):
            for s in ss.values():
                s.close()
        self.slots, self.plain = {}, {}

    # ---- solving ----
    def _solve_plain(self, req: Request) -> Result:
        """One request through the drop-in path of its pattern.  A solve the
        library reports as failed (an error code) is a result with that code
        and an infinite residual; the worker goes on with the next request."""
        p = self.patterns[req.pattern]
        s = self.plain[req.pattern]
        try:
            p.set_pattern(s, req.k)
            rc = s.execute()
        except H.HykktError as e:
            self.failures.append(str(e))
            return Result(req.rid, req.pattern, 0.0, e.code, float("inf"), None, req.k)
        st = s.stats()
        sol = tuple(np.copy(a) for a in s.solution()) if self.keep else None
        return Result(req.rid, req.pattern, 0.0, rc, st["resid4"], sol, req.k, st["pivot"])

    def _solve_batch(self, reqs) -> list:
        """Requests of one pattern as one batch on its slot."""
        pid = reqs[0].pattern
        p, b = self.patterns[pid], self.slots[pid]
        try:
            for j, r in enumerate(reqs):
                p.bind_system(b, j, r.k)
            rcs = b.execute_batch(len(reqs))
        except H.HykktError as e:
            self.failures.append(str(e))
            return [Result(r.rid, pid, 0.0, e.code, float("inf"), None, r.k, -1, len(reqs))
                    for r in reqs]
        out = []
        for j, r in enumerate(reqs):
            st = b.stats_k(j)
            sol = tuple(np.copy(a) for a in b.solution_k(j)) if self.keep else None
            rc = rcs[j]
            if rc < 0:
                self.failures.append(f"request {r.rid}: status {rc}")
            out.append(Result(r.rid, pid, 0.0, rc, st["resid4"] if rc >= 0 else float("inf"),
                              sol, r.k, st["pivot"], len(reqs)))
        return out

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
        # admitted requests of each pattern not yet taken by a worker
        waiting = {pid: 0 for pid in self.served}
        limit = {pid: self.queue_batches * self.capacity[pid] for pid in self.served}
        if plain:
            queues = {None: queue.Queue()}
            keys = [None]
        else:
            queues = {pid: queue.Queue() for pid in self.served}
            keys = list(self.served)

        def late():
            return deadline is not None and time.perf_counter() - t0 > deadline

        def worker(key):
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
                with lock:
                    waiting[req.pattern] -= 1
                if late():
                    continue   # past the deadline: pending, not run
                if key is None:
                    res = [self._solve_plain(req)]
                else:
                    batch, done = [req], False
                    while len(batch) < self.capacity[key]:
                        try:
                            nxt = q.get_nowait()
                        except queue.Empty:
                            break
                        if nxt is None:
                            done = True
                            break
                        with lock:
                            waiting[nxt.pattern] -= 1
                        batch.append(nxt)
                    res = self._solve_batch(batch)
                t = time.perf_counter() - t0
                reqs = {r.rid: r for r in ([req] if key is None else batch)}
                for r in res:
                    r.t_arrive = reqs[r.rid].t_arrive
                    r.latency_s = t - r.t_arrive
                with lock:
                    results.extend(res)
                if key is not None and done:
                    return

        threads = [threading.Thread(target=worker, args=(k,), daemon=True) for k in keys]
        for t in threads:
            t.start()
        arrived = 0
        for req in requests:
            dt = req.t_arrive - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            if late():
                break
            arrived += 1
            key = None if plain else req.pattern
            with lock:
                full = req.pattern not in waiting or waiting[req.pattern] >= limit[req.pattern]
                if not full:
                    waiting[req.pattern] += 1
            if full:
                rejected[0] += 1
                continue
            queues[key].put(req)
        if deadline is not None:
            while not late() and any(q.qsize() for q in queues.values()):
                time.sleep(0.01)
        for q in queues.values():
            q.put(None)
        stop.set()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        pending = arrived - len(results) - rejected[0]   # admitted, not finished
        return results, rejected[0], pending, wall


def poisson_requests(patterns, rate: float, seconds: float, seed: int = 0, pool: int = POOL,
                     served=None):
    """Poisson arrivals at `rate` per second for `seconds`; each picks a
    pattern in the served mix's shares.  A pattern's i-th request is its
    interior-point step i: pool entry i mod pool, whose Ds spread grows with
    the entry."""
    rng = np.random.default_rng(seed)
    ids = served if served is not None else [p.pid for p in patterns]
    by_id = {p.pid: p for p in patterns}
    sizes = np.array([by_id[i].nbus for i in ids])
    share = {n: p for n, p in SERVED_SIZES}
    w = np.array([share.get(int(n), 1.0) / max(np.sum(sizes == n), 1) for n in sizes])
    w /= w.sum()
    t, reqs, rid = 0.0, [], 0
    step = {i: 0 for i in ids}
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        pid = int(ids[rng.choice(len(ids), p=w)])
        reqs.append(Request(rid, pid, step[pid] % pool, t))
        step[pid] += 1
        rid += 1
    return reqs


def summarize(results, rejected, pending, wall, tol=1e-8, patterns=None, solo_s=None,
              warmup=0.0):
    """Counts, throughput (results within `tol` per wall second) and latency
    quantiles of the requests that arrived after `warmup` seconds; with
    `patterns`, also per bus count.  With `solo_s` (seconds of each pattern's
    solve run alone), also the solo seconds of the results within `tol` per
    wall second: a throughput that does not depend on which sizes were
    completed (1.0 for one solve at a time, back to back)."""
    lat = np.array([r.latency_s for r in results if r.t_arrive >= warmup]) * 1e3
    ok = [r for r in results if r.rc == 0 and r.resid4 <= tol]
    out = {"completed": len(results), "within_tol": len(ok),
           "failed": sum(1 for r in results if r.rc != 0),
           "non_positive_pivot": sum(1 for r in results if r.pivot >= 0), "rejected": rejected,
           "pending": pending, "wall_s": wall,
           "solves_per_s": len(ok) / wall if wall > 0 else 0.0,
           "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
           "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
           "max_resid4": max((r.resid4 for r in results), default=None),
           "mean_batch": float(np.mean([r.batch for r in results])) if results else None,
           "latency_after_s": warmup}
    done = {}
    for r in ok:
        done[str(r.pattern)] = done.get(str(r.pattern), 0) + 1
    out["within_tol_by_pattern"] = done
    if solo_s is not None:
        out["solo_s_per_s"] = (sum(solo_s[r.pattern] for r in ok if r.pattern in solo_s) / wall
                               if wall > 0 else 0.0)
    if patterns is not None:
        by = {}
        pmap = patterns if isinstance(patterns, dict) else {p.pid: p for p in patterns}
        for n in sorted({p.nbus for p in pmap.values()}):
            sel = [r for r in results if pmap[r.pattern].nbus == n and r.t_arrive >= warmup]
            ls = np.array([r.latency_s for r in sel]) * 1e3
            by[str(n)] = {"completed": len(sel),
                          "p50_ms": float(np.percentile(ls, 50)) if len(ls) else None,
                          "p99_ms": float(np.percentile(ls, 99)) if len(ls) else None}
        out["by_buses"] = by
    return out
