"""Batched form: many independent KKT systems of uneven size, solved together.

A batch is a set of systems over a few distinct sparsity patterns (several
value draws per pattern, each with its own gamma).  Every pattern gets one
resident batch slot (symbolic analysis, permutation, and per-system factor
storage and workspace for all its draws; admission by HBM budget, as in the
served form).  All the draws of a pattern are solved in one pass through the
batch engine -- formation, Ruiz, the numeric Cholesky, the triangular sweeps
and the device-resident Schur CG each launch once for all of them -- and the
patterns' passes run concurrently on their own streams, so systems of
different sizes overlap on the device.  The same systems run one at a time
through the drop-in path are the plain comparison; each result is
bit-identical either way.  A system whose H_gamma is not positive definite
(the generator's random H can make H~ indefinite on the null space of Jc)
fails at a non-positive pivot, as in the reference, and is counted as such.
"""
from __future__ import annotations

import numpy as np

from hykkt_amd.served import SWEEP_SHARE, Request, Server, ServedPattern, summarize

# 256 systems: 32 patterns x 8 value draws, bus counts uniform on [200, 5000]
# (branches 1.3x buses), gamma log-uniform on [1e4, 1e6] per system, Ds
# log-uniform on [1e-4, 1e4].
BATCH_PATTERNS, BATCH_DRAWS = 32, 8
BATCH_BUSES = (200, 5000)
BATCH_GAMMA = (1e4, 1e6)
BATCH_DS = (1e-4, 1e4)


def make_batch(npat: int = BATCH_PATTERNS, draws: int = BATCH_DRAWS, seed: int = 0,
               buses=BATCH_BUSES, gamma=BATCH_GAMMA, ds=BATCH_DS):
    """The batch's patterns, each with `draws` value sets resident in HBM."""
    rng = np.random.default_rng([seed, 4099])
    pats = []
    for pid in range(npat):
        nbus = int(rng.integers(buses[0], buses[1] + 1))
        gs = np.exp(rng.uniform(np.log(gamma[0]), np.log(gamma[1]), size=draws))
        pats.append(ServedPattern(pid, nbus, seed=1000 * seed + pid, pool=draws,
                                  gamma=float(gs[0]), ds_ranges=[ds] * draws, gammas=gs))
    return pats


def batch_requests(patterns):
    """Every value draw of every pattern, all present at the start."""
    reqs, rid = [], 0
    for p in patterns:
        for k in range(len(p.pool)):
            reqs.append(Request(rid, p.pid, k, 0.0))
            rid += 1
    return reqs


def make_server(patterns, hbm_budget: float = 200e9, keep_solutions: bool = False,
                sweep_share: int = SWEEP_SHARE):
    """One batch slot per pattern holding all its draws."""
    draws = max(len(p.pool) for p in patterns)
    return Server(patterns, hbm_budget, capacity=draws, keep_solutions=keep_solutions,
                  sweep_share=sweep_share)


def run_batch(server: Server, patterns, plain: bool = False):
    """Solve the whole batch; returns (results, wall seconds)."""
    res, rejected, pending, wall = server.run(batch_requests(patterns), plain=plain)
    if rejected or pending:
        raise RuntimeError(f"batch: {rejected} rejected, {pending} not finished")
    return res, wall


def summarize_batch(results, wall, patterns, tol=1e-8):
    out = summarize(results, 0, 0, wall, tol=tol)
    out["systems"] = len(results)
    out["systems_per_s"] = len(results) / wall if wall > 0 else 0.0
    out["buses_min_max"] = [min(p.nbus for p in patterns), max(p.nbus for p in patterns)]
    return out
