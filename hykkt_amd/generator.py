"""Seeded synthetic OPF-like KKT systems (the benchmark workloads).

A grid of `nbus` buses is a random spanning tree (bus i attaches to a uniformly
drawn earlier bus) plus distinct random extra edges up to `nbranch` branches.
Each bus carries 4 primal variables (n = 4*nbus).  Blocks:

  H   (n x n)  symmetric: per bus a symmetrised 4x4 N(0,1) block, per branch a
               4x4 N(0,1) coupling block at (a,b) and its transpose at (b,a),
               plus Dx ~ U[1e-2, 1e2] on the diagonal (the "H + Dx" block).
  Jc  (2*nbus x n)    two rows per bus over the variables of the bus and of
                      its adjacent buses, N(0,1).
  Jd  (2*nbranch x n) two rows per branch over both endpoint buses' variables.
  Ds  (d x d)  diagonal, log-uniform on [ds_lo, ds_hi].
  rx, rs, ry, ryd ~ N(0,1).

Same-pattern sequences rescale every nonzero of H, Jc, Jd (and H's mirror
entry by the same factor, keeping H symmetric) by 1 + 0.05*N(0,1), drawn
afresh from the base values at each step (non-cumulative).

The PRNG is numpy's PCG64 (`numpy.random.default_rng(seed)`), so every system
is reproducible from (sizes, seed, step).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp


@dataclass
class CsrBlock:
    n: int
    m: int
    rp: np.ndarray   # int32 (n+1)
    ci: np.ndarray   # int32 (nnz)
    v: np.ndarray    # float64 (nnz)

    @property
    def nnz(self) -> int:
        return int(self.rp[-1])


@dataclass
class KKTSystem:
    H: CsrBlock
    Ds: CsrBlock
    Jc: CsrBlock
    Jd: CsrBlock
    rx: np.ndarray
    rs: np.ndarray
    ry: np.ndarray
    ryd: np.ndarray
    gamma: float

    @property
    def sizes(self):
        return self.H.n, self.Jc.n, self.Jd.n


def _to_block(M: sp.spmatrix) -> CsrBlock:
    M = sp.csr_matrix(M)
    M.sum_duplicates()
    M.sort_indices()
    return CsrBlock(M.shape[0], M.shape[1], M.indptr.astype(np.int32),
                    M.indices.astype(np.int32), M.data.astype(np.float64))


def grid_edges(nbus: int, nbranch: int, rng: np.random.Generator) -> np.ndarray:
    """Random spanning tree by attachment plus distinct extra edges (a < b)."""
    if nbranch < nbus - 1:
        raise ValueError("need at least nbus-1 branches for a spanning tree")
    if nbranch > nbus * (nbus - 1) // 2:
        raise ValueError("too many branches for a simple graph")
    child = np.arange(1, nbus, dtype=np.int64)
    parent = (rng.random(nbus - 1) * child).astype(np.int64)
    a = np.minimum(parent, child)
    b = np.maximum(parent, child)
    keys = set((a * nbus + b).tolist())
    extra = []
    need = nbranch - (nbus - 1)
    while need > 0:
        u = rng.integers(0, nbus, size=2 * need + 16)
        w = rng.integers(0, nbus, size=2 * need + 16)
        for x, y in zip(u.tolist(), w.tolist()):
            if x == y:
                continue
            lo, hi = (x, y) if x < y else (y, x)
            k = lo * nbus + hi
            if k in keys:
                continue
            keys.add(k)
            extra.append((lo, hi))
            need -= 1
            if need == 0:
                break
    e = np.array([(int(x), int(y)) for x, y in zip(a, b)] + extra, dtype=np.int64)
    return e.reshape(-1, 2)


class GridPattern:
    """Sparsity pattern and base values of one synthetic KKT system."""

    def __init__(self, nbus: int, nbranch: int, seed: int = 0, gamma: float = 1e4,
                 ds_range=(1e-4, 1e4)):
        self.nbus, self.nbranch, self.seed = nbus, nbranch, seed
        self.gamma = float(gamma)
        self.ds_range = ds_range
        rng = np.random.default_rng(seed)
        E = grid_edges(nbus, nbranch, rng)
        self.edges = E
        n = 4 * nbus
        blk = np.arange(4)
        # ---- H: bus blocks (symmetrised) + branch coupling blocks + Dx ----
        G = rng.standard_normal((nbus, 4, 4))
        G = 0.5 * (G + G.transpose(0, 2, 1))
        G[:, blk, blk] += rng.uniform(1e-2, 1e2, size=(nbus, 4))
        Cb = rng.standard_normal((len(E), 4, 4))
        bi = (4 * np.arange(nbus))[:, None, None] + blk[None, :, None]
        bj = (4 * np.arange(nbus))[:, None, None] + blk[None, None, :]
        ea = (4 * E[:, 0])[:, None, None] + blk[None, :, None]
        eb = (4 * E[:, 1])[:, None, None] + blk[None, None, :]
        rows = np.concatenate([np.broadcast_to(bi, G.shape).ravel(),
                               np.broadcast_to(ea, Cb.shape).ravel(),
                               np.broadcast_to(eb.transpose(0, 2, 1), Cb.shape).ravel()])
        cols = np.concatenate([np.broadcast_to(bj, G.shape).ravel(),
                               np.broadcast_to(eb, Cb.shape).ravel(),
                               np.broadcast_to(ea.transpose(0, 2, 1), Cb.shape).ravel()])
        vals = np.concatenate([G.ravel(), Cb.ravel(), Cb.transpose(0, 2, 1).ravel()])
        self.H = _to_block(sp.coo_matrix((vals, (rows, cols)), shape=(n, n)))
        # ---- Jc: 2 rows per bus over own + adjacent buses' variables ----
        nbr = [[b] for b in range(nbus)]
        for x, y in E.tolist():
            nbr[x].append(y)
            nbr[y].append(x)
        jr, jc = [], []
        for b in range(nbus):
            vs = np.sort(np.concatenate([4 * np.array(nbr[b])[:, None] + blk[None, :]]).ravel())
            for r in (2 * b, 2 * b + 1):
                jr.append(np.full(len(vs), r))
                jc.append(vs)
        jr = np.concatenate(jr)
        jc = np.concatenate(jc)
        self.Jc = _to_block(sp.coo_matrix((rng.standard_normal(len(jr)), (jr, jc)),
                                          shape=(2 * nbus, n)))
        # ---- Jd: 2 rows per branch over both endpoints' variables ----
        nb = len(E)
        vs = np.concatenate([4 * E[:, :1] + blk[None, :], 4 * E[:, 1:2] + blk[None, :]], axis=1)
        dr = np.repeat(np.arange(2 * nb), 8)
        dc = np.repeat(vs, 2, axis=0).ravel()
        self.Jd = _to_block(sp.coo_matrix((rng.standard_normal(len(dr)), (dr, dc)),
                                          shape=(2 * nb, n)))
        self.n, self.m, self.d = n, 2 * nbus, 2 * nb
        # mirror map of H (position of (j,i) for each (i,j)) for symmetric rescaling
        Hm = sp.csr_matrix((np.arange(self.H.nnz, dtype=np.float64) + 1.0, self.H.ci, self.H.rp),
                           shape=(n, n))
        self._h_mirror = (Hm.T.tocsr().sorted_indices().data - 1.0).astype(np.int64)
        self._rng_base = rng.bit_generator.state

    def _ds(self, rng, lo, hi):
        return np.exp(rng.uniform(np.log(lo), np.log(hi), size=self.d))

    def system(self, step: int = -1, ds_range=None, perturb: bool | None = None) -> KKTSystem:
        """step < 0: base values.  step >= 0: values perturbed by 1+0.05*N(0,1)
        (independently per step) and fresh rhs.  ds_range overrides Ds spread."""
        rng = np.random.default_rng([self.seed, 7919, step + 1])
        perturb = (step >= 0) if perturb is None else perturb
        hv, jcv, jdv = self.H.v.copy(), self.Jc.v.copy(), self.Jd.v.copy()
        if perturb:
            f = 1.0 + 0.05 * rng.standard_normal(self.H.nnz)
            idx = np.arange(self.H.nnz)
            f = np.where(idx <= self._h_mirror, f, f[self._h_mirror])
            hv *= f
            jcv *= 1.0 + 0.05 * rng.standard_normal(self.Jc.nnz)
            jdv *= 1.0 + 0.05 * rng.standard_normal(self.Jd.nnz)
        lo, hi = ds_range if ds_range is not None else self.ds_range
        ds = self._ds(rng, lo, hi)
        d = self.d
        Ds = CsrBlock(d, d, np.arange(d + 1, dtype=np.int32), np.arange(d, dtype=np.int32), ds)
        H = CsrBlock(self.n, self.n, self.H.rp, self.H.ci, hv)
        Jc = CsrBlock(self.m, self.n, self.Jc.rp, self.Jc.ci, jcv)
        Jd = CsrBlock(self.d, self.n, self.Jd.rp, self.Jd.ci, jdv)
        return KKTSystem(H, Ds, Jc, Jd, rng.standard_normal(self.n), rng.standard_normal(d),
                         rng.standard_normal(self.m), rng.standard_normal(d), self.gamma)


def ds_spread_schedule(step: int, nsteps: int = 20, start=(1e-2, 1e2), end=(1e-4, 1e8)):
    """Ds range growing geometrically from `start` to `end` over `nsteps` steps."""
    t = 0.0 if nsteps <= 1 else step / (nsteps - 1)
    lo = start[0] * (end[0] / start[0]) ** t
    hi = start[1] * (end[1] / start[1]) ** t
    return lo, hi


# Named workloads (sizes and constants of the benchmark configurations).
CONFIGS = {
    "grid500": dict(nbus=500, nbranch=700, gamma=1e4, ds_range=(1e-4, 1e4)),
    "grid10k_seq20": dict(nbus=10_000, nbranch=13_000, gamma=1e5, ds_range=(1e-2, 1e2)),
    "grid80k": dict(nbus=80_000, nbranch=110_000, gamma=1e6, ds_range=(1e-6, 1e6)),
}


def workload(name: str, seed: int = 0) -> GridPattern:
    c = CONFIGS[name]
    return GridPattern(c["nbus"], c["nbranch"], seed=seed, gamma=c["gamma"],
                       ds_range=c["ds_range"])
