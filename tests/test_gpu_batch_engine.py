"""The batched form on the GPU (hykkt_create_batch / hykkt_bind_system /
hykkt_execute_batch): systems of one sparsity pattern solved together, every
kernel launch covering all of them.  Each system must come out bit for bit
as it does alone through the drop-in hykkt_execute -- solution, CG
iterations, refinement passes and residuals -- whatever else is in the batch,
including systems that stop CG at other iterations, refine a different number
of times, fail with a non-positive pivot or carry a zero row.  Needs an
MI355X."""
import numpy as np
import pytest

import _oracle as O
import hykkt_amd as H
from hykkt_amd.generator import GridPattern

pytestmark = pytest.mark.gpu


def draws(g, k, ds_ranges, gammas):
    out = []
    for i in range(k):
        sy = g.system(i, ds_range=ds_ranges[i % len(ds_ranges)])
        sy.gamma = gammas[i % len(gammas)]
        out.append(sy)
    return out


def set_pattern(s, g):
    d = g.d
    s.set_H(g.H.rp, g.H.ci, g.H.v, g.n, g.n, g.H.nnz)
    s.set_Ds(np.arange(d + 1), np.arange(d), np.ones(d), d, d, d)
    s.set_Jc(g.Jc.rp, g.Jc.ci, g.Jc.v, g.m, g.n, g.Jc.nnz)
    s.set_Jd(g.Jd.rp, g.Jd.ci, g.Jd.v, g.d, g.n, g.Jd.nnz)


def alone(g, sy, **opts):
    s = H.HykktSolver(sy.gamma)
    if opts:
        s.options(**opts)
    d = g.d
    s.set_H(sy.H.rp, sy.H.ci, sy.H.v, g.n, g.n, sy.H.nnz)
    s.set_Ds(np.arange(d + 1), np.arange(d), sy.Ds.v, d, d, d)
    s.set_Jc(sy.Jc.rp, sy.Jc.ci, sy.Jc.v, g.m, g.n, sy.Jc.nnz)
    s.set_Jd(sy.Jd.rp, sy.Jd.ci, sy.Jd.v, g.d, g.n, sy.Jd.nnz)
    for name in ("rx", "rs", "ry", "ryd"):
        v = getattr(sy, name)
        getattr(s, "set_" + name)(v, len(v))
    try:
        rc = s.execute()
    except H.HykktError as e:
        return e.code, None, None, s
    return rc, s.stats(), np.concatenate(s.solution()), s


def batched(g, systems, capacity=None, device=False, **opts):
    b = H.BatchSolver(systems[0].gamma, capacity or len(systems))
    if device:
        b.options(inputs_on_device=1)
        D = H.DeviceArray
        i32 = lambda a: D(np.asarray(a, np.int32))
        d = g.d
        b._pat = [i32(g.H.rp), i32(g.H.ci), i32(np.arange(d + 1)), i32(np.arange(d)),
                  i32(g.Jc.rp), i32(g.Jc.ci), i32(g.Jd.rp), i32(g.Jd.ci)]
        P = b._pat
        ones = D(np.ones(max(g.H.nnz, g.Jc.nnz, g.Jd.nnz, d)))
        b._pat.append(ones)
        b.set_H(P[0], P[1], ones, g.n, g.n, g.H.nnz)
        b.set_Ds(P[2], P[3], ones, d, d, d)
        b.set_Jc(P[4], P[5], ones, g.m, g.n, g.Jc.nnz)
        b.set_Jd(P[6], P[7], ones, g.d, g.n, g.Jd.nnz)
    else:
        set_pattern(b, g)
    if opts:
        b.options(**opts)
    for k, sy in enumerate(systems):
        arrs = [sy.H.v, sy.Ds.v, sy.Jc.v, sy.Jd.v, sy.rx, sy.rs, sy.ry, sy.ryd]
        if device:
            arrs = [H.DeviceArray(np.asarray(a, np.float64)) for a in arrs]
        b.bind_system(k, sy.gamma, *arrs)
    rc = b.execute_batch(len(systems))
    return b, rc


def check_same(g, systems, b, rc, **opts):
    for k, sy in enumerate(systems):
        rc1, st1, x1, _ = alone(g, sy, **opts)
        assert rc[k] == rc1, (k, rc[k], rc1)
        if st1 is None:
            continue
        st = b.stats_k(k)
        for key in ("cg_iters", "cg_status", "refine_steps", "pivot", "resid2", "resid4",
                    "resid4_first", "cg_final_norm"):
            assert st[key] == st1[key] or (np.isnan(st[key]) and np.isnan(st1[key])), (k, key)
        assert np.array_equal(np.concatenate(b.solution_k(k)), x1), k


def test_batch_bitwise_equal_to_alone():
    """Five draws of a 300-bus pattern with different gammas and Ds spreads
    (so CG stops at different iterations and refinement differs)."""
    g = GridPattern(300, 390, seed=41, gamma=1e5)
    systems = draws(g, 5, [(1e-2, 1e2), (1e-4, 1e8), (1e-3, 1e5), (1e-4, 1e6), (1.0, 1.0)],
                    [1e4, 1e5, 1e6, 3e4, 1e5])
    b, rc = batched(g, systems)
    print("batch status", rc, [b.stats_k(k)["cg_iters"] for k in range(5)],
          [b.stats_k(k)["refine_steps"] for k in range(5)], flush=True)
    assert all(r == 0 for r in rc)
    assert all(b.stats_k(k)["resid4"] <= 1e-8 for k in range(5))
    check_same(g, systems, b, rc)


def test_batch_partial_count_and_reuse():
    """A batch of capacity 6 run with 3 systems, then with 6 (the first three
    again, their results unchanged), then with 1."""
    g = GridPattern(120, 156, seed=7, gamma=1e4)
    systems = draws(g, 6, [(1e-3, 1e3), (1e-4, 1e4)], [1e4, 2e4, 5e4])
    b, rc = batched(g, systems[:3], capacity=6)
    first = [np.concatenate(b.solution_k(k)) for k in range(3)]
    for k, sy in enumerate(systems[3:], start=3):
        b.bind_system(k, sy.gamma, sy.H.v, sy.Ds.v, sy.Jc.v, sy.Jd.v, sy.rx, sy.rs, sy.ry, sy.ryd)
    rc6 = b.execute_batch(6)
    assert rc6[:3] == rc and all(r == 0 for r in rc6)
    for k in range(3):
        assert np.array_equal(np.concatenate(b.solution_k(k)), first[k])
    check_same(g, systems, b, rc6)
    assert b.execute_batch(1) == [0]
    assert np.array_equal(np.concatenate(b.solution_k(0)), first[0])


def test_batch_device_inputs_and_oracle():
    """Device-resident inputs give the same bits as host inputs; one system
    agrees with the CPU oracle to 1e-8."""
    g = GridPattern(200, 260, seed=3, gamma=1e5)
    systems = draws(g, 4, [(1e-3, 1e4), (1e-4, 1e8)], [1e5, 3e5])
    bh, rch = batched(g, systems)
    bd, rcd = batched(g, systems, device=True)
    assert rch == rcd == [0, 0, 0, 0]
    for k in range(4):
        assert np.array_equal(np.concatenate(bh.solution_k(k)), np.concatenate(bd.solution_k(k)))
    sy = systems[2]
    c = lambda m: O.Csr(m.n, m.m, m.rp, m.ci, m.v)
    ref = O.execute(c(sy.H), sy.Ds.v, c(sy.Jc), c(sy.Jd), sy.rx, sy.rs, sy.ry, sy.ryd, sy.gamma,
                    perm=bh.permutation())
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    x = np.concatenate(bh.solution_k(2))
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"batched system vs oracle: {err:.3e}", flush=True)
    assert err < 1e-8


def test_batch_with_failing_systems():
    """In one batch: a system with a zero Jc row (argument error), one whose
    H is negated (non-positive pivot), and healthy ones; the healthy ones come
    out as alone, the failing ones report as they do alone."""
    g = GridPattern(100, 130, seed=9, gamma=1e4)
    systems = draws(g, 4, [(1e-3, 1e3)], [1e4])
    bad_row = systems[1]
    r = 5
    bad_row.Jc.v = bad_row.Jc.v.copy()
    bad_row.Jc.v[bad_row.Jc.rp[r]:bad_row.Jc.rp[r + 1]] = 0.0
    bad_piv = systems[2]
    bad_piv.H.v = -np.abs(bad_piv.H.v)
    bad_piv.gamma = 0.0
    b, rc = batched(g, systems)
    assert rc[0] == 0 and rc[3] == 0
    assert rc[1] == -1           # HYKKT_E_ARG: the zero row
    assert rc[2] == 1 and b.stats_k(2)["pivot"] >= 0
    # the failed factorization's system skips the solve phase: zero solution,
    # no CG iterations, relative residual 1
    st2 = b.stats_k(2)
    assert st2["cg_iters"] == 0 and st2["refine_steps"] == 0
    assert abs(st2["resid4"] - 1.0) < 1e-12
    assert not np.any(np.concatenate(b.solution_k(2)))
    rc2, st2a, x2a, _ = alone(g, bad_piv)
    assert rc2 == 1 and st2a["cg_iters"] == 0 and not np.any(x2a)
    rc1, _, _, _ = alone(g, bad_row)
    assert rc1 == -1
    check_same(g, [systems[0]], b, rc[:1])
    rc3, st3, x3, _ = alone(g, systems[3])
    assert rc3 == 0 and np.array_equal(np.concatenate(b.solution_k(3)), x3)


def test_batch_refine_and_cg_options():
    """refine_max = 0 (the reference's single pass) and a fixed CG check
    interval give the same per-system results alone and batched."""
    g = GridPattern(150, 195, seed=13, gamma=1e5)
    systems = draws(g, 3, [(1e-4, 1e8), (1e-2, 1e2)], [1e5, 1e6])
    for opts in (dict(refine_max=0), dict(cg_check_every=2), dict(cg_check_every=1)):
        b, rc = batched(g, systems, **opts)
        check_same(g, systems, b, rc, **opts)


def test_sweep_share_same_bits():
    """Sweeps bounded to a share of the device's workgroup slots (persistent
    workgroups taking the items in order) give the same bits as unbounded
    ones, batched and alone."""
    g = GridPattern(400, 520, seed=21, gamma=1e5)
    systems = draws(g, 3, [(1e-4, 1e8), (1e-2, 1e2)], [1e5, 3e4])
    b1, rc1 = batched(g, systems)
    b8, rc8 = batched(g, systems, sweep_share=8)
    assert rc1 == rc8
    for k in range(len(systems)):
        assert np.array_equal(np.concatenate(b1.solution_k(k)), np.concatenate(b8.solution_k(k)))
    check_same(g, systems, b8, rc8, sweep_share=8)
    with pytest.raises(H.HykktError) as ex:
        b8.options(sweep_share=0)
    assert ex.value.code == -1
