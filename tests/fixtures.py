"""Loaders for the reference's fixture systems (src/mats) and the CTest cases
that use them (examples/CMakeLists.txt:22-86)."""
import _oracle as O

# (directory, tag, skip_lines) -- examples/CMakeLists.txt:22-86
A200 = ("a200", "ACTIVSg200_AC", 2)
A2000 = ("a2000", "ACTIVSg2000_AC", 3)


def files(kind, step):
    d, tag, skip = kind
    f = lambda blk: f"{d}/block_{blk}_{tag}_{step}.mtx"
    return dict(H=f("H_matrix"), Ds=f("Dd_matrix"), Jc=f("J_matrix"), Jd=f("Jd_matrix"),
                rx=f("rx"), rs=f("rs"), ry=f("ry"), ryd=f("ryd"), skip=skip)


def load(mats, kind, step):
    fl = files(kind, step)
    sk = fl["skip"]
    H = O.read_mm(mats(fl["H"]), sk, True)
    Ds = O.read_mm(mats(fl["Ds"]), sk, False)
    Jc = O.read_mm(mats(fl["Jc"]), sk, False)
    Jd = O.read_mm(mats(fl["Jd"]), sk, False)
    vec = lambda k: O.read_rhs(mats(fl[k]))
    return dict(H=H, Ds=Ds, Jc=Jc, Jd=Jd, rx=vec("rx"), rs=vec("rs"), ry=vec("ry"),
                ryd=vec("ryd"), skip=sk, files=fl)


def hgamma_in_factor_order(solver, nnz_hgam, n, hgam_natural):
    """The oracle's H_gamma values (natural CSR order) rearranged the way
    debug selector 10 returns what the factorization received: P*Hg*P^T CSR
    order (selectors 15/16, value map 17), NaN above the diagonal."""
    import numpy as np
    prp = solver.debug(15, n + 1, np.int32)
    pci = solver.debug(16, nnz_hgam, np.int32)
    pmap = solver.debug(17, nnz_hgam, np.int32)
    rows = np.repeat(np.arange(n), np.diff(prp))
    want = hgam_natural[pmap].copy()
    want[pci > rows] = np.nan
    return want


def same_entries(got, want):
    """Same NaN positions, and the other entries equal bit for bit."""
    import numpy as np
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    if got.shape != want.shape:
        return False
    nan = np.isnan(want)
    return (np.array_equal(np.isnan(got), nan)
            and np.array_equal(got[~nan].view(np.int64), want[~nan].view(np.int64)))
