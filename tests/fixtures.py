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
