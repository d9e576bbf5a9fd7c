"""The reference's example drivers rebuilt on include/HykktSolver.hpp:
hykkt_amd/bin/solver_driver (examples/cuSolver_driver_solver.cpp) and
hykkt_amd/bin/solver_emulator (examples/cuSolver_driver_solver_emulator.cpp),
run with the CTest fixture systems (examples/CMakeLists.txt:22-86)."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import fixtures as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hykkt_amd", "bin")


def run(name, *args, out_dir=None):
    env = dict(os.environ)
    if out_dir is not None:
        env["HYKKT_DRIVER_SOLUTION_DIR"] = str(out_dir)
    return subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True,
                          text=True, timeout=120, env=env)


def check_vs_oracle(out_dir, tag, mats, kind, step, gamma=1e4, tol=1e-6):
    """The solution a driver wrote for `tag` against the oracle's solve of the
    same fixture system with the same ordering (relative difference < tol,
    the bound of the other fixture comparisons)."""
    s = F.load(mats, kind, step)
    rd = lambda w, t=np.float64: np.fromfile(os.path.join(out_dir, f"{tag}_{w}.bin"), dtype=t)
    perm = rd("perm", np.int32)
    mine = np.concatenate([rd(w) for w in ("dx", "ds", "dyc", "dyd")])
    ref = O.execute(s["H"], s["Ds"].v, s["Jc"], s["Jd"], s["rx"], s["rs"], s["ry"], s["ryd"],
                    gamma, perm=perm)
    xo = np.concatenate([ref["dx"], ref["ds"], ref["dyc"], ref["dyd"]])
    assert len(mine) == len(xo) and len(perm) == len(ref["dx"])
    err = np.linalg.norm(mine - xo) / np.linalg.norm(xo)
    print(f"{tag} ({kind[0]} step {step}) vs oracle: rel. difference {err:.3e}", flush=True)
    assert err < tol
    return err


@pytest.mark.parametrize("name", ["solver_driver", "solver_emulator"])
def test_driver_argument_count(name):
    """Wrong argument count: the reference's message and return -1."""
    r = run(name, "x")
    assert r.returncode == 255
    assert "Incorrect number of inputs. Exiting ..." in r.stdout


def _emulator_dir(mats, tmp_path, kind, steps):
    """The emulator's file naming (block_{H,Ds,Jc,Jd}_matrix_*, block_{rx,rs,ry,rd}_rhs_*)
    as links to the fixture files."""
    names = {"H": "block_H_matrix", "Ds": "block_Ds_matrix", "Jc": "block_Jc_matrix",
             "Jd": "block_Jd_matrix", "rx": "block_rx_rhs", "rs": "block_rs_rhs",
             "ry": "block_ry_rhs", "ryd": "block_rd_rhs"}
    for st in steps:
        fl = F.files(kind, st)
        for k, pre in names.items():
            os.symlink(mats(fl[k]), tmp_path / f"{pre}_{kind[1]}_{st}.mtx")
    return tmp_path


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [F.A200, F.A2000])
def test_solver_driver_two_systems(mats, kind, tmp_path):
    """CTest `solver` cases: system 09 bound, solved, overwritten in place by
    system 10, solved again; both pass (resid4 <= 1e-8 printed) and both
    solutions the driver wrote agree with the oracle's solves of 09 and 10."""
    a, b = F.files(kind, "09"), F.files(kind, "10")
    keys = ("H", "Ds", "Jc", "Jd", "rx", "rs", "ry", "ryd")
    r = run("solver_driver", *[mats(a[k]) for k in keys], *[mats(b[k]) for k in keys],
            a["skip"], "1e4", out_dir=tmp_path)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("system ")]
    assert len(lines) == 2
    for l in lines:
        assert float(l.split("resid4")[1].split()[0]) <= 1e-8
    check_vs_oracle(tmp_path, "system1", mats, kind, "09")
    check_vs_oracle(tmp_path, "system2", mats, kind, "10")


@pytest.mark.gpu
def test_solver_emulator_sequence(mats, tmp_path):
    """Same-pattern sequence 09,10 passes; switching to step 00, whose Jd is
    empty, is the reference's "new nonzero structure" failure (return 1)."""
    d = _emulator_dir(mats, tmp_path, F.A200, ["09", "10", "00"])
    out = tmp_path / "solutions"
    out.mkdir()
    r = run("solver_emulator", d, F.A200[1], "09,10", F.A200[2], "1e4", out_dir=out)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.count("status 0") == 2
    for step in ("09", "10"):
        check_vs_oracle(out, f"sequence_{step}", mats, F.A200, step)
    r = run("solver_emulator", d, F.A200[1], "09,00", F.A200[2], "1e4")
    assert r.returncode == 1
