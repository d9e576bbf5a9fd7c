"""The reference's example drivers rebuilt on include/HykktSolver.hpp:
hykkt_amd/bin/solver_driver (examples/cuSolver_driver_solver.cpp) and
hykkt_amd/bin/solver_emulator (examples/cuSolver_driver_solver_emulator.cpp),
run with the CTest fixture systems (examples/CMakeLists.txt:22-86)."""
import os
import subprocess

import pytest

import fixtures as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hykkt_amd", "bin")


def run(name, *args):
    return subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True,
                          text=True, timeout=120)


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
def test_solver_driver_two_systems(mats, kind):
    """CTest `solver` cases: system 09 bound, solved, overwritten in place by
    system 10, solved again; both pass (resid4 <= 1e-8 printed)."""
    a, b = F.files(kind, "09"), F.files(kind, "10")
    keys = ("H", "Ds", "Jc", "Jd", "rx", "rs", "ry", "ryd")
    r = run("solver_driver", *[mats(a[k]) for k in keys], *[mats(b[k]) for k in keys],
            a["skip"], "1e4")
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("system ")]
    assert len(lines) == 2
    for l in lines:
        assert float(l.split("resid4")[1].split()[0]) <= 1e-8


@pytest.mark.gpu
def test_solver_emulator_sequence(mats, tmp_path):
    """Same-pattern sequence 09,10 passes; switching to step 00, whose Jd is
    empty, is the reference's "new nonzero structure" failure (return 1)."""
    d = _emulator_dir(mats, tmp_path, F.A200, ["09", "10", "00"])
    r = run("solver_emulator", d, F.A200[1], "09,10", F.A200[2], "1e4")
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.count("status 0") == 2
    r = run("solver_emulator", d, F.A200[1], "09,00", F.A200[2], "1e4")
    assert r.returncode == 1
