import gzip
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def mats(tmp_path_factory):
    """The reference's MatrixMarket fixtures (src/mats), unpacked to a temp dir.
    Returns a function rel_path -> absolute path of the plain .mtx file."""
    out = tmp_path_factory.mktemp("mats")
    src = os.path.join(GOLDEN, "mats")
    for d in sorted(os.listdir(src)):
        os.makedirs(out / d, exist_ok=True)
        for f in sorted(os.listdir(os.path.join(src, d))):
            with gzip.open(os.path.join(src, d, f), "rb") as fi, open(out / d / f[:-3], "wb") as fo:
                shutil.copyfileobj(fi, fo)
    return lambda rel: str(out / rel)
