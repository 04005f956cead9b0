import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

DATA = os.path.join(ROOT, "tests", "golden", "data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdlg.so on the device)")
    config.addinivalue_line("markers", "slow: minutes of CPU time; runs only with DLG_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("DLG_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU pin; set DLG_SLOW=1 to run")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def read_bin(name: str) -> np.ndarray:
    """int32 pairs of a reference data file (rows = filesize / 8)."""
    a = np.fromfile(os.path.join(DATA, name), dtype=np.int32)
    return a[: a.size // 2 * 2].reshape(-1, 2)


def read_pairs_txt(name: str) -> set:
    out = set()
    with open(os.path.join(DATA, name)) as f:
        for line in f:
            p = line.split()
            if len(p) >= 2:
                out.add((int(p[0]), int(p[1])))
    return out


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
