import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP engine)")
    config.addinivalue_line("markers", "slow: runs for more than a few seconds on the CPU")


def structure(name):
    from axtb_amd.structures import read_xyz
    return read_xyz(os.path.join(GOLDEN, "structures", name + ".xyz"))


def d4_fixture(name):
    rec, numbers = {}, None
    for line in open(os.path.join(GOLDEN, "d4", name + ".txt")):
        if line.startswith("#") or not line.strip():
            continue
        key, *vals = line.split()
        if key == "numbers":
            numbers = np.array([int(v) for v in vals], dtype=np.int32)
        else:
            rec[key] = np.array([float(v) for v in vals])
    rec["positions"] = rec["positions"].reshape(-1, 3)
    return numbers, rec
