import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")

# Element and multi-element cases of axtb's reference matrix (tests/golden/elements.json) whose
# iteration count the reference's round-off sets: the oracle's count changes when the molecule is
# rotated (ReO4-, tetrahedral: 14 or 15; mb16-04: 60 or 61; mbheavy-113: 149-153), the SCF runs for
# hundreds of iterations (mbheavy-103: 212, mbheavy-112: 170), the last step is borderline
# (Ni(CO)4: residual 2.23e-11 at iteration 17 against the bound 2e-11, converged at 18), or a
# symmetry-breaking component that the reference's round-off seeds decays slowly before Broyden
# removes it (tetrahedral OsO4: residual 1.2e-9, 7.1e-10, then 2.7e-13 at iteration 14; arithmetic
# that is symmetric between the equivalent atoms does not seed it and stops at 13). Energies,
# charges and dipoles are still checked.
ELEMENT_ROUNDOFF_ITERATIONS = {"el-75-Re", "el-28-Ni", "el-76-Os", "mb16-04", "mbheavy-103", "mbheavy-112",
                               "mbheavy-113"}
# mbheavy-101: the reference's Broyden converges in 249 of its 250 iterations; a run that is one
# iteration slower ends ScfNotConverged, so the case does not pin anything.
ELEMENT_EDGE_OF_MAX_ITER = {"mbheavy-101"}


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
