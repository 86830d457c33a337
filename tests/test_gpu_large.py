"""The benchmark molecules on the GPU against the oracle's committed results (tests/golden/
oracle_<name>.npz, written by make_oracle_fixtures.py from the same seeded geometries).

Bars (the benchmark's): the FP64 path |dE| <= 1e-8 Eh, max |dg| <= 1e-7 Eh/bohr, the same SCF
iteration count, charges to 1e-7 and orbital energies to 1e-7 Eh; the mixed-precision path
|dE| <= 1e-6 Eh, max |dg| <= 1e-5 Eh/bohr, charges to 1e-5 and orbital energies to 1e-5 Eh; FP64-refine
(the mixed path's refinement in FP64) the FP64 bars on energy and gradient, charges and orbital
energies to 1e-6.
(H2O)20 (60 atoms), C100H202 (302 atoms, n = 602), (H2O)200 (600 atoms, n = 1200) and
(H2O)1000 (3000 atoms, n = 6000).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

BARS = {"fp64": (1e-8, 1e-7, 1e-7, 1e-7), "mixed": (1e-6, 1e-5, 1e-5, 1e-5), "fp64-refine": (1e-8, 1e-7, 1e-6, 1e-6)}


@pytest.fixture(scope="module")
def ctx():
    from axtb_amd import _lib
    return _lib.Context(0)


def fixture(name):
    f = np.load(os.path.join(GOLDEN, f"oracle_{name}.npz"))
    return {k: f[k] for k in f.files}


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp64-refine"])
@pytest.mark.parametrize("name", ["water20", "c100h202", "water200", "water1000"])
def test_against_oracle_fixture(ctx, name, precision):
    from axtb_amd import _lib
    o = fixture(name)
    print(f"[{name} {precision}] n_atoms {len(o['numbers'])}", flush=True)
    r = ctx.engrad_batch([(o["numbers"], o["positions"])],
                         _lib.make_opts(accuracy=float(o["accuracy"]), precision=precision))[0]
    assert r["status"] == 0, ctx.last_error()
    d = [abs(r["energy"] - float(o["energy"])), float(np.abs(r["gradient"] - o["gradient"]).max()),
         float(np.abs(r["charges"] - o["charges"]).max()),
         float(np.abs(np.sort(r["orbital_energies"]) - o["orbital_energies"]).max())]
    print(f"[{name} {precision}] |dE| {d[0]:.2e} |dg| {d[1]:.2e} |dq| {d[2]:.2e} |deps| {d[3]:.2e} "
          f"it {r['iterations']}/{int(o['iterations'])}", flush=True)
    bars = BARS[precision]
    for v, bar, what in zip(d, bars, ("dE", "dg", "dq", "deps")):
        assert v <= bar, f"{what} {v:.3e} > {bar:.0e}"
    if precision == "fp64":
        assert r["iterations"] == int(o["iterations"])
        assert abs(r["efermi"] - float(o["efermi"])) <= 1e-7
