"""Host-side mirror of axtb's GFN2 calculator API over the HIP engine.

``Calculator`` follows axtb's ``Calculator`` / ``Method`` (crates/axtb/src/calculator.rs:92-128,
378-478; crates/base/src/method/mod.rs:48-154): the molecule (numbers, charge, uhf) is fixed at
construction, ``singlepoint(positions, want)`` returns a ``SinglePoint`` with the fields of
crates/base/src/method/results.rs:39-230 that the path computes, and errors are raised with axtb's
error kinds (crates/base/src/error.rs:8-24). Positions are in bohr.
"""
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib


@dataclass
class Options:
    """axtb::Options of a GFN2 calculation (crates/axtb/src/lib.rs:92-153)."""
    accuracy: float = 1.0
    etemp: float = 300.0
    max_iter: int = 250
    damping: float = 0.4
    guess: str = "eeq"
    dispersion: bool = True
    precision: str = "fp64"
    device: int = 0


@dataclass
class ScfReport:
    iterations: int
    residual: float
    energy_change: float
    eigensolver: str = "hip-jacobi"


@dataclass
class Electronic:
    shell_charges: np.ndarray
    atomic_dipoles: np.ndarray
    atomic_quadrupoles: np.ndarray
    levels: np.ndarray
    fermi_level: float
    scf: ScfReport


@dataclass
class SinglePoint:
    energy: float
    terms: Dict[str, float]
    gradient: Optional[np.ndarray]
    charges: np.ndarray
    dipole: np.ndarray
    electronic: Optional[Electronic] = None
    timings: List = field(default_factory=list)

    def term(self, name):
        return self.terms.get(name)


class Want:
    ENERGY = "energy"
    GRADIENT = "gradient"


class Calculator:
    """GFN2-xTB on the GPU for one molecule (numbers, charge, uhf fixed at construction)."""

    def __init__(self, numbers, positions, charge=0.0, uhf=None, options: Optional[Options] = None, context=None):
        self.numbers = np.ascontiguousarray(numbers, dtype=np.int32)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        if len(self.positions) != len(self.numbers):
            raise _lib.AxtbError(_lib.ERR_INPUT, "one position per atom")
        self.charge = float(charge)
        self.uhf = uhf
        self.options = options or Options()
        self.ctx = context or _lib.Context(self.options.device)

    def singlepoint(self, positions=None, want=Want.GRADIENT) -> SinglePoint:
        x = self.positions if positions is None else np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        o = self.options
        opts = _lib.make_opts(accuracy=o.accuracy, etemp=o.etemp, max_iter=o.max_iter, damping=o.damping,
                              guess=o.guess, gradient=(want == Want.GRADIENT), precision=o.precision,
                              dispersion=o.dispersion)
        r = self.ctx.engrad_batch([(self.numbers, x, self.charge, self.uhf)], opts)[0]
        if r["status"]:
            raise _lib.AxtbError(r["status"], self.ctx.last_error())
        dipole = (r["charges"][:, None] * x).sum(0) + r["dipoles"].sum(0)
        el = Electronic(r["shell_charges"], r["dipoles"], r["quadrupoles"], r["orbital_energies"], r["efermi"],
                        ScfReport(r["iterations"], r["residual"], r["energy_change"]))
        return SinglePoint(r["energy"], r["terms"], r["gradient"], r["charges"], dipole, el)


def singlepoint(numbers, positions, charge=0.0, uhf=None, options=None, want=Want.GRADIENT):
    """axtb::singlepoint (crates/axtb/src/lib.rs:197-205)."""
    return Calculator(numbers, positions, charge, uhf, options).singlepoint(positions, want)
