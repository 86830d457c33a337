"""Host-side mirror of axtb's GFN2 calculator API over the HIP engine.

``Calculator`` follows axtb's ``Calculator`` / ``Method`` (crates/axtb/src/calculator.rs:92-128,
378-478; crates/base/src/method/mod.rs:48-154): the molecule (numbers, charge, uhf) is fixed at
construction, ``singlepoint(positions, want)`` returns a ``SinglePoint`` with the fields of
crates/base/src/method/results.rs:39-230 that the path computes, and errors are raised with axtb's
error kinds (crates/base/src/error.rs:8-24). Positions are in bohr.

Like axtb's calculator, each single point starts from the previous geometry's converged moments
(shell charges, atomic dipoles and quadrupoles) unless the restart policy says "fresh", or from
moments given by ``set_wavefunction``; a warm start that does not converge runs once more from the
initial guess (calculator.rs:185-206, crates/base/src/method/retry.rs:150-158), and an error
forgets the moments (calculator.rs:387-395).
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
    # the SCF start of the next single point (ScfSettings::restart, crates/base/src/method/settings.rs:5-41):
    # "previous" (the default, None) the previous geometry's converged moments, "fresh" the initial guess
    restart: Optional[str] = None
    # implicit solvation (SolvationSpec, crates/axtb/src/solvation.rs:14-22): None, "alpb" or "gbsa",
    # the solvent's name (xtb's names and aliases) and the reference state (gsolv, bar1M, reference)
    solvation: Optional[str] = None
    solvent: Optional[str] = None
    solvation_state: str = "gsolv"


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
        self._start = None
        self._given = False
        self._nsh = None

    def singlepoint(self, positions=None, want=Want.GRADIENT) -> SinglePoint:
        x = self.positions if positions is None else np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        o = self.options
        opts = _lib.make_opts(accuracy=o.accuracy, etemp=o.etemp, max_iter=o.max_iter, damping=o.damping,
                              guess=o.guess, gradient=(want == Want.GRADIENT), precision=o.precision,
                              dispersion=o.dispersion, solvation=o.solvation, solvent=o.solvent,
                              state=o.solvation_state)
        start = self._take_start()
        r = self.ctx.engrad_batch([(self.numbers, x, self.charge, self.uhf)], opts, starts=[start])[0]
        if r["status"]:
            self.reset_wavefunction()
            raise _lib.AxtbError(r["status"], self.ctx.last_error())
        self._start = {k: np.array(r[k]) for k in ("shell_charges", "dipoles", "quadrupoles")}
        dipole = (r["charges"][:, None] * x).sum(0) + r["dipoles"].sum(0)
        el = Electronic(r["shell_charges"], r["dipoles"], r["quadrupoles"], r["orbital_energies"], r["efermi"],
                        ScfReport(r["iterations"], r["residual"], r["energy_change"]))
        return SinglePoint(r["energy"], r["terms"], r["gradient"], r["charges"], dipole, el)

    def _take_start(self):
        """The moments the next SCF starts from (None: the initial guess): given ones always, the
        previous ones when the restart policy reuses them (calculator.rs:185-192)."""
        reuse = self._given or (self.options.restart or "previous") == "previous"
        self._given = False
        start, self._start = self._start, None
        return start if reuse else None

    def wavefunction(self):
        """The moments the next single point may start from, or None (calculator.rs:413-423)."""
        return None if self._start is None else {k: v.copy() for k, v in self._start.items()}

    def set_wavefunction(self, wf):
        """Start the next SCF from the moments of ``wf`` ({"shell_charges", "dipoles", "quadrupoles"});
        they must have this molecule's sizes and be finite (calculator.rs:424-430, 482-509)."""
        self._install(wf)
        self._given = True

    def set_history(self, wf):
        """The moments of ``wf`` as the previous ones (the restart policy decides; None forgets them;
        calculator.rs:432-442)."""
        if wf is None:
            self.reset_wavefunction()
        else:
            self._install(wf)
            self._given = False

    def reset_wavefunction(self):
        """Forget the moments: the next single point starts from the initial guess (calculator.rs:444-447)."""
        self._start = None
        self._given = False

    def _install(self, wf):
        nat = len(self.numbers)
        if self._nsh is None:
            self._nsh = _lib.basis_size(self.numbers)[0]
        d = {k: np.asarray(wf[k], dtype=np.float64) for k in ("shell_charges", "dipoles", "quadrupoles")}
        if d["shell_charges"].size != self._nsh or d["dipoles"].size != 3 * nat or d["quadrupoles"].size != 6 * nat:
            raise _lib.AxtbError(_lib.ERR_INPUT, f"guess needs {self._nsh} shell charges and {nat} atomic "
                                                 f"dipoles/quadrupoles")
        if not all(np.isfinite(v).all() for v in d.values()):
            raise _lib.AxtbError(_lib.ERR_INPUT, "guess with non-finite charges or multipoles")
        self._start = {"shell_charges": d["shell_charges"].reshape(-1).copy(),
                       "dipoles": d["dipoles"].reshape(nat, 3).copy(),
                       "quadrupoles": d["quadrupoles"].reshape(nat, 6).copy()}


def singlepoint(numbers, positions, charge=0.0, uhf=None, options=None, want=Want.GRADIENT):
    """axtb::singlepoint (crates/axtb/src/lib.rs:197-205)."""
    return Calculator(numbers, positions, charge, uhf, options).singlepoint(positions, want)
