#!/usr/bin/env python3
"""Oracle results of the large benchmark molecules, for the GPU parity tests.

Runs the C++ FP64 oracle (oracle/liboracle.so, OpenMP) once on the seeded geometries of the
benchmark configuration and writes, per molecule, ``oracle_<name>.npz`` here: numbers and
positions (bohr, the exact input), energy, gradient, atomic charges, orbital energies, the Fermi
level and the SCF iteration count at accuracy 1e-2 (SCC energy convergence 1e-8 Eh), 300 K, D4,
EEQ start. The GPU tests compare against these files, so the oracle need not run on the GPU box
for molecules it takes minutes to compute.

    python tests/golden/make_oracle_fixtures.py water200 c100h202 water1000
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from axtb_amd import structures  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def make(name, acc=1e-2):
    z, x = structures.config(name)
    t0 = time.perf_counter()
    o = oracle.engrad(z, x, acc=acc)
    dt = time.perf_counter() - t0
    path = os.path.join(HERE, f"oracle_{name}.npz")
    np.savez_compressed(path, numbers=z, positions=x, energy=o["energy"], gradient=o["gradient"], charges=o["qat"],
                        orbital_energies=o["emo"], efermi=o["efermi"], iterations=o["iterations"], accuracy=acc,
                        terms=np.array([o["terms"][k] for k in oracle.TERMS]))
    print(f"{name}: {len(z)} atoms, E {o['energy']:.12f} Eh, {o['iterations']} iterations, {dt:.0f} s -> {path}",
          flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or ["water200", "c100h202"]:
        make(n)
