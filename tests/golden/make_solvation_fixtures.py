#!/usr/bin/env python3
"""The implicit-solvation cases of axtb's reference matrix, as inputs and axtb's own results.

The geometries are generated by the reference's own case builder
(tests/reference_matrix/cases.py, imported from the reference checkout in the build container);
the energies, per-term energies (solvation among them), charges, dipoles and iteration counts are
axtb's results in tests/reference_matrix/results.json (accuracy 1e-6, axtb conventions: the
corrected ALPB/GBSA model). Writes solvation.json here: the molecules once, then per case the
model, solvent, reference state, molecule, charge, spin and results.

    python tests/golden/make_solvation_fixtures.py [/path/to/reference]
"""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
sys.path.insert(0, str(REF / "tests" / "reference_matrix"))
import cases  # noqa: E402  (the reference's case builder)

HERE = Path(__file__).parent

results = json.loads((REF / "tests/reference_matrix/results.json").read_text())["cases"]
mols, out = {}, {}
for c in cases.solvation():
    r = results.get(c["id"], {}).get("programs", {}).get("axtb")
    if not r or not r.get("ok"):
        continue
    model = "alpb" if "alpb" in c else "gbsa"
    key = json.dumps([c["numbers"], c["positions"]])
    if key not in mols:
        mols[key] = len(mols)
    out[c["id"]] = {"model": model, "solvent": c[model], "state": c["solvation_state"], "molecule": mols[key],
                    "charge": c["charge"], "uhf": c["uhf"], "accuracy": c["accuracy"], "energy": r["energy"],
                    "iterations": r["iterations"], "terms": r["terms"], "charges": r["charges"],
                    "dipole": r["dipole"]}
molecules = [None] * len(mols)
for key, i in mols.items():
    numbers, positions = json.loads(key)
    molecules[i] = {"numbers": numbers, "positions": positions}
(HERE / "solvation.json").write_text(json.dumps({"molecules": molecules, "cases": out}, indent=0, sort_keys=True))
print(len(out), "cases,", len(molecules), "molecules")
