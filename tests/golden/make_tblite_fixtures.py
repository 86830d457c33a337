#!/usr/bin/env python3
"""axtb's results in the tblite conventions (program "axtb[tblite]" of the reference matrix) for
every case of the matrix: molecules of any element H-Rn, gas phase or with ALPB/GBSA solvation,
any charge and spin, electronic temperature and accuracy.

The geometries are generated by the reference's own case builder
(tests/reference_matrix/cases.py, imported from the reference checkout in the build container);
the energies, terms, charges, dipoles and iteration counts are axtb's results in
tests/reference_matrix/results.json. Writes tblite.json here: the molecules once, then per case
its molecule, charge, spin, options and results.

    python tests/golden/make_tblite_fixtures.py [/path/to/reference]
"""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
sys.path.insert(0, str(REF / "tests" / "reference_matrix"))
import cases  # noqa: E402  (the reference's case builder)

HERE = Path(__file__).parent


OPTS = ("accuracy", "etemp", "max_iter", "dispersion", "alpb", "gbsa", "solvation_state")
results = json.loads((REF / "tests/reference_matrix/results.json").read_text())["cases"]
mols, out = {}, {}
for c in cases.all_cases():
    r = results.get(c["id"], {}).get("programs", {}).get("axtb[tblite]")
    if not r or not r.get("ok") or "expect" in c:
        continue
    key = json.dumps([c["numbers"], c["positions"]])
    if key not in mols:
        mols[key] = len(mols)
    out[c["id"]] = {"category": c["category"], "molecule": mols[key], "charge": c["charge"], "uhf": c["uhf"],
                    "options": {k: c[k] for k in OPTS if k in c}, "energy": r["energy"],
                    "iterations": r.get("iterations"), "terms": r.get("terms"), "charges": r.get("charges"),
                    "dipole": r.get("dipole")}
molecules = [None] * len(mols)
for key, i in mols.items():
    numbers, positions = json.loads(key)
    molecules[i] = {"numbers": numbers, "positions": positions}
(HERE / "tblite.json").write_text(json.dumps({"molecules": molecules, "cases": out}, indent=0, sort_keys=True))
print(len(out), "cases,", len(molecules), "molecules")
