#!/usr/bin/env python3
"""The element cases of axtb's reference matrix (one or two molecules for every element H-Rn) and
its random multi-element molecules (MB16-43-style light and heavy-element sets), as inputs and
axtb's own results.

The geometries are generated by the reference's own case builder
(tests/reference_matrix/cases.py, imported from the reference checkout in the build container);
the energies, per-term energies, charges, dipoles and iteration counts are axtb's results in
tests/reference_matrix/results.json (accuracy 1e-6, axtb conventions). Writes elements.json here.

    python tests/golden/make_element_fixtures.py [/path/to/reference]
"""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
sys.path.insert(0, str(REF / "tests" / "reference_matrix"))
import cases  # noqa: E402  (the reference's case builder)

HERE = Path(__file__).parent


results = json.loads((REF / "tests/reference_matrix/results.json").read_text())["cases"]
out = {}
for c in cases.elements() + cases.multielement():
    r = results.get(c["id"], {}).get("programs", {}).get("axtb")
    if not r or not r.get("ok"):
        continue
    out[c["id"]] = {"label": c["label"], "numbers": c["numbers"], "positions": c["positions"], "charge": c["charge"],
                    "uhf": c["uhf"], "multiplicity": c.get("multiplicity"), "accuracy": c["accuracy"],
                    "energy": r["energy"], "iterations": r["iterations"], "terms": r["terms"],
                    "charges": r["charges"], "dipole": r["dipole"]}
(HERE / "elements.json").write_text(json.dumps(out, indent=1, sort_keys=True))
print(len(out), sorted(out))
