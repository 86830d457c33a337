#!/usr/bin/env python3
"""The element cases of axtb's reference matrix that the GPU engine covers (every atom an
element whose GFN2 basis has only s and p shells), as inputs and axtb's own results.

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
ROOT = HERE.parent.parent


def sp_only_elements():
    """Elements whose GFN2 basis has only s and p shells (the engine's integral kernels), from
    the generated parameter header."""
    import re
    text = (ROOT / "axtb_amd" / "csrc" / "gfn2_tables.h").read_text()

    def table(name):
        body = re.search(name + r"\[[^=]*=\s*\{(.*?)\};", text, re.S).group(1)
        return [int(v) for v in re.findall(r"-?\d+", body)]
    nsh, ang = table("GFN2_NSHELL"), table("GFN2_ANG_SHELL")
    return [z for z in range(1, 87) if all(ang[3 * (z - 1) + s] <= 1 for s in range(nsh[z - 1]))]


results = json.loads((REF / "tests/reference_matrix/results.json").read_text())["cases"]
sp = set(sp_only_elements())
out = {}
for c in cases.elements():
    r = results.get(c["id"], {}).get("programs", {}).get("axtb")
    if not r or not r.get("ok") or not set(c["numbers"]) <= sp:
        continue
    out[c["id"]] = {"label": c["label"], "numbers": c["numbers"], "positions": c["positions"], "charge": c["charge"],
                    "uhf": c["uhf"], "multiplicity": c.get("multiplicity"), "accuracy": c["accuracy"],
                    "energy": r["energy"], "iterations": r["iterations"], "terms": r["terms"],
                    "charges": r["charges"], "dipole": r["dipole"]}
(HERE / "elements.json").write_text(json.dumps(out, indent=1, sort_keys=True))
print(len(out), sorted(out))
