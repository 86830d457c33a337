#!/usr/bin/env python3
"""Collect axtb's committed reference numbers for the GFN2 E+grad path into golden.json.

Reads (as data) from the reference checkout:
  tests/reference_matrix/results.json      axtb energies, per-term energies, charges, dipoles,
                                           iteration counts at accuracy 1e-6 (struct-* cases)
  bench/results/final-9985wx/gfn2-eg-t1.json  axtb CLI energies + iterations at default accuracy 1
and copies nothing else. The structures these refer to are in structures/ (copies of the
reference's tests/structures files); the generated alkanes are rebuilt by tests/geometry.py.
"""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
out = {"acc1e-6": {}, "acc1": {}}
cases = json.loads((REF / "tests/reference_matrix/results.json").read_text())["cases"]
for cid in ("struct-water", "struct-caffeine", "struct-taxol"):
    a = cases[cid]["programs"]["axtb"]
    out["acc1e-6"][cid.split("-", 1)[1]] = {k: a[k] for k in ("energy", "iterations", "terms", "charges", "dipole")}
runs = json.loads((REF / "bench/results/final-9985wx/gfn2-eg-t1.json").read_text())["runs"]
for r in runs:
    if r.get("program") == "axtb CLI" and r.get("ok"):
        out["acc1"][r["molecule"]] = {"energy": r["energy"], "iterations": r["iterations"]}
(Path(__file__).parent / "golden.json").write_text(json.dumps(out, indent=1, sort_keys=True))
print(sorted(out["acc1"]), sorted(out["acc1e-6"]))
