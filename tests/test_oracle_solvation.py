"""The oracle's implicit solvation (ALPB / GBSA) pinned by the reference's own data.

- The model alone at xtb's converged charges against the components xtb 6.7.1 prints
  (crates/solvation/tests/data/*.ref, copied to tests/golden/solvation_ref; the reference's bar
  0.6e-12 Eh, crates/solvation/tests/xtb_reference.rs:11-13), in xtb 6.7.1's conventions
  (its Angstrom factor, its reference-state density and its Born-radius quirk).
- The solvated SCF in the axtb conventions (the corrected model) against axtb's results for every
  reference-matrix solvation case (H/C/N/O molecules and the Mg2+ and Fe3+ aqua ions;
  tests/golden/solvation.json, written by
  make_solvation_fixtures.py): energy, the solvation term, charges, dipole, iteration count.
- The gradient of the solvated energy by central differences (the reference's 2e-7 bar,
  crates/axtb/tests/consistency.rs:26-58), and the model's fixed-charge gradient by central
  differences of its energy.
"""
import glob
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

oracle = pytest.importorskip("oracle")

REFS = sorted(glob.glob(os.path.join(GOLDEN, "solvation_ref", "*.ref")))
SOLV = json.load(open(os.path.join(GOLDEN, "solvation.json")))


def parse_ref(path):
    c = {"nang": 230, "atoms": []}
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] in ("case", "model", "solvent", "state"):
            c[t[0]] = t[1]
        elif t[0] == "nang":
            c["nang"] = int(t[1])
        elif t[0] == "atom":
            c["atoms"].append([float(v) for v in t[1:6]])
        elif t[0] in ("gsolv", "gelec", "gsasa", "ghb", "gshift"):
            c[t[0]] = float(t[1])
    a = np.array(c["atoms"])
    c["numbers"], c["positions"], c["charges"] = a[:, 0].astype(np.int32), a[:, 1:4], a[:, 4]
    return c


def test_reference_files_present():
    assert len(REFS) == 74
    assert len(SOLV["cases"]) == 483


@pytest.mark.parametrize("path", REFS, ids=[os.path.basename(p)[:-4] for p in REFS])
def test_model_against_xtb(path):
    c = parse_ref(path)
    r = oracle.solvation_parts(c["numbers"], c["positions"], c["charges"], c["model"], c["solvent"], c["state"],
                               conventions="xtb671", nang=c["nang"])
    for k in ("gsolv", "gelec", "gsasa", "ghb", "gshift"):
        assert abs(r[k] - c[k]) <= 0.6e-12, (k, r[k], c[k])


def test_scf_cases():
    """All 483 cases in one test (about 2 s)."""
    worst = np.zeros(4)
    for cid, c in sorted(SOLV["cases"].items()):
        m = SOLV["molecules"][c["molecule"]]
        r = oracle.engrad(np.array(m["numbers"]), np.array(m["positions"]), charge=float(c["charge"]), uhf=c["uhf"],
                          acc=c["accuracy"], gradient=False, solvation=c["model"], solvent=c["solvent"],
                          state=c["state"])
        d = [abs(r["energy"] - c["energy"]), abs(r["terms"]["solvation"] - c["terms"]["solvation"]),
             np.abs(r["qat"] - np.array(c["charges"])).max() if c["charges"] is not None else 0.0,
             np.abs(r["dipole"] - np.array(c["dipole"])).max()]
        worst = np.maximum(worst, d)
        assert d[0] <= 1e-8 and d[1] <= 1e-8 and d[2] <= 1e-7 and d[3] <= 1e-6, (cid, d)
        assert r["iterations"] == c["iterations"], cid
    print("solvated SCF vs axtb: max |dE| %.2e |dsolv| %.2e |dq| %.2e |ddip| %.2e" % tuple(worst))


def _fd_check(z, x, charge, h=1e-4, **kw):
    r = oracle.engrad(z, x, charge=charge, acc=1e-8, **kw)
    g = r["gradient"]
    worst = 0.0
    rng = np.random.default_rng(5)
    for _ in range(4):
        d = rng.standard_normal(x.shape)
        d /= np.linalg.norm(d)
        ep = oracle.engrad(z, x + h * d, charge=charge, acc=1e-8, gradient=False, **kw)["energy"]
        em = oracle.engrad(z, x - h * d, charge=charge, acc=1e-8, gradient=False, **kw)["energy"]
        worst = max(worst, abs((ep - em) / (2 * h) - np.sum(g * d)))
    return worst


@pytest.mark.parametrize("model,solvent,species,charge", [("alpb", "water", 0, 0.0), ("gbsa", "thf", 1, 1.0),
                                                          ("alpb", "toluene", 2, -1.0)])
def test_gradient_finite_differences(model, solvent, species, charge):
    mols = {json.dumps(m["numbers"]): m for m in SOLV["molecules"]}
    m = sorted(mols.values(), key=lambda m: len(m["numbers"]))[species]
    z, x = np.array(m["numbers"], dtype=np.int32), np.array(m["positions"])
    err = _fd_check(z, x, charge, solvation=model, solvent=solvent)
    print(model, solvent, len(z), "atoms: FD error %.2e" % err)
    assert err <= 2e-7


def test_model_gradient_fixed_charges():
    """gsolv at fixed charges and its gradient (Born kernel, radii chain, surface, H-bond and ALPB
    shape terms) on taxol in ALPB water and caffeine in GBSA water, xtb's charges."""
    for name, conv in (("taxol_alpb_water", "corrected"), ("caffeine_gbsa_water", "corrected"),
                       ("zncl4_alpb_water", "xtb")):
        c = parse_ref(os.path.join(GOLDEN, "solvation_ref", name + ".ref"))
        z, x, q = c["numbers"], c["positions"], c["charges"]
        r = oracle.solvation_parts(z, x, q, c["model"], c["solvent"], c["state"], conventions=conv, gradient=True)
        rng = np.random.default_rng(1)
        h = 1e-4
        worst = 0.0
        for _ in range(3):
            d = rng.standard_normal(x.shape)
            d /= np.linalg.norm(d)
            ep = oracle.solvation_parts(z, x + h * d, q, c["model"], c["solvent"], c["state"], conventions=conv)
            em = oracle.solvation_parts(z, x - h * d, q, c["model"], c["solvent"], c["state"], conventions=conv)
            worst = max(worst, abs((ep["gsolv"] - em["gsolv"]) / (2 * h) - np.sum(r["gradient"] * d)))
        print(name, "fixed-charge gradient FD error %.2e" % worst)
        assert worst <= 1e-8


def test_unknown_solvent_and_model_table():
    m = SOLV["molecules"][0]
    z, x = np.array(m["numbers"]), np.array(m["positions"])
    with pytest.raises(oracle.OracleError) as e:
        oracle.engrad(z, x, solvation="alpb", solvent="unobtainium")
    assert e.value.code == 1
    with pytest.raises(oracle.OracleError) as e:  # aniline has ALPB parameters only
        oracle.engrad(z, x, solvation="gbsa", solvent="aniline")
    assert e.value.code == 1
    # aliases and case of the name (lib.rs:255-271)
    a = oracle.engrad(z, x, acc=1e-6, gradient=False, solvation="alpb", solvent="H2O")["energy"]
    b = oracle.engrad(z, x, acc=1e-6, gradient=False, solvation="alpb", solvent="water")["energy"]
    assert a == b
