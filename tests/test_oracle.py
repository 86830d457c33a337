"""The CPU oracle against axtb's committed reference data, closed forms and finite differences.

Golden data (tests/golden): axtb's energies, per-term energies, charges and iteration counts
(reference_matrix/results.json at accuracy 1e-6; the axtb CLI runs of bench/results/final-9985wx
at accuracy 1), D4 fixtures of dftd4 4.3.1 / tblite 0.8.0 (crates/d4/tests/fixtures).
"""
import json
import os

import numpy as np
import pytest

import oracle
from axtb_amd import structures
from conftest import ELEMENT_EDGE_OF_MAX_ITER, ELEMENT_ROUNDOFF_ITERATIONS, GOLDEN, d4_fixture, structure

GOLD = json.load(open(os.path.join(GOLDEN, "golden.json")))


@pytest.mark.parametrize("name", ["water", "caffeine", "taxol"])
def test_energy_terms_charges_acc1e6(name):
    g = GOLD["acc1e-6"][name]
    z, x = structure(name)
    r = oracle.engrad(z, x, acc=1e-6, gradient=False)
    assert abs(r["energy"] - g["energy"]) < 1e-9
    assert r["iterations"] == g["iterations"]
    for k in ("repulsion", "dispersion", "electronic", "isotropic_es", "third_order", "anisotropic", "entropy"):
        assert abs(r["terms"][k] - g["terms"][k]) < 1e-9, k
    if g["charges"] is not None:
        assert np.abs(r["qat"] - np.array(g["charges"])).max() < 1e-8
    dip = (r["qat"][:, None] * x).sum(0) + r["dipoles"].sum(0)
    assert np.abs(dip - np.array(g["dipole"])).max() < 1e-7


@pytest.mark.parametrize("name", ["water", "caffeine", "C20H42", "taxol"])
def test_energy_acc1(name):
    g = GOLD["acc1"][name]
    z, x = structures.alkane_bench(20) if name == "C20H42" else structure(name)
    r = oracle.engrad(z, x, acc=1.0, gradient=False)
    assert r["iterations"] == g["iterations"]
    assert abs(r["energy"] - g["energy"]) < 1e-9


@pytest.mark.slow
@pytest.mark.parametrize("name", ["C50H102", "(H2O)100", "C100H202", "(H2O)200"])
def test_energy_acc1_large(name):
    g = GOLD["acc1"][name]
    files = {"C50H102": "c50h102", "(H2O)100": "water100", "(H2O)200": "water200"}
    z, x = structures.alkane_bench(100) if name == "C100H202" else structure(files[name])
    r = oracle.engrad(z, x, acc=1.0, gradient=False)
    assert r["iterations"] == g["iterations"]
    assert abs(r["energy"] - g["energy"]) < 1e-8


@pytest.mark.parametrize("name", ["water", "benzene", "alkane_c40"])
def test_d4_fixtures(name):
    z, rec = d4_fixture(name)
    x = rec["positions"]
    q = rec["d4.q"]
    # EEQ-2019 charges of dftd4 (multicharge) = the SCF start's EEQ model
    assert np.abs(oracle.eeq(z, x) - q).max() < 1e-9
    for cset in ("dftd4", "tbcut", "wide"):
        c = rec[cset + ".cutoff"]  # disp2 disp3 cn width2 width3
        r = oracle.d4(z, x, q, charge_model="eeq", radii="mctc", cutoffs=(c[2], c[0], c[1], c[3], c[4]))
        assert abs(r["e2"] - rec[cset + ".e2"][0]) <= 1e-12 * abs(rec[cset + ".e2"][0])
        assert abs(r["e3"] - rec[cset + ".e3"][0]) <= 1e-11 * abs(rec[cset + ".e3"][0])
        if cset == "dftd4":
            n = len(z)
            r30 = oracle.d4(z, x, q, charge_model="eeq", radii="mctc", cutoffs=(30.0, c[0], c[1], c[3], c[4]))
            assert np.abs(r30["cn"] - rec["d4.cn"]).max() <= 1e-12 * np.abs(rec["d4.cn"]).max()
            assert np.abs(r30["c6"].ravel() - rec["d4.c6"]).max() <= 1e-12 * np.abs(rec["d4.c6"]).max()
            assert np.abs(r30["alpha"] - rec["d4.alpha"]).max() <= 1e-12 * np.abs(rec["d4.alpha"]).max()
    if "tblite_gfn2.e3" in rec:
        # tblite's GFN2 D4 (GFN2 reference charges, tblite's cutoffs): three-body energy
        r = oracle.d4(z, x, np.zeros(len(z)), charge_model="gfn2", radii="mctc",
                      cutoffs=(30.0, 50.0, 25.0, 0.05, 0.05))
        ref = rec["tblite_gfn2.e3"][0]
        assert abs(r["e3"] - ref) <= 1e-12 * abs(ref)
        r = oracle.d4(z, x, rec["tblite_gfn2.q"], charge_model="gfn2", radii="mctc",
                      cutoffs=(30.0, 50.0, 25.0, 0.05, 0.05))
        ref = rec["tblite_gfn2.e2_hf"][0]
        assert abs(r["e2"] - ref) <= 1e-8 * abs(ref)


def test_d4_gradient_fd():
    """Two- and three-body D4 gradient at fixed charges against central differences."""
    z, rec = d4_fixture("water")
    x = rec["positions"]
    q = rec["d4.q"]
    cut = (40.0, 72.0, 40.0, 1.0, 1.0)
    g = oracle.d4(z, x, q, charge_model="gfn2", radii="codata", cutoffs=cut, gradient=True)["gradient"]
    h = 1e-5
    for i in range(len(z)):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            ep = oracle.d4(z, xp, q, "gfn2", "codata", cut)
            em = oracle.d4(z, xm, q, "gfn2", "codata", cut)
            fd = ((ep["e2"] + ep["e3"]) - (em["e2"] + em["e3"])) / (2 * h)
            assert abs(fd - g[i, k]) < 1e-9


def test_tblite_scc_charges():
    """GFN2 Mulliken charges of tblite 0.8.0 (d4 fixture) at ~1e-6: axtb's conventions differ
    from tblite's only in constants (CODATA 2022 radii, kB) and cutoffs."""
    z, rec = d4_fixture("water")
    r = oracle.engrad(z, rec["positions"], acc=1e-6, gradient=False)
    assert np.abs(r["qat"] - rec["tblite_gfn2.q"]).max() < 1e-5


def test_gradient_fd_water():
    """Analytic gradient against central differences, h = 1e-4 bohr (tests/consistency.rs:26-58)."""
    z, x = structure("water")
    r = oracle.engrad(z, x, acc=1e-8)
    g = r["gradient"]
    assert np.abs(g.sum(0)).max() < 1e-12
    h = 1e-4
    fd = np.zeros_like(g)
    for i in range(len(z)):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd[i, k] = (oracle.engrad(z, xp, acc=1e-8, gradient=False)["energy"]
                        - oracle.engrad(z, xm, acc=1e-8, gradient=False)["energy"]) / (2 * h)
    assert np.abs(fd - g).max() < 2e-7


def test_gradient_fd_alkane_noise():
    z, x = structures.alkane(3, rng=np.random.default_rng(3), noise=0.1)
    r = oracle.engrad(z, x, acc=1e-8)
    g = r["gradient"]
    h = 1e-4
    for (i, k) in [(0, 0), (1, 2), (4, 1), (7, 0), (10, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        fd = (oracle.engrad(z, xp, acc=1e-8, gradient=False)["energy"]
              - oracle.engrad(z, xm, acc=1e-8, gradient=False)["energy"]) / (2 * h)
        assert abs(fd - g[i, k]) < 2e-7


def test_invariants_charged():
    """Sum of charges = molecular charge; rotating the molecule leaves E unchanged and rotates g."""
    z, x = structure("caffeine")
    r = oracle.engrad(z, x, charge=1.0, acc=1e-6)
    assert abs(r["qat"].sum() - 1.0) < 1e-10
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    r2 = oracle.engrad(z, x @ R.T, charge=1.0, acc=1e-6)
    assert abs(r2["energy"] - r["energy"]) < 1e-9
    assert np.abs(r2["gradient"] - r["gradient"] @ R.T).max() < 1e-7


def test_self_overlap_and_eigh():
    z, x = structure("caffeine")
    S, _ = oracle.overlap(z, x)
    assert np.abs(np.diag(S) - 1.0).max() < 1e-9  # STO-nG contraction norm (integrals.rs:1021-1034)
    assert np.abs(S - S.T).max() < 1e-15
    w, V = oracle.eigh(S)
    assert np.abs(S @ V - V * w).max() < 1e-12
    assert np.abs(V.T @ V - np.eye(len(w))).max() < 1e-12
    assert np.allclose(w, np.linalg.eigvalsh(S), atol=1e-12)


def test_errors():
    z, x = structure("water")
    with pytest.raises(oracle.OracleError) as e:
        oracle.engrad(z, x, charge=9.0)
    assert e.value.code == 1
    with pytest.raises(oracle.OracleError) as e:
        oracle.engrad(z, x, max_iter=2)
    assert e.value.code == 3
    with pytest.raises(oracle.OracleError) as e:
        oracle.engrad(z, x, charge=0.5)
    assert e.value.code == 1


WATER_RS = ([8, 1, 1], np.array([[0.0, 0.0, 0.22], [0.0, 1.43, -0.89], [0.0, -1.43, -0.89]]))


def bad_start(nsh, nat):
    """The start the reference's calculator test cannot recover from in 9 iterations
    (crates/axtb/tests/calculator_threads.rs:26-36)."""
    return {"qsh": np.array([4.0 if i % 2 == 0 else -4.0 for i in range(nsh)]),
            "dipoles": np.array([[3.0, -3.0, 3.0]] * nat), "quadrupoles": np.array([[5.0, 0.0, -5.0, 0.0, 0.0, 0.0]] * nat)}


def test_warm_start_golden_iterations():
    """A given SCF start (xtb.rs:519-529) against the counts of the reference's own test
    (calculator_threads.rs:19): water converges in 8 iterations from the EEQ start and needs 11
    from the bad start (so it fails at max_iter 9), to the same energy; its own converged moments
    are a fixed point (2 iterations, the same energy)."""
    z, x = WATER_RS
    f = oracle.engrad(z, x)
    assert f["iterations"] == 8
    b = oracle.engrad(z, x, start=bad_start(len(f["qsh"]), 3))
    assert b["iterations"] == 11
    assert abs(b["energy"] - f["energy"]) < 1e-9
    with pytest.raises(oracle.OracleError) as e:
        oracle.engrad(z, x, start=bad_start(len(f["qsh"]), 3), max_iter=9)
    assert e.value.code == 3
    w = oracle.engrad(z, x, start={"qsh": f["qsh"], "dipoles": f["dipoles"], "quadrupoles": f["quadrupoles"]})
    assert w["iterations"] <= 2
    assert abs(w["energy"] - f["energy"]) < 1e-10
    # a start that drops the multipoles is no fixed point: more iterations, the same energy
    q = oracle.engrad(z, x, start={"qsh": f["qsh"], "dipoles": 0 * f["dipoles"], "quadrupoles": 0 * f["quadrupoles"]})
    assert q["iterations"] > w["iterations"]
    assert abs(q["energy"] - f["energy"]) < 1e-9


def test_mixer_plan_stretched_water():
    """The reference's golden case for the mixer plan (crates/axtb/tests/mixer.rs:56-92): water with
    both O-H bonds at 3 A, acc 1e-4. Broyden oscillates for all 250 iterations; DIIS from the start
    converges to -4.6080231096 Eh; auto is DIIS's result, bit for bit, reported as diis-fallback.
    With max_iter 5 both attempts fail and the error counts the iterations of both."""
    z, x = structure("scan-H2O-3.0")
    rep = {}
    with pytest.raises(oracle.OracleError) as e:
        oracle.engrad(z, x, acc=1e-4, mixer="broyden", report=rep)
    assert e.value.code == 3 and rep["attempts"] == [("broyden", 250, False)]
    d = oracle.engrad(z, x, acc=1e-4, mixer="diis")
    a = oracle.engrad(z, x, acc=1e-4)
    assert a["energy"] == d["energy"] and np.array_equal(a["gradient"], d["gradient"])
    assert a["iterations"] == d["iterations"]
    assert abs(a["energy"] - -4.6080231096) < 1e-8
    assert d["mixer_used"] == "diis" and a["mixer_used"] == "diis-fallback"
    assert a["attempts"] == [("broyden", 250, False), ("diis", d["iterations"], True)]
    rep = {}
    with pytest.raises(oracle.OracleError) as e:
        oracle.engrad(z, x, acc=1e-4, max_iter=5, gradient=False, report=rep)
    assert e.value.code == 3 and rep["iterations"] == 10


def test_mixer_auto_is_broyden_where_broyden_converges():
    """mixer.rs:37-53: auto equals Broyden bit for bit where Broyden converges, and DIIS alone lands
    on the same solution."""
    z, x = structure("caffeine")
    b = oracle.engrad(z, x, acc=1e-6, mixer="broyden")
    a = oracle.engrad(z, x, acc=1e-6)
    assert a["energy"] == b["energy"] and np.array_equal(a["gradient"], b["gradient"])
    assert a["mixer_used"] == "broyden" and a["attempts"] == [("broyden", b["iterations"], True)]
    d = oracle.engrad(z, x, acc=1e-6, mixer="diis")
    assert abs(d["energy"] - b["energy"]) < 1e-10 and d["mixer_used"] == "diis"


def test_diis_linear_fixed_point_and_coefficients():
    """scf/mixing.rs:523-577: on q = A q + b DIIS converges in about as many steps as A has distinct
    eigenvalues, also where plain damped mixing diverges (eigenvalue -4); the Pulay coefficients
    sum to one and cancel the residuals; a residual parallel to the newest one gets no weight."""
    eig = np.array([0.9, 0.5, -0.5, -0.95, -4.0, 0.2])
    b = np.array([1.0, -2.0, 0.5, 3.0, -1.0, 0.25])
    steps, xq = oracle.diis_linear(eig, b, 0.4, max_steps=2 * len(eig))
    assert 0 <= steps <= 2 * len(eig)
    assert np.abs(xq - b / (1.0 - eig)).max() < 1e-10
    c = oracle.diis_coefficients([[1.0, 0.0], [-3.0, 0.0]])
    assert abs(c.sum() - 1.0) < 1e-14 and abs(c[0] * 1.0 + c[1] * -3.0) < 1e-14
    c = oracle.diis_coefficients([[2.0, 2.0], [1.0, 1.0], [1.0, 1.0]])
    assert c[1] == 0.0 and abs(c.sum() - 1.0) < 1e-14


def test_engrad_many_matches_single():
    """The batch entry (one molecule per thread) gives each molecule's single-point result."""
    mols = structures.batch(6, seed=3)
    e, g, st = oracle.engrad_many(mols, acc=1e-2)
    assert (st == 0).all()
    off = 0
    for i, (z, x) in enumerate(mols):
        r = oracle.engrad(z, x, acc=1e-2)
        assert abs(e[i] - r["energy"]) < 1e-10
        assert np.abs(g[off:off + len(z)] - r["gradient"]).max() < 1e-10
        off += len(z)


ELEMENTS = json.load(open(os.path.join(GOLDEN, "elements.json")))


@pytest.mark.parametrize("cid", sorted(set(ELEMENTS) - ELEMENT_EDGE_OF_MAX_ITER))
def test_element_cases(cid):
    """Every element H-Rn (s, p and d shells) in axtb's reference-matrix molecules, and its random
    light and heavy-element molecules (tests/reference_matrix/results.json, accuracy 1e-6, charges
    and unpaired electrons as given): energy, iterations, charges, dipole. Pins the parameter
    tables, the STO-nG expansions and the d-shell integrals, which the engine shares with the
    oracle or restates."""
    c = ELEMENTS[cid]
    z = np.array(c["numbers"], dtype=np.int32)
    x = np.array(c["positions"])
    r = oracle.engrad(z, x, charge=float(c["charge"]), uhf=c["uhf"], acc=c["accuracy"], gradient=False)
    assert abs(r["energy"] - c["energy"]) < 1e-8, (r["energy"], c["energy"])
    if cid not in ELEMENT_ROUNDOFF_ITERATIONS:
        assert r["iterations"] == c["iterations"]
    if c["charges"] is not None:
        assert np.abs(r["qat"] - np.array(c["charges"])).max() < 1e-7
    if c["dipole"] is not None:
        assert np.abs(r["dipole"] - np.array(c["dipole"])).max() < 1e-6
