"""Warm starts on the GPU: the SCF from given moments (ScfGuess, crates/axtb/src/xtb.rs:204-215,
519-529), the retry from the fresh start when a warm start does not converge
(crates/axtb/src/calculator.rs:185-206, crates/base/src/method/retry.rs:150-158) and the
calculator's restart policy, through the C ABI (axtbhip_gfn2_engrad_batch_guess).

Bars: as the FP64 parity tests, |dE| <= 1e-8 Eh and |dg| <= 1e-7 Eh/bohr against the oracle
started from the same moments, with the same SCF iteration counts; the reference's own counts
for water (calculator_threads.rs:19-41).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WATER = ([8, 1, 1], np.array([[0.0, 0.0, 0.22], [0.0, 1.43, -0.89], [0.0, -1.43, -0.89]]))


def bad_start(nsh, nat):
    """The start of calculator_threads.rs:26-36: shell charges +-4, dipoles (3, -3, 3),
    quadrupoles (5, 0, -5, 0, 0, 0)."""
    return {"shell_charges": np.array([4.0 if i % 2 == 0 else -4.0 for i in range(nsh)]),
            "dipoles": np.array([[3.0, -3.0, 3.0]] * nat),
            "quadrupoles": np.array([[5.0, 0.0, -5.0, 0.0, 0.0, 0.0]] * nat)}


def as_oracle(st):
    return {"qsh": st["shell_charges"], "dipoles": st["dipoles"], "quadrupoles": st["quadrupoles"]}


@pytest.fixture(scope="module")
def ctx():
    from axtb_amd import _lib
    return _lib.Context(0)


def test_water_golden_and_retry(ctx):
    """calculator_threads.rs:15-41: 8 iterations fresh, 11 from the bad start; at max_iter 9 the
    warm start fails and the retry from the EEQ start gives the fresh result."""
    import oracle
    from axtb_amd import _lib
    from axtb_amd.calculator import Calculator, Options
    z, x = WATER
    print("water fresh / bad start", flush=True)
    fresh = ctx.engrad(z, x)
    assert fresh["iterations"] == 8
    st = bad_start(len(fresh["shell_charges"]), 3)
    w = ctx.engrad_batch([(z, x)], _lib.make_opts(), starts=[st])[0]
    o = oracle.engrad(z, x, start=as_oracle(st))
    assert w["status"] == 0 and w["iterations"] == 11 == o["iterations"]
    assert abs(w["energy"] - o["energy"]) <= 1e-8
    assert np.abs(w["gradient"] - o["gradient"]).max() <= 1e-7
    assert abs(w["energy"] - fresh["energy"]) <= 1e-9
    # max_iter 9: the warm start does not converge, the retry does (same start, same bits)
    calc = Calculator(z, x, options=Options(max_iter=9), context=ctx)
    calc.set_wavefunction(st)
    r = calc.singlepoint(x)
    f9 = ctx.engrad(z, x, max_iter=9)
    assert r.electronic.scf.iterations == f9["iterations"] == 8
    assert r.energy == f9["energy"]
    assert np.array_equal(r.gradient, f9["gradient"])


def test_calculator_restart_policy(ctx):
    """The calculator starts each single point from the previous one's moments (restart None:
    previous), "fresh" from the EEQ start every time; errors and non-finite starts."""
    import oracle
    from axtb_amd import _lib
    from axtb_amd.calculator import Calculator, Options
    from axtb_amd import structures as S
    z, x = S.alkane(10, rng=np.random.default_rng(3), noise=0.05)
    print("C10H22 restart policy", flush=True)
    calc = Calculator(z, x, context=ctx)
    a = calc.singlepoint(x)
    wf = calc.wavefunction()
    b = calc.singlepoint(x)  # from its own converged moments
    o = oracle.engrad(z, x, start=as_oracle(wf))
    assert b.electronic.scf.iterations == o["iterations"] <= 2 < a.electronic.scf.iterations
    assert abs(b.energy - o["energy"]) <= 1e-8 and abs(b.energy - a.energy) <= 1e-9
    fresh = Calculator(z, x, options=Options(restart="fresh"), context=ctx)
    fresh.singlepoint(x)
    c = fresh.singlepoint(x)
    assert c.electronic.scf.iterations == a.electronic.scf.iterations and c.energy == a.energy
    # calculator.rs:482-509 / tests/validation.rs:150-172: sizes and finite values
    bad = dict(wf)
    bad["quadrupoles"] = wf["quadrupoles"].copy()
    bad["quadrupoles"][1, 3] = np.nan
    with pytest.raises(_lib.AxtbError) as e:
        calc.set_wavefunction(bad)
    assert e.value.code == _lib.ERR_INPUT
    with pytest.raises(_lib.AxtbError):
        calc.set_wavefunction({**wf, "quadrupoles": wf["quadrupoles"][:-1]})
    # the engine's own check of a non-finite start (C ABI)
    r = ctx.engrad_batch([(z, x)], _lib.make_opts(), starts=[bad])[0]
    assert r["status"] == _lib.ERR_INPUT


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_trajectory_batch(ctx, precision):
    """A trajectory step of a batch: every molecule moved by 0.02 bohr and started from its
    previous geometry's moments. Against the oracle from the same starts (FP64: the parity bars
    and iteration counts; mixed: against FP64 at 1e-6 Eh, 1e-5 Eh/bohr) and against the fresh
    start (the same energies, fewer iterations)."""
    import oracle
    from axtb_amd import _lib
    from axtb_amd import structures as S
    mols = S.batch(256, seed=11)
    rng = np.random.default_rng(5)
    opts = _lib.make_opts(accuracy=0.01, precision=precision)
    opts64 = _lib.make_opts(accuracy=0.01)
    print(f"trajectory batch of 256 ({precision})", flush=True)
    r0 = ctx.engrad_batch(mols, opts64, raise_on_error=True)
    mols1 = [(z, x + rng.normal(scale=0.02, size=x.shape)) for z, x in mols]
    starts = [{k: r[k] for k in ("shell_charges", "dipoles", "quadrupoles")} for r in r0]
    warm = ctx.engrad_batch(mols1, opts, raise_on_error=True, starts=starts)
    fresh = ctx.engrad_batch(mols1, opts64, raise_on_error=True)
    it_w = np.mean([r["iterations"] for r in warm])
    it_f = np.mean([r["iterations"] for r in fresh])
    de = max(abs(w["energy"] - f["energy"]) for w, f in zip(warm, fresh))
    print(f"iterations warm {it_w:.2f} fresh {it_f:.2f}; warm vs fresh max |dE| {de:.1e}", flush=True)
    assert it_w < it_f - 1.0  # the oracle: 9.4 against 11.8 on every 16th molecule
    # the resident batch with the same starts (axtbhip_batch_prepare_guess): the same bits, twice
    ctx.prepare(mols1, opts, starts=starts)
    for _ in range(2):
        ctx.run()
        res = ctx.results(raise_on_error=True)
        assert all(a["energy"] == b["energy"] and np.array_equal(a["gradient"], b["gradient"])
                   and a["iterations"] == b["iterations"] for a, b in zip(res, warm))
    # the same step as an update of the resident batch of the previous geometry
    # (axtbhip_batch_update): the same bits; a mismatched update fails and keeps the batch
    ctx.prepare(mols, opts)
    ctx.run()
    ctx.update(mols1, starts=starts)
    bad = list(mols1)
    bad[3] = (bad[3][0][:-1], bad[3][1][:-1])
    with pytest.raises(_lib.AxtbError) as e:
        ctx.update(bad, starts=None)
    assert e.value.code == _lib.ERR_INPUT
    ctx.run()
    res = ctx.results(raise_on_error=True)
    assert all(a["energy"] == b["energy"] and np.array_equal(a["gradient"], b["gradient"])
               and a["iterations"] == b["iterations"] for a, b in zip(res, warm))
    if precision == "mixed":
        w64 = ctx.engrad_batch(mols1, opts64, raise_on_error=True, starts=starts)
        for w, f in zip(warm, w64):
            assert abs(w["energy"] - f["energy"]) <= 1e-6
            assert np.abs(w["gradient"] - f["gradient"]).max() <= 1e-5
        return
    for i in range(0, 256, 32):
        z, x = mols1[i]
        o = oracle.engrad(z, x, acc=0.01, start=as_oracle(starts[i]))
        assert warm[i]["iterations"] == o["iterations"]
        assert abs(warm[i]["energy"] - o["energy"]) <= 1e-8
        assert np.abs(warm[i]["gradient"] - o["gradient"]).max() <= 1e-7


def test_session_lineages(ctx):
    """Session.compute_many (crates/axtb-py/src/engine.rs:403-470, 836-843): two trajectories of
    C10H22 as two lineages, advanced side by side in GPU batches, give the results of two
    calculators stepped one geometry at a time (each from its previous moments)."""
    from axtb_amd.calculator import Calculator
    from axtb_amd.session import Session
    from axtb_amd import structures as S
    z, x0 = S.alkane(10, rng=np.random.default_rng(7), noise=0.05)
    rng = np.random.default_rng(8)
    traj = {l: [x0 + rng.normal(scale=0.03, size=x0.shape) for _ in range(3)] for l in ("p", "q")}
    print("session of two C10H22 lineages", flush=True)
    reqs = [dict(positions=traj[l][k], lineage=l) for k in range(3) for l in ("p", "q")]
    out = Session(z, context=ctx).compute_many(reqs)
    assert [r["from"] for r in out] == [None, None, "p", "q", "p", "q"]
    for l in ("p", "q"):
        calc = Calculator(z, x0, context=ctx)
        got = [r for q, r in zip(reqs, out) if q["lineage"] == l]
        for x, r in zip(traj[l], got):
            sp = calc.singlepoint(x)
            assert sp.electronic.scf.iterations == r["iterations"]
            assert abs(sp.energy - r["energy"]) <= 1e-12
            assert np.abs(sp.gradient - r["gradient"]).max() <= 1e-12
