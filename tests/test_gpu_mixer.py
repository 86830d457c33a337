"""The SCF mixer plan, the SCF report and the error paths of the engine, on the GPU.

The reference's golden case (crates/axtb/tests/mixer.rs:56-92): water with both O-H bonds at
3 A, acc 1e-4; Broyden does not converge in 250 iterations, DIIS from the start converges to
-4.6080231096 Eh, and auto (the default) reports that result as diis-fallback.
"""
import os

import numpy as np
import pytest

from conftest import structure

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from axtb_amd import _lib
    return _lib.Context(0)


def test_stretched_water_falls_back_to_diis(ctx):
    import oracle
    from axtb_amd import _lib
    z, x = structure("scan-H2O-3.0")
    print("[mixer] stretched water, broyden alone", flush=True)
    b = ctx.engrad_batch([(z, x, 0.0, 0)], _lib.make_opts(accuracy=1e-4, mixer="broyden"))[0]
    assert b["status"] == _lib.ERR_SCF and b["iterations"] == 250
    assert b["attempts"] == [("broyden", 250, False)]
    print("[mixer] stretched water, diis alone and auto", flush=True)
    d = ctx.engrad_batch([(z, x, 0.0, 0)], _lib.make_opts(accuracy=1e-4, mixer="diis"))[0]
    a = ctx.engrad_batch([(z, x, 0.0, 0)], _lib.make_opts(accuracy=1e-4))[0]
    o = oracle.engrad(z, x, acc=1e-4)
    print(f"[mixer] auto E {a['energy']:.10f} it {a['iterations']} (oracle {o['energy']:.10f} it {o['iterations']})",
          flush=True)
    assert d["status"] == 0 and a["status"] == 0
    assert a["energy"] == d["energy"] and np.array_equal(a["gradient"], d["gradient"])
    assert abs(a["energy"] - -4.6080231096) <= 1e-8
    assert abs(a["energy"] - o["energy"]) <= 1e-8
    assert np.abs(a["gradient"] - o["gradient"]).max() <= 1e-7
    assert d["mixer_used"] == "diis" and a["mixer_used"] == "diis-fallback"
    assert a["attempts"] == [("broyden", 250, False), ("diis", d["iterations"], True)]
    assert a["iterations"] == d["iterations"]
    # both attempts failing: the error counts the iterations of both (mixer.rs:84-91)
    f = ctx.engrad_batch([(z, x, 0.0, 0)], _lib.make_opts(accuracy=1e-4, max_iter=5, gradient=False))[0]
    assert f["status"] == _lib.ERR_SCF and f["iterations"] == 10


def test_auto_is_broyden_where_broyden_converges(ctx):
    """mixer.rs:37-53, in a batch with a molecule that needs the fallback: the molecules Broyden
    converges are bit for bit Broyden's; the molecule that falls back matches the oracle."""
    import oracle
    from axtb_amd import _lib
    zc, xc = structure("caffeine")
    zw, xw = structure("scan-H2O-3.0")
    mols = [(zc, xc), (zw, xw, 0.0, 0)]
    a = ctx.engrad_batch(mols, _lib.make_opts(accuracy=1e-4))
    b = ctx.engrad_batch(mols, _lib.make_opts(accuracy=1e-4, mixer="broyden"))
    assert a[0]["energy"] == b[0]["energy"] and np.array_equal(a[0]["gradient"], b[0]["gradient"])
    assert a[0]["mixer_used"] == "broyden" and a[0]["attempts"] == [("broyden", b[0]["iterations"], True)]
    assert a[1]["mixer_used"] == "diis-fallback" and b[1]["status"] == _lib.ERR_SCF
    assert abs(a[1]["energy"] - oracle.engrad(zw, xw, acc=1e-4)["energy"]) <= 1e-8
    d = ctx.engrad_batch(mols[:1], _lib.make_opts(accuracy=1e-6, mixer="diis"))[0]
    o = oracle.engrad(zc, xc, acc=1e-6)
    assert abs(d["energy"] - o["energy"]) <= 1e-8 and d["mixer_used"] == "diis"


def test_report_fields(ctx):
    """SinglePoint.dipole, the HOMO-LUMO gap and the offending element of an unsupported molecule."""
    import oracle
    from axtb_amd import _lib
    z, x = structure("caffeine")
    r = ctx.engrad(z, x, accuracy=1e-6)
    o = oracle.engrad(z, x, acc=1e-6)
    assert np.abs(r["dipole"] - o["dipole"]).max() <= 1e-7
    assert abs(r["gap"] - o["gap"]) <= 1e-7
    bad = ctx.engrad_batch([(np.array([88, 1], dtype=np.int32), np.array([[0.0, 0, 0], [0, 0, 3.0]]))],
                           _lib.make_opts())[0]
    assert bad["status"] == _lib.ERR_ELEMENT and bad["element"] == 88


def test_single_entry_with_start(ctx):
    """axtbhip_gfn2_engrad_guess: a warm start of one molecule equals the batch entry's."""
    from axtb_amd import _lib
    z, x = structure("caffeine")
    r0 = ctx.engrad(z, x, accuracy=1e-2)
    st = {k: r0[k] for k in ("shell_charges", "dipoles", "quadrupoles")}
    x1 = x + 0.01
    a = ctx.engrad(z, x1, accuracy=1e-2, start=st)
    b = ctx.engrad_batch([(z, x1)], _lib.make_opts(accuracy=1e-2), starts=[st])[0]
    assert a["energy"] == b["energy"] and a["iterations"] == b["iterations"]
    assert a["iterations"] < r0["iterations"]


def test_unconverged_jacobi_is_a_numerical_error(monkeypatch):
    """A batched FP64 Jacobi solve that does not converge within its sweeps (here the limit is
    lowered to one sweep, read when the context is created; every kernel still ends normally)
    fails its molecule with Error::Numerical instead of continuing with unconverged vectors."""
    from axtb_amd import _lib
    monkeypatch.setenv("AXTB_HIP_JACOBI_SWEEPS", "1")
    c = _lib.Context(0)
    z, x = structure("caffeine")
    r = c.engrad_batch([(z, x)], _lib.make_opts(accuracy=1e-2))[0]
    assert r["status"] == _lib.ERR_NUMERICAL
    assert "eigensolver" in c.last_error()
    c.close()


def test_resident_batch_is_invalidated_by_a_one_shot_call():
    """prepare(4) then a one-shot batch of 16 on the same context: the resident calls report
    ERR_INPUT (no stale or out-of-range results); results() after update() needs a run()."""
    from axtb_amd import _lib
    from axtb_amd import structures as S
    c = _lib.Context(0)
    mols = S.batch(16, seed=5)
    opts = _lib.make_opts(accuracy=1e-2)
    c.prepare(mols[:4], opts)
    c.run()
    assert all(r["status"] == 0 for r in c.results())
    c.engrad_batch(mols, opts)
    for call in (c.run, c.results, c.gradient):
        with pytest.raises(_lib.AxtbError) as e:
            call()
        assert e.value.code == _lib.ERR_INPUT
    c.prepare(mols[:4], opts)
    c.run()
    c.update([(z, np.asarray(x) + 0.01) for z, x in mols[:4]])
    with pytest.raises(_lib.AxtbError) as e:
        c.results()
    assert e.value.code == _lib.ERR_INPUT
    with pytest.raises(_lib.AxtbError):
        c.update([(z, x, 1.0) for z, x in mols[:4]])  # another charge
    c.close()
