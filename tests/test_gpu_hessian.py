"""The semi-numerical Hessian (axtbhip_gfn2_hessian: 6 N displaced gradients as one GPU batch,
each SCF started from the reference's moments) against the same construction on the CPU oracle,
and against what the mathematics fixes: symmetry, the translational sum rule (each row of a
Hessian sums to zero over the atoms), and the diagonal against second differences of energies."""
import numpy as np
import pytest

from conftest import structure

pytestmark = pytest.mark.gpu


def oracle_hessian(z, x, step=0.005, acc=0.01):
    """Central differences of the oracle's analytic gradients, every coordinate forward by the step
    and back by twice the step, each SCF from the reference geometry's moments (fd.rs:163-209,
    calculator.rs:266-296)."""
    import oracle
    ref = oracle.engrad(z, x, acc=acc)
    st = {"qsh": ref["qsh"], "dipoles": ref["dipoles"], "quadrupoles": ref["quadrupoles"]}
    n3 = 3 * len(z)
    H = np.zeros((n3, n3))
    for y in range(n3):
        xp = np.array(x, dtype=float).reshape(-1)
        xp[y] = xp[y] + step
        xm = xp.copy()
        xm[y] = xm[y] - 2.0 * step
        gp = oracle.engrad(z, xp.reshape(-1, 3), acc=acc, start=st)["gradient"].reshape(-1)
        gm = oracle.engrad(z, xm.reshape(-1, 3), acc=acc, start=st)["gradient"].reshape(-1)
        H[y] = (gp - gm) / (2.0 * step)
    return 0.5 * (H + H.T), ref


@pytest.mark.parametrize("name", ["water", "caffeine"])
def test_numerical_hessian(name):
    from axtb_amd import _lib
    z, x = structure(name)
    import time
    ctx = _lib.Context(0)
    ctx.hessian(z, x, accuracy=1.0)  # warm-up (code objects, buffers)
    t0 = time.perf_counter()
    h = ctx.hessian(z, x, accuracy=1.0)
    t1 = time.perf_counter()
    Ho, ref = oracle_hessian(z, x)
    t2 = time.perf_counter()
    n3 = 3 * len(z)
    d = np.abs(h["hessian"] - Ho).max()
    print(f"[hessian {name}] {h['displacements']} displacements, max |dH| vs oracle {d:.2e}, "
          f"asymmetry {h['max_asymmetry']:.2e}; GPU {t1 - t0:.3f} s, oracle one thread per displacement in "
          f"sequence {t2 - t1:.2f} s", flush=True)
    assert h["displacements"] == 2 * n3
    assert d <= 1e-6
    assert abs(h["energy"] - ref["energy"]) <= 1e-8
    assert np.abs(h["gradient"] - ref["gradient"]).max() <= 1e-7
    assert np.array_equal(h["hessian"], h["hessian"].T)
    # translations: sum over the atoms B of H(A x, B y) vanishes (to the step's and SCF's accuracy)
    rows = h["hessian"].reshape(n3, len(z), 3).sum(axis=1)
    assert np.abs(rows).max() <= 1e-4 * max(1.0, np.abs(h["hessian"]).max())
    ctx.close()


def test_hessian_diagonal_against_energy_differences():
    """H(y, y) against (E(x + h e_y) - 2 E(x) + E(x - h e_y)) / h^2 of the engine's own energies."""
    from axtb_amd import _lib
    z, x = structure("water")
    ctx = _lib.Context(0)
    h = ctx.hessian(z, x, accuracy=1e-4, hessian_accuracy=1e-4)
    step = 0.005
    mols = []
    for y in range(9):
        for s in (1.0, -1.0):
            xp = np.array(x, dtype=float).reshape(-1)
            xp[y] += s * step
            mols.append((z, xp.reshape(-1, 3)))
    e = [r["energy"] for r in ctx.engrad_batch(mols, _lib.make_opts(accuracy=1e-6, gradient=False))]
    e0 = ctx.engrad(z, x, accuracy=1e-6)["energy"]
    for y in range(9):
        fd2 = (e[2 * y] - 2.0 * e0 + e[2 * y + 1]) / step ** 2
        assert abs(fd2 - h["hessian"][y, y]) <= 1e-4 * max(1.0, abs(fd2)), (y, fd2, h["hessian"][y, y])
    ctx.close()


def test_hessian_input_errors():
    """calculator validation (tests/validation.rs:136-148): a non-positive or non-finite step or accuracy."""
    from axtb_amd import _lib
    z, x = structure("water")
    ctx = _lib.Context(0)
    for bad in (0.0, -0.005, float("nan"), float("inf")):
        for kw in ({"step": bad}, {"hessian_accuracy": bad}):
            with pytest.raises(_lib.AxtbError) as e:
                ctx.hessian(z, x, **kw)
            assert e.value.code == _lib.ERR_INPUT
    ctx.close()
