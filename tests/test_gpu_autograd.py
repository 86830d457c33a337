"""Forces of C30H62 conformers through torch.autograd (the deferred adjoint step on the GPU),
checked against the FP64 oracle's analytic gradient, against central finite differences of the
engine's own energies (h = 1e-4 bohr, the step of the differentiation workload), and between the
FP64 and mixed-precision paths. Bars: FP64 |dg| <= 1e-7 Eh/bohr vs the oracle, mixed <= 1e-5."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _conformers(n, seed):
    from axtb_amd import structures as S
    mols = S.conformers(n, seed=seed)
    return mols[0][0], np.stack([x for _, x in mols])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_conformer_forces(precision):
    import torch
    import oracle
    from axtb_amd import autograd, _lib
    z, X = _conformers(256, seed=11)
    pos = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    e = autograd.energies(pos, z, accuracy=1e-2, precision=precision)
    # powers of two: grad_output scales the gradient exactly
    w = torch.tensor([2.0 ** (i % 3 - 1) for i in range(len(e))], dtype=torch.float64)
    (w * e).sum().backward()
    print(f"{precision}: 256 x C30H62 energies + forces {time.perf_counter() - t0:.2f} s", flush=True)
    g = pos.grad.numpy() / w.numpy()[:, None, None]
    # the deferred adjoint equals a run that computes the gradient with the energy (same batch)
    ref = _lib.Context(0).engrad_batch([(z, X[i]) for i in range(len(X))],
                                       _lib.make_opts(accuracy=1e-2, precision=precision))
    for i in range(len(X)):
        assert ref[i]["energy"] == e[i].item()
        assert np.array_equal(ref[i]["gradient"], g[i])
    bar_e, bar_g = (1e-8, 1e-7) if precision == "fp64" else (1e-6, 1e-5)
    for i in (0, 97, 255):
        o = oracle.engrad(z, X[i], acc=1e-2)
        de, dg = abs(e[i].item() - o["energy"]), np.abs(g[i] - o["gradient"]).max()
        print(f"conformer {i}: |dE| {de:.2e} |dg| {dg:.2e}", flush=True)
        assert de <= bar_e and dg <= bar_g


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_conformer_finite_differences(precision):
    """32 conformers: the gradient along three random unit directions each, and along single
    coordinates of a C and an H atom, against central differences of the energy."""
    import torch
    from axtb_amd import autograd
    z, X = _conformers(32, seed=12)
    rng = np.random.default_rng(5)
    h = 1e-4
    dirs = []
    for i in range(32):
        for _ in range(3):
            d = rng.standard_normal(X[i].shape)
            dirs.append((i, d / np.linalg.norm(d)))
        for a, k in ((1, int(rng.integers(3))), (len(z) - 1, int(rng.integers(3)))):
            d = np.zeros_like(X[i])
            d[a, k] = 1.0
            dirs.append((i, d))
    pos = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    autograd.energies(pos, z, accuracy=1e-6, precision=precision).sum().backward()
    g = pos.grad.numpy()
    disp = np.stack([X[i] + s * h * d for i, d in dirs for s in (1.0, -1.0)])
    ed = autograd.energies(torch.tensor(disp), z, accuracy=1e-6, precision=precision).numpy()
    worst = 0.0
    for q, (i, d) in enumerate(dirs):
        fd = (ed[2 * q] - ed[2 * q + 1]) / (2 * h)
        an = float((g[i] * d).sum())
        worst = max(worst, abs(fd - an))
    print(f"{precision}: max |FD - analytic| over {len(dirs)} directions {worst:.2e} Eh/bohr", flush=True)
    assert worst <= 2e-7  # the reference's bar (crates/axtb/tests/consistency.rs:47-57)
