"""GFN2-xTB energies as a differentiable torch function.

``GFN2Energy.apply(positions, numbers, charge, accuracy, device, precision)`` returns the energies
[nmol] of conformers of one molecule (positions [nmol, nat, 3] in bohr). The forward pass runs
the engine's batched SCC without the gradient and leaves the converged state (P, the orbitals,
the potentials) resident on the GPU; backward runs the adjoint step on that state
(``axtbhip_batch_adjoint``: W and P contracted with dS/dR and dH0/dR recomputed per shell-pair
block, plus the classical and CN-chain terms, as ``XtbModel::gradient_functional``,
crates/axtb/src/xtb.rs:392-421) weighted by grad_output, grad_output[i] * dE_i/dR. Nothing of size 3 N^2 is stored.
If the context ran another batch between forward and backward, backward recomputes the state.
"""
import itertools

import numpy as np
import torch

from . import _lib

_ctx = {}
_serial = itertools.count(1)


def _context(device):
    if device not in _ctx:
        _ctx[device] = [_lib.Context(device), 0]  # context, serial of the batch it holds
    return _ctx[device]


def _check(res, c):
    bad = [r["status"] for r in res if r["status"]]
    if bad:
        raise _lib.AxtbError(bad[0], c.last_error())


class GFN2Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, numbers, charge=0.0, accuracy=1e-2, device=0, precision="fp64"):
        pos = positions.detach().to("cpu", torch.float64).numpy()
        nums = np.asarray(numbers, dtype=np.int32)
        mols = [(nums, pos[i], charge) for i in range(pos.shape[0])]
        slot = _context(device)
        c = slot[0]
        c.prepare(mols, _lib.make_opts(accuracy=accuracy, gradient=False, precision=precision))
        c.run()
        res = c.results()
        _check(res, c)
        slot[1] = next(_serial)
        ctx.serial, ctx.device, ctx.mols = slot[1], device, mols
        ctx.args = (accuracy, precision)
        return torch.tensor([r["energy"] for r in res], dtype=torch.float64, device=positions.device)

    @staticmethod
    def backward(ctx, grad_e):
        slot = _context(ctx.device)
        c = slot[0]
        if slot[1] != ctx.serial:  # the context holds another batch now: recompute this one
            accuracy, precision = ctx.args
            c.prepare(ctx.mols, _lib.make_opts(accuracy=accuracy, gradient=False, precision=precision))
            c.run()
            slot[1] = ctx.serial
        # the vector-Jacobian product on the device state (axtbhip_batch_adjoint): grad_e[i] dE_i/dR
        dE = grad_e.detach().to("cpu", torch.float64).numpy()
        g = c.adjoint(dE).reshape(len(ctx.mols), -1, 3)
        return torch.tensor(g, dtype=grad_e.dtype, device=grad_e.device), None, None, None, None, None


def energies(positions, numbers, charge=0.0, accuracy=1e-2, device=0, precision="fp64"):
    """Energies [nmol] of conformers positions [nmol, nat, 3] (bohr), differentiable in positions."""
    return GFN2Energy.apply(positions, numbers, charge, accuracy, device, precision)
