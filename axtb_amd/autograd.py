"""GFN2-xTB energies as a differentiable torch function.

``GFN2Energy.apply(positions, numbers, charges)`` returns the energies of a batch of molecules
(positions: list of [nat, 3] tensors in bohr, or one tensor [nmol, nat, 3] for conformers of one
molecule). The forward pass runs the engine's batched SCF and analytic gradient kernels; backward
returns grad_output * dE/dR from that gradient, so no 3 N^2 derivative data is stored.
"""
import numpy as np
import torch

from . import _lib

_ctx = {}


def _context(device):
    if device not in _ctx:
        _ctx[device] = _lib.Context(device)
    return _ctx[device]


class GFN2Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, numbers, charge=0.0, accuracy=1e-2, device=0):
        pos = positions.detach().to("cpu", torch.float64).numpy()
        nums = np.asarray(numbers, dtype=np.int32)
        mols = [(nums, pos[i], charge) for i in range(pos.shape[0])]
        res = _context(device).engrad_batch(mols, _lib.make_opts(accuracy=accuracy, gradient=True))
        bad = [r["status"] for r in res if r["status"]]
        if bad:
            raise _lib.AxtbError(bad[0], _context(device).last_error())
        e = torch.tensor([r["energy"] for r in res], dtype=torch.float64, device=positions.device)
        g = torch.tensor(np.stack([r["gradient"] for r in res]), dtype=torch.float64, device=positions.device)
        ctx.save_for_backward(g)
        return e

    @staticmethod
    def backward(ctx, grad_e):
        (g,) = ctx.saved_tensors
        return grad_e[:, None, None].to(g.dtype) * g, None, None, None, None


def energies(positions, numbers, charge=0.0, accuracy=1e-2, device=0):
    """Energies [nmol] of conformers positions [nmol, nat, 3] (bohr), differentiable in positions."""
    return GFN2Energy.apply(positions, numbers, charge, accuracy, device)
