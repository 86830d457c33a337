"""The entry point that allocates nothing (axtbhip_gfn2_engrad_noalloc) against axtbhip_gfn2_engrad:
two calls in a row, the second with a smaller molecule, each to the FP64 bar (|dE| <= 1e-8 Eh,
max |dg| <= 1e-7 Eh/bohr, the same iterations); the size limit is an input error."""
import numpy as np
import pytest

from conftest import structure

pytestmark = pytest.mark.gpu


def compare(ctx, z, x, **kw):
    from axtb_amd import _lib
    a = ctx.engrad(z, x, **kw)
    b = _lib.engrad_noalloc(ctx, z, x, **kw)
    assert b["status"] == 0, ctx.last_error()
    de = abs(a["energy"] - b["energy"])
    dg = float(np.abs(a["gradient"] - b["gradient"]).max())
    print(f"[noalloc] nat {len(z)} |dE| {de:.2e} |dg| {dg:.2e} it {b['iterations']}/{a['iterations']} "
          f"bitwise {a['energy'] == b['energy'] and np.array_equal(a['gradient'], b['gradient'])}", flush=True)
    assert de <= 1e-8 and dg <= 1e-7 and a["iterations"] == b["iterations"]
    assert np.abs(a["charges"] - b["charges"]).max() <= 1e-7


def test_noalloc_matches_engrad():
    from axtb_amd import _lib
    from axtb_amd import structures as S
    ctx = _lib.Context(0)
    z, x = S.alkane(40, noise=0.05)  # 242 basis functions, near the limit
    compare(ctx, z, x, accuracy=1e-2)
    zc, xc = structure("caffeine")  # smaller, after the larger one: the scratch holds the old contents
    compare(ctx, zc, xc, accuracy=1e-2)
    compare(ctx, zc, xc, accuracy=1e-2, precision="mixed")
    # too large: rejected before anything runs
    zb, xb = S.alkane(50)
    r = _lib.engrad_noalloc(ctx, zb, xb, accuracy=1e-2)
    assert r["status"] == _lib.ERR_INPUT and "256" in ctx.last_error()
    ctx.close()
