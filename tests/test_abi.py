"""The C-ABI library loads and exports every symbol its headers declare (no GPU calls)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "axtb_amd", "libaxtb_hip.so")


def declared():
    names = []
    for h in ("axtb_hip.h", "axtb_hip_test.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(axtbhip_\w+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libaxtb_hip.so not built")
def test_exports():
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_lists_exports():
    from axtb_amd import _lib
    for n in _lib.EXPORTS:
        assert n in declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libaxtb_hip.so not built")
def test_version_and_basis_size():
    from axtb_amd import _lib
    import numpy as np
    L = _lib.lib()
    assert b"gfx950" in L.axtbhip_version()
    z = np.array([8, 1, 1], dtype=np.int32)
    x = np.zeros((3, 3))
    m = _lib.Mol(3, z.ctypes.data, x.ctypes.data, 0.0, -1)
    nsh, nao = ctypes.c_int(), ctypes.c_int()
    assert L.axtbhip_basis_size(ctypes.byref(m), ctypes.byref(nsh), ctypes.byref(nao)) == 0
    assert (nsh.value, nao.value) == (4, 6)
    # d shells: Fe (3d 4s 4p), S (3s 3p 3d), Au (5d 6s 6p): 3 shells and 9 functions each
    for zz in (26, 16, 79):
        z1 = np.array([zz], dtype=np.int32)
        m1 = _lib.Mol(1, z1.ctypes.data, x.ctypes.data, 0.0, -1)
        assert L.axtbhip_basis_size(ctypes.byref(m1), ctypes.byref(nsh), ctypes.byref(nao)) == 0
        assert (nsh.value, nao.value) == (3, 9), zz
    # Fr lies beyond the GFN2 parametrisation (UnsupportedElement), 119 and 0 are no elements (Input)
    for zz, rc in ((87, _lib.ERR_ELEMENT), (119, _lib.ERR_INPUT), (0, _lib.ERR_INPUT)):
        z1 = np.array([zz], dtype=np.int32)
        m1 = _lib.Mol(1, z1.ctypes.data, x.ctypes.data, 0.0, -1)
        assert L.axtbhip_basis_size(ctypes.byref(m1), ctypes.byref(nsh), ctypes.byref(nao)) == rc, zz


def test_workloads():
    """The synthetic workloads: sizes of the benchmark configurations and their geometry rules."""
    import numpy as np
    from axtb_amd import structures as S
    z, x = S.config("water1000")
    assert len(z) == 3000
    z, x = S.config("water20")
    d = np.linalg.norm(x[0] - x[1]) / S.AATOAU
    assert abs(d - 0.9572) < 1e-9
    z, x = S.alkane(100)
    assert len(z) == 302 and (z == 6).sum() == 100
    b = S.batch(64, seed=1)
    assert len(b) == 64 and all(9 <= len(m[0]) <= 62 for m in b)


def test_bench_single_gpu_only():
    """bench.py measures one GPU: --gpus N > 1 exits with a message, without touching a GPU."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_conformer_workload():
    """The differentiation workload: C30H62 conformers, all-trans plus 0.1 A Gaussian noise."""
    import numpy as np
    from axtb_amd import structures as S
    c = S.conformers(4, seed=3)
    assert len(c) == 4 and all(len(z) == 92 for z, _ in c)
    ref = S.alkane(30)[1]
    dev = np.concatenate([x - ref for _, x in c]).ravel() / S.AATOAU
    assert 0.07 < dev.std() < 0.13
