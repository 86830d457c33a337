"""ctypes wrappers of the kernel-level test hooks (include/axtb_hip_test.h), for parity tests."""
import ctypes

import numpy as np

from ._lib import Mol, lib, make_opts


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _mol(z, x, charge=0.0):
    z = np.ascontiguousarray(z, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    return Mol(len(z), z.ctypes.data, x.ctypes.data, float(charge), -1), (z, x)


def _rc(rc):
    if rc:
        raise RuntimeError(f"test hook failed with status {rc}")


def gemm(A, B, ta=False, tb=False, dk=None):
    """op(A) diag(dk) op(B) on the device (column-major arrays)."""
    Af = np.asfortranarray(A, dtype=np.float64)
    Bf = np.asfortranarray(B, dtype=np.float64)
    m = Af.shape[1] if ta else Af.shape[0]
    k = Af.shape[0] if ta else Af.shape[1]
    n = Bf.shape[0] if tb else Bf.shape[1]
    C = np.zeros((m, n), order="F")
    d = None if dk is None else np.ascontiguousarray(dk, dtype=np.float64)
    L = lib()
    L.axtbhip_test_gemm.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    _rc(L.axtbhip_test_gemm(m, n, k, int(ta), int(tb), _p(Af), Af.shape[0], _p(Bf), Bf.shape[0], _p(d), _p(C), m))
    return C


def gemm_upd32(U, B):
    """U + U (B - I) with the product in FP32 on the device."""
    Uf = np.asfortranarray(U, dtype=np.float64)
    Bf = np.asfortranarray(B, dtype=np.float64)
    m, n = Uf.shape
    C = np.zeros((m, n), order="F")
    L = lib()
    L.axtbhip_test_gemm_upd32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    _rc(L.axtbhip_test_gemm_upd32(m, n, _p(Uf), m, _p(Bf), n, _p(C), m))
    return C


def eigh(A):
    """Eigenpairs of symmetric matrices A [nb, n, n] (or one [n, n]) by the device Jacobi solver."""
    A = np.asarray(A, dtype=np.float64)
    one = A.ndim == 2
    if one:
        A = A[None]
    nb, n, _ = A.shape
    buf = np.ascontiguousarray(A.transpose(0, 2, 1))
    w = np.zeros((nb, n))
    Z = np.zeros((nb, n, n))
    sw = np.zeros(nb, dtype=np.int32)
    _rc(lib().axtbhip_test_eigh(n, nb, _p(buf), _p(w), _p(Z), _p(sw)))
    Z = Z.transpose(0, 2, 1)
    return (w[0], Z[0], sw[0]) if one else (w, Z, sw)


def dense_eigh(A, k=None):
    """All eigenvalues and the first k eigenvectors by the dense device solver."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    k = n if k is None else k
    buf = np.ascontiguousarray(A.T)
    w = np.zeros(n)
    Z = np.zeros((k, n))
    _rc(lib().axtbhip_test_dense_eigh(n, _p(buf), k, _p(w), _p(Z)))
    return w, Z.T.copy()


def setup(z, x, charge=0.0, **kw):
    mol, keep = _mol(z, x, charge)
    nsh, nao = ctypes.c_int(), ctypes.c_int()
    lib().axtbhip_basis_size(ctypes.byref(mol), ctypes.byref(nsh), ctypes.byref(nao))
    nat, nsh, nao = len(z), nsh.value, nao.value
    nn = nao * nao
    o = {k: np.zeros(v) for k, v in dict(S=nn, H0=nn, D=3 * nn, Q=6 * nn, cn=nat, qstart=nsh, X=nn, e3=nat).items()}
    opts = make_opts(**kw)
    _rc(lib().axtbhip_test_setup(ctypes.byref(mol), ctypes.byref(opts),
                                 *[_p(o[k]) for k in ("S", "H0", "D", "Q", "cn", "qstart", "X", "e3")]))
    for k in ("S", "H0", "X"):
        o[k] = o[k].reshape(nao, nao).T.copy()
    o["D"] = o["D"].reshape(3, nao, nao).transpose(0, 2, 1).copy()
    o["Q"] = o["Q"].reshape(6, nao, nao).transpose(0, 2, 1).copy()
    return o


def iter1(z, x, charge=0.0, **kw):
    mol, keep = _mol(z, x, charge)
    nsh, nao = ctypes.c_int(), ctypes.c_int()
    lib().axtbhip_basis_size(ctypes.byref(mol), ctypes.byref(nsh), ctypes.byref(nao))
    nat, nsh, nao = len(z), nsh.value, nao.value
    nn = nao * nao
    o = {k: np.zeros(v) for k, v in dict(F=nn, w=nao, P=nn, qsh=nsh, vsh=nsh, vat=nat, vdp=3 * nat,
                                          vqp=6 * nat).items()}
    opts = make_opts(**kw)
    _rc(lib().axtbhip_test_iter1(ctypes.byref(mol), ctypes.byref(opts),
                                 *[_p(o[k]) for k in ("F", "w", "P", "qsh", "vsh", "vat", "vdp", "vqp")]))
    for k in ("F", "P"):
        o[k] = o[k].reshape(nao, nao).T.copy()
    return o
