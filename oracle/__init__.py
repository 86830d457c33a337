"""CPU FP64 oracle of the GFN2-xTB path: TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg load this, as the
checker or the CPU baseline, never as the product. ``liboracle.so`` is built by ``make -C oracle``
(``__graft_entry__.build()`` does it).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

ERRORS = {1: "input", 2: "numerical", 3: "scf_not_converged", 4: "unsupported_element"}


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _lib = ctypes.CDLL(path)
        _lib.oracle_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().oracle_last_error().decode())


def sizes(numbers):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    nsh, nao = ctypes.c_int(), ctypes.c_int()
    _check(lib().oracle_sizes(len(z), _p(z), ctypes.byref(nsh), ctypes.byref(nao)))
    return nsh.value, nao.value


TERMS = ("repulsion", "dispersion", "electronic", "isotropic_es", "third_order", "anisotropic", "entropy", "band")


def engrad(numbers, positions, charge=0.0, acc=1.0, etemp=300.0, max_iter=250, guess="eeq", gradient=True,
           dispersion=True, start=None):
    """GFN2-xTB energy (+ gradient) of one molecule; positions in bohr [N,3]. ``start``: a given
    SCF start {"qsh", "dipoles", "quadrupoles"} (the moments of a previous result), else ``guess``."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    nat = len(z)
    nsh, nao = sizes(z)
    e = ctypes.c_double()
    it = ctypes.c_int()
    terms = np.zeros(8)
    g = np.zeros((nat, 3))
    qat = np.zeros(nat)
    qsh = np.zeros(nsh)
    dp = np.zeros((nat, 3))
    qp = np.zeros((nat, 6))
    emo = np.zeros(nao)
    gs = [None, None, None]
    if start is not None:
        gs = [np.ascontiguousarray(start[k], dtype=np.float64).reshape(-1) for k in ("qsh", "dipoles", "quadrupoles")]
        if gs[0].size != nsh or gs[1].size != 3 * nat or gs[2].size != 6 * nat:
            raise OracleError(1, "the start does not have the molecule's sizes")
    rc = lib().oracle_engrad_guess(nat, _p(z), _p(x), ctypes.c_double(charge), ctypes.c_double(acc),
                                   ctypes.c_double(etemp), max_iter, 0 if guess == "eeq" else 1, int(gradient),
                                   int(dispersion), _p(gs[0]), _p(gs[1]), _p(gs[2]), ctypes.byref(e), _p(terms),
                                   _p(g), _p(qat), _p(qsh), _p(dp), _p(qp), _p(emo), ctypes.byref(it))
    _check(rc)
    return {"energy": e.value, "terms": dict(zip(TERMS, terms)), "gradient": g if gradient else None,
            "qat": qat, "qsh": qsh, "dipoles": dp, "quadrupoles": qp, "emo": emo, "iterations": it.value}


def eeq(numbers, positions, charge=0.0):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions, dtype=np.float64)
    q = np.zeros(len(z))
    _check(lib().oracle_eeq(len(z), _p(z), _p(x), ctypes.c_double(charge), _p(q)))
    return q


def d4(numbers, positions, q, charge_model="eeq", radii="mctc", cutoffs=(40.0, 72.0, 40.0, 1.0, 1.0),
       gradient=False):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions, dtype=np.float64)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    n = len(z)
    cn, c6, alpha = np.zeros(n), np.zeros((n, n)), np.zeros(n)
    e2, e3 = ctypes.c_double(), ctypes.c_double()
    g = np.zeros((n, 3)) if gradient else None
    cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
    _check(lib().oracle_d4(n, _p(z), _p(x), _p(qq), 0 if charge_model == "eeq" else 1, 0 if radii == "mctc" else 1,
                           _p(cut), _p(cn), _p(c6), _p(alpha), ctypes.byref(e2), ctypes.byref(e3), _p(g)))
    return {"cn": cn, "c6": c6, "alpha": alpha, "e2": e2.value, "e3": e3.value, "gradient": g}


def overlap(numbers, positions, acc=1.0):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions, dtype=np.float64)
    nsh, nao = sizes(z)
    S = np.zeros((nao, nao))
    cn = np.zeros(len(z))
    _check(lib().oracle_overlap(len(z), _p(z), _p(x), ctypes.c_double(acc), _p(S), _p(cn)))
    return S.T.copy(), cn  # stored column-major


def eigh(a):
    a = np.array(a, dtype=np.float64)
    n = a.shape[0]
    buf = np.ascontiguousarray(a.T)
    w = np.zeros(n)
    _check(lib().oracle_eigh(n, _p(buf), _p(w)))
    return w, buf.T.copy()


def debug(numbers, positions, charge=0.0, acc=1.0):
    """Setup and first-iteration intermediates of the oracle (S, H0, D, Q, cn, qstart, potentials,
    F, eigenvalues, P, output shell charges)."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions, dtype=np.float64)
    nat = len(z)
    nsh, nao = sizes(z)
    nn = nao * nao
    o = {k: np.zeros(v) for k, v in dict(S=nn, H0=nn, D=3 * nn, Q=6 * nn, cn=nat, qstart=nsh, vsh=nsh, vat=nat,
                                          vdp=3 * nat, vqp=6 * nat, F=nn, w=nao, P=nn, qsh=nsh).items()}
    _check(lib().oracle_debug(nat, _p(z), _p(x), ctypes.c_double(charge), ctypes.c_double(acc),
                              *[_p(o[k]) for k in ("S", "H0", "D", "Q", "cn", "qstart", "vsh", "vat", "vdp", "vqp",
                                                   "F", "w", "P", "qsh")]))
    for k in ("S", "H0", "F", "P"):
        o[k] = o[k].reshape(nao, nao).T.copy()
    o["D"] = o["D"].reshape(3, nao, nao).transpose(0, 2, 1).copy()
    o["Q"] = o["Q"].reshape(6, nao, nao).transpose(0, 2, 1).copy()
    return o
