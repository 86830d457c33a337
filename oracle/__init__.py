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


TERMS = ("repulsion", "dispersion", "electronic", "isotropic_es", "third_order", "anisotropic", "entropy", "band",
         "solvation")
SOLV_MODELS = {None: 0, "alpb": 1, "gbsa": 2}
SOLV_STATES = {"gsolv": 0, "bar1M": 1, "bar1m": 1, "reference": 2}
SOLV_CONVENTIONS = {"xtb": 0, "xtb671": 1, "corrected": 2}


MIXERS = {"broyden": 0, "diis": 1, "auto": 2, "default": -1}  # default: the conventions' plan
CONVENTIONS = {"axtb": 0, "tblite": 3}
MIXER_USED = {0: "broyden", 1: "diis", 2: "diis-fallback"}


def engrad(numbers, positions, charge=0.0, acc=1.0, etemp=300.0, max_iter=250, guess="eeq", gradient=True,
           dispersion=True, start=None, mixer="default", report=None, solvation=None, solvent=None, state="gsolv",
           conventions="axtb", uhf=None):
    """GFN2-xTB energy (+ gradient) of one molecule; positions in bohr [N,3]. ``start``: a given
    SCF start {"qsh", "dipoles", "quadrupoles"} (the moments of a previous result), else ``guess``.
    ``mixer``: the SCF mixer plan (auto = Broyden, then DIIS from the start: the axtb conventions).
    ``report``: a dict that receives the SCF report (iterations, mixer_used, attempts) also when
    the calculation fails. ``solvation``: None, "alpb" or "gbsa" with ``solvent`` (a name xtb
    accepts) and the reference ``state`` (gsolv, bar1M, reference). ``conventions``: "axtb" or
    "tblite" (conventions.rs:244-353); "default" mixer: the conventions' (auto for axtb, Broyden
    for tblite)."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    nat = len(z)
    nsh, nao = sizes(z)
    e = ctypes.c_double()
    terms = np.zeros(len(TERMS))
    g = np.zeros((nat, 3))
    qat = np.zeros(nat)
    qsh = np.zeros(nsh)
    dp = np.zeros((nat, 3))
    qp = np.zeros((nat, 6))
    emo = np.zeros(nao)
    info = np.zeros(7, dtype=np.int32)
    scal = np.zeros(5)
    gs = [None, None, None]
    if start is not None:
        gs = [np.ascontiguousarray(start[k], dtype=np.float64).reshape(-1) for k in ("qsh", "dipoles", "quadrupoles")]
        if gs[0].size != nsh or gs[1].size != 3 * nat or gs[2].size != 6 * nat:
            raise OracleError(1, "the start does not have the molecule's sizes")
    rc = lib().oracle_singlepoint(nat, _p(z), _p(x), ctypes.c_double(charge), ctypes.c_double(acc),
                                  ctypes.c_double(etemp), max_iter, 0 if guess == "eeq" else 1, int(gradient),
                                  int(dispersion), MIXERS[mixer], _p(gs[0]), _p(gs[1]), _p(gs[2]), ctypes.byref(e),
                                  _p(terms), _p(g), _p(qat), _p(qsh), _p(dp), _p(qp), _p(emo), _p(info), _p(scal),
                                  SOLV_MODELS[solvation], (solvent or "").encode(), SOLV_STATES[state],
                                  CONVENTIONS[conventions], -1 if uhf is None else int(uhf))
    rep = {"iterations": int(info[0]), "mixer_used": MIXER_USED.get(int(info[1])),
           "attempts": [(("broyden", "diis")[0 if (mixer != "diis" and a == 0) else 1], int(info[3 + 2 * a]),
                         bool(info[4 + 2 * a])) for a in range(int(info[2]))]}
    if report is not None:
        report.update(rep)
    _check(rc)
    return {"energy": e.value, "terms": dict(zip(TERMS, terms)), "gradient": g if gradient else None,
            "qat": qat, "qsh": qsh, "dipoles": dp, "quadrupoles": qp, "emo": emo, "efermi": scal[0],
            "dipole": scal[1:4].copy(), "gap": scal[4], **rep}


def solvation_parts(numbers, positions, charges, model, solvent, state="gsolv", conventions="corrected", nang=230,
                    gradient=False):
    """The ALPB/GBSA model alone at fixed charges (positions in bohr): {gsolv, gelec, gsasa, ghb,
    gshift} as xtb prints them, Born radii, surface areas, and the gradient of gsolv at fixed
    charges if asked."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(charges, dtype=np.float64)
    n = len(z)
    out, brad, sasa = np.zeros(5), np.zeros(n), np.zeros(n)
    g = np.zeros((n, 3)) if gradient else None
    _check(lib().oracle_solv_parts(n, _p(z), _p(x), _p(q), SOLV_MODELS[model], solvent.encode(), SOLV_STATES[state],
                                   SOLV_CONVENTIONS[conventions], nang, _p(out), _p(brad), _p(sasa), _p(g)))
    r = dict(zip(("gsolv", "gelec", "gsasa", "ghb", "gshift"), out))
    r.update(born_radii=brad, sasa=sasa, gradient=g)
    return r


def engrad_many(mols, acc=1.0, etemp=300.0, gradient=True):
    """Energies (and gradients) of many neutral molecules [(numbers, bohr)], one molecule per
    OpenMP thread (OMP_NUM_THREADS). Returns (energies, gradients or None, status)."""
    nats = np.array([len(m[0]) for m in mols], dtype=np.int32)
    off = np.zeros(len(mols), dtype=np.int64)
    off[1:] = np.cumsum(nats)[:-1]
    z = np.ascontiguousarray(np.concatenate([np.asarray(m[0], dtype=np.int32) for m in mols]))
    x = np.ascontiguousarray(np.concatenate([np.asarray(m[1], dtype=np.float64).reshape(-1, 3) for m in mols]))
    e = np.zeros(len(mols))
    g = np.zeros_like(x) if gradient else None
    st = np.zeros(len(mols), dtype=np.int32)
    lib().oracle_engrad_many(len(mols), _p(nats), _p(off), _p(z), _p(x), ctypes.c_double(acc), ctypes.c_double(etemp),
                             int(gradient), _p(e), _p(g), _p(st))
    return e, g, st


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


def diis_linear(eig, b, damp=0.4, max_steps=100):
    """The oracle's DIIS mixer on q = diag(eig) q + b: (steps, final input)."""
    eig = np.ascontiguousarray(eig, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(eig))
    steps = lib().oracle_diis_linear(len(eig), _p(eig), _p(b), ctypes.c_double(damp), max_steps, _p(x))
    return steps, x


def diis_coefficients(residuals):
    """Pulay coefficients of the residuals (rows, oldest first)."""
    r = np.ascontiguousarray(residuals, dtype=np.float64)
    c = np.zeros(r.shape[0])
    lib().oracle_diis_coefficients(r.shape[0], r.shape[1], _p(r), _p(c))
    return c


def set_threads(n):
    """OpenMP threads of the following oracle calls."""
    lib().oracle_set_threads(int(n))
