"""ctypes binding of libaxtb_hip.so (the C ABI declared in include/axtb_hip.h).

The library is built in-tree (``axtb_amd/libaxtb_hip.so``, see ``axtb_amd.build``). Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libaxtb_hip.so")

OK, ERR_INPUT, ERR_NUMERICAL, ERR_SCF, ERR_ELEMENT, ERR_DEVICE = range(6)
ERROR_KINDS = {1: "input", 2: "numerical", 3: "scf_not_converged", 4: "unsupported_element", 5: "device"}
TERMS = ("repulsion", "dispersion", "electronic", "isotropic_es", "third_order", "anisotropic", "entropy", "band",
         "solvation")
EXPORTS = ("axtbhip_version", "axtbhip_context_create", "axtbhip_context_destroy", "axtbhip_last_error",
           "axtbhip_basis_size", "axtbhip_gfn2_engrad", "axtbhip_gfn2_engrad_guess", "axtbhip_gfn2_engrad_batch",
           "axtbhip_gfn2_engrad_batch_guess", "axtbhip_batch_adjoint", "axtbhip_gfn2_engrad_noalloc",
           "axtbhip_gfn2_hessian",
           "axtbhip_batch_prepare", "axtbhip_batch_prepare_guess", "axtbhip_batch_update",
           "axtbhip_batch_run", "axtbhip_batch_gradient", "axtbhip_batch_results", "axtbhip_profile_enable",
           "axtbhip_profile_report", "axtbhip_synchronize")


class AxtbError(RuntimeError):
    """An error of the engine; ``kind`` is axtb's error kind (crates/base/src/error.rs:8-24)."""

    def __init__(self, code, msg):
        super().__init__(f"{ERROR_KINDS.get(code, code)}: {msg}")
        self.code = code
        self.kind = ERROR_KINDS.get(code, "unknown")


class Mol(ctypes.Structure):
    _fields_ = [("nat", ctypes.c_int), ("numbers", ctypes.c_void_p), ("positions", ctypes.c_void_p),
                ("charge", ctypes.c_double), ("uhf", ctypes.c_int)]


class Guess(ctypes.Structure):
    _fields_ = [("shell_charges", ctypes.c_void_p), ("dipoles", ctypes.c_void_p), ("quadrupoles", ctypes.c_void_p)]


class Opts(ctypes.Structure):
    _fields_ = [("accuracy", ctypes.c_double), ("etemp", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("damping", ctypes.c_double), ("guess", ctypes.c_int), ("want_gradient", ctypes.c_int),
                ("precision", ctypes.c_int), ("dispersion", ctypes.c_int), ("mixer", ctypes.c_int),
                ("solvation", ctypes.c_int), ("solvent", ctypes.c_char_p), ("solvation_state", ctypes.c_int),
                ("conventions", ctypes.c_int)]


# axtb_opts.mixer: the conventions' default (axtb: Broyden then DIIS; tblite: Broyden), or a plan
MIXERS = {"default": 0, "broyden": 1, "diis": 2, "auto": 3}
CONVENTIONS = {"axtb": 0, "xtb": 1, "xtb-6.7.1": 2, "tblite": 3}
MIXER_USED = {0: "broyden", 1: "diis", 2: "diis-fallback"}


class HessianOut(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("gradient", ctypes.c_void_p), ("hessian", ctypes.c_void_p),
                ("dipole_derivatives", ctypes.c_void_p), ("max_asymmetry", ctypes.c_double),
                ("displacements", ctypes.c_int), ("status", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("gradient", ctypes.c_void_p), ("charges", ctypes.c_void_p),
                ("shell_charges", ctypes.c_void_p), ("dipoles", ctypes.c_void_p), ("quadrupoles", ctypes.c_void_p),
                ("orbital_energies", ctypes.c_void_p), ("terms", ctypes.c_double * 9), ("efermi", ctypes.c_double),
                ("residual", ctypes.c_double), ("energy_change", ctypes.c_double), ("iterations", ctypes.c_int),
                ("status", ctypes.c_int), ("nsh", ctypes.c_int), ("nao", ctypes.c_int),
                ("dipole", ctypes.c_double * 3), ("gap", ctypes.c_double), ("mixer_used", ctypes.c_int),
                ("nattempts", ctypes.c_int), ("attempt_iterations", ctypes.c_int * 2),
                ("attempt_converged", ctypes.c_int * 2), ("element", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIBPATH)
        L.axtbhip_version.restype = ctypes.c_char_p
        L.axtbhip_context_create.restype = ctypes.c_void_p
        L.axtbhip_context_create.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.axtbhip_context_destroy.argtypes = [ctypes.c_void_p]
        L.axtbhip_last_error.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.axtbhip_basis_size.argtypes = [ctypes.POINTER(Mol), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.axtbhip_gfn2_engrad.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.POINTER(Opts),
                                          ctypes.POINTER(Result)]
        L.axtbhip_gfn2_engrad_guess.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.POINTER(Opts),
                                                ctypes.POINTER(Guess), ctypes.POINTER(Result)]
        L.axtbhip_gfn2_engrad_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.POINTER(Opts),
                                                  ctypes.POINTER(Result)]
        L.axtbhip_gfn2_hessian.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.POINTER(Opts),
                                           ctypes.POINTER(Guess), ctypes.c_double, ctypes.c_double,
                                           ctypes.POINTER(HessianOut)]
        L.axtbhip_batch_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.axtbhip_gfn2_engrad_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.c_int,
                                                ctypes.POINTER(Opts), ctypes.POINTER(Result)]
        L.axtbhip_gfn2_engrad_batch_guess.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.c_int,
                                                      ctypes.POINTER(Guess), ctypes.POINTER(Opts),
                                                      ctypes.POINTER(Result)]
        L.axtbhip_batch_prepare.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.c_int, ctypes.POINTER(Opts)]
        L.axtbhip_batch_prepare_guess.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.c_int,
                                                  ctypes.POINTER(Guess), ctypes.POINTER(Opts)]
        L.axtbhip_batch_update.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.c_int, ctypes.POINTER(Guess)]
        L.axtbhip_batch_run.argtypes = [ctypes.c_void_p]
        L.axtbhip_batch_gradient.argtypes = [ctypes.c_void_p]
        L.axtbhip_batch_results.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mol), ctypes.c_int,
                                            ctypes.POINTER(Result)]
        L.axtbhip_profile_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.axtbhip_profile_report.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.axtbhip_synchronize.argtypes = [ctypes.c_void_p]
        L.axtbhip_tuning_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


# axtb_opts.precision: FP64 with a full solve per SCC iteration, mixed FP32/FP64, and the mixed
# path's refinement with every step in FP64
PRECISIONS = {"fp64": 0, "mixed": 1, "fp64-refine": 2}


# implicit solvation (axtb_opts.solvation, .solvation_state)
SOLVATION = {None: 0, "alpb": 1, "gbsa": 2}
SOLVATION_STATES = {"gsolv": 0, "bar1M": 1, "bar1m": 1, "reference": 2}


def make_opts(accuracy=1.0, etemp=300.0, max_iter=250, damping=0.4, guess="eeq", gradient=True, precision="fp64",
              dispersion=True, mixer="default", solvation=None, solvent=None, state="gsolv", conventions="axtb"):
    """The options of a calculation; ``solvation`` None, "alpb" or "gbsa" with ``solvent`` (a name
    xtb accepts) and the reference ``state`` (gsolv, bar1M, reference); ``conventions`` "axtb" or
    "tblite"."""
    return Opts(float(accuracy), float(etemp), int(max_iter), float(damping), 0 if guess == "eeq" else 1,
                int(bool(gradient)), PRECISIONS[precision], int(bool(dispersion)), MIXERS[mixer],
                SOLVATION[solvation], (solvent or "").encode(), SOLVATION_STATES[state], CONVENTIONS[conventions])


class Context:
    """A device context (one HIP stream and its buffers) of the engine."""

    def __init__(self, device=0, tuning="default"):
        """tuning: "default" applies the committed autotuner cache (axtb_amd/tuning_gfx950.json) if
        present, a path applies that cache, None keeps the built-in launch rules."""
        err = ctypes.create_string_buffer(512)
        self._h = lib().axtbhip_context_create(int(device), err, 512)
        if not self._h:
            raise AxtbError(ERR_DEVICE, err.value.decode())
        if tuning is not None:
            from . import autotune
            path = autotune.DEFAULT if tuning == "default" else tuning
            if tuning != "default" or os.path.exists(path):
                autotune.apply(self, path)

    def tune(self, knob, variant, bucket, value):
        """One launch-shape entry of the tuning table (axtbhip_tuning_set); 0 or an error code."""
        return lib().axtbhip_tuning_set(self._h, knob.encode(), int(variant), int(bucket), int(value))

    def close(self):
        if self._h:
            lib().axtbhip_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self):
        buf = ctypes.create_string_buffer(1024)
        lib().axtbhip_last_error(self._h, buf, 1024)
        return buf.value.decode()

    def engrad_batch(self, molecules, opts, raise_on_error=False, starts=None):
        """molecules: list of (numbers, positions_bohr[, charge[, uhf]]). Returns a list of dicts.

        starts: None, or one given SCF start per molecule (None or a dict with "shell_charges",
        "dipoles", "quadrupoles", e.g. a previous result); a warm start that does not converge is
        computed once more from the fresh start."""
        mols, res, keep, outs = self._marshal(molecules)
        if starts is None:
            rc = lib().axtbhip_gfn2_engrad_batch(self._h, mols, len(molecules), ctypes.byref(opts), res)
        else:
            if len(starts) != len(molecules):
                raise AxtbError(ERR_INPUT, "one start per molecule")
            gs = self._starts(mols, starts, keep)
            rc = lib().axtbhip_gfn2_engrad_batch_guess(self._h, mols, len(molecules), gs, ctypes.byref(opts), res)
        return self._unmarshal(rc, res, outs, opts, raise_on_error)

    def _starts(self, mols, starts, keep):
        """The axtb_guess array of per-molecule starts (None: no start); checks the sizes."""
        if len(starts) != len(mols):
            raise AxtbError(ERR_INPUT, "one start per molecule")
        gs = (Guess * len(mols))()
        for i, st in enumerate(starts):
            if st is None:
                continue
            nat = mols[i].nat
            arrs = [np.ascontiguousarray(st[k], dtype=np.float64).reshape(-1)
                    for k in ("shell_charges", "dipoles", "quadrupoles")]
            nsh = ctypes.c_int()
            lib().axtbhip_basis_size(ctypes.byref(mols[i]), ctypes.byref(nsh), None)
            if arrs[0].size != nsh.value or arrs[1].size != 3 * nat or arrs[2].size != 6 * nat:
                raise AxtbError(ERR_INPUT, f"start {i} needs {nsh.value} shell charges and {nat} atomic "
                                           f"dipoles/quadrupoles")
            keep += arrs
            gs[i] = Guess(arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data)
        return gs

    def prepare(self, molecules, opts, starts=None):
        """Upload a batch once (with optional per-molecule starts); run() then computes on the
        resident inputs, results() downloads."""
        self._prep = self._marshal(molecules) + (opts,)
        if starts is None:
            rc = lib().axtbhip_batch_prepare(self._h, self._prep[0], len(molecules), ctypes.byref(opts))
        else:
            gs = self._starts(self._prep[0], starts, self._prep[2])
            rc = lib().axtbhip_batch_prepare_guess(self._h, self._prep[0], len(molecules), gs, ctypes.byref(opts))
        if rc:
            raise AxtbError(rc, self.last_error())

    def update(self, molecules, starts=None):
        """New positions (and starts) for the prepared batch's molecules (same atoms, same order);
        the next run() computes them. Raises on any mismatch and leaves the batch unchanged."""
        mols, res, keep, outs = self._marshal(molecules)
        gs = self._starts(mols, starts, keep) if starts is not None else None
        rc = lib().axtbhip_batch_update(self._h, mols, len(molecules), gs)
        if rc:
            raise AxtbError(rc, self.last_error())

    def run(self):
        self._grad_ready = False
        rc = lib().axtbhip_batch_run(self._h)
        if rc:
            raise AxtbError(rc, self.last_error())

    def adjoint(self, dE):
        """dE[i] * dE_i/dR for the prepared molecules, packed in input order ([sum nat_i, 3])."""
        mols = self._prep[0]
        dE = np.ascontiguousarray(dE, dtype=np.float64)
        if dE.shape != (len(mols),):
            raise AxtbError(ERR_INPUT, "one dE per prepared molecule")
        g = np.zeros((sum(m.nat for m in mols), 3))
        rc = lib().axtbhip_batch_adjoint(self._h, dE.ctypes.data, g.ctypes.data)
        if rc:
            raise AxtbError(rc, self.last_error())
        self._grad_ready = True
        return g

    def gradient(self):
        """The analytic gradient of the last run() (which may have run without gradient)."""
        rc = lib().axtbhip_batch_gradient(self._h)
        if rc:
            raise AxtbError(rc, self.last_error())
        self._grad_ready = True

    def results(self, raise_on_error=False):
        mols, res, keep, outs, opts = self._prep
        rc = lib().axtbhip_batch_results(self._h, mols, len(outs), res)
        return self._unmarshal(rc, res, outs, opts, raise_on_error,
                               with_gradient=opts.want_gradient or getattr(self, "_grad_ready", False))

    def profile(self, on=True):
        lib().axtbhip_profile_enable(self._h, int(on))

    def profile_report(self):
        """{kernel: (launches, total_ms, work)} of the launches since the last report; the ms
        covered by the union of each kernel's launch intervals go to `self.last_spans`."""
        buf = ctypes.create_string_buffer(1 << 16)
        lib().axtbhip_profile_report(self._h, buf, 1 << 16)
        out, self.last_spans = {}, {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms, work, span = line.rsplit(" ", 4)
            out[name] = (int(cnt), float(ms), float(work))
            self.last_spans[name] = float(span)
        return out

    def synchronize(self):
        lib().axtbhip_synchronize(self._h)

    def _marshal(self, molecules):
        n = len(molecules)
        mols = (Mol * n)()
        res = (Result * n)()
        keep, outs = [], []
        for i, m in enumerate(molecules):
            z = np.ascontiguousarray(m[0], dtype=np.int32)
            x = np.ascontiguousarray(m[1], dtype=np.float64).reshape(-1, 3)
            charge = float(m[2]) if len(m) > 2 else 0.0
            uhf = int(m[3]) if len(m) > 3 and m[3] is not None else -1
            nat = len(z)
            mols[i] = Mol(nat, z.ctypes.data, x.ctypes.data, charge, uhf)
            nsh, nao = ctypes.c_int(), ctypes.c_int()
            lib().axtbhip_basis_size(ctypes.byref(mols[i]), ctypes.byref(nsh), ctypes.byref(nao))
            o = {"gradient": np.zeros((nat, 3)), "charges": np.zeros(nat), "shell_charges": np.zeros(max(nsh.value, 1)),
                 "dipoles": np.zeros((nat, 3)), "quadrupoles": np.zeros((nat, 6)),
                 "orbital_energies": np.zeros(max(nao.value, 1))}
            r = res[i]
            for k, v in o.items():
                setattr(r, k, v.ctypes.data)
            keep += [z, x]
            outs.append(o)
        return mols, res, keep, outs

    def _unmarshal(self, rc, res, outs, opts, raise_on_error, with_gradient=None):
        n = len(outs)
        # a nonzero return with no molecule's own status is an error of the call itself
        if rc == ERR_DEVICE or (raise_on_error and rc) or (rc and not any(res[i].status for i in range(n))):
            raise AxtbError(rc, self.last_error())
        results = []
        for i in range(n):
            r = res[i]
            o = outs[i]
            kinds = ("broyden", "diis")
            first = 1 if (r.mixer_used == 1 and r.nattempts == 1) else 0
            d = {"status": r.status, "energy": r.energy, "iterations": r.iterations, "efermi": r.efermi,
                 "residual": r.residual, "energy_change": r.energy_change,
                 "terms": dict(zip(TERMS, list(r.terms))), "nsh": r.nsh, "nao": r.nao,
                 "dipole": np.array(list(r.dipole)), "gap": r.gap,
                 "mixer_used": MIXER_USED.get(r.mixer_used) if r.nattempts else None,
                 "attempts": [(kinds[first if a == 0 else 1], r.attempt_iterations[a], bool(r.attempt_converged[a]))
                              for a in range(r.nattempts)],
                 "element": r.element or None}
            if r.status == 0:
                d.update(o)
                d["shell_charges"] = o["shell_charges"][:r.nsh]
                d["orbital_energies"] = o["orbital_energies"][:r.nao]
                if not (opts.want_gradient if with_gradient is None else with_gradient):
                    d["gradient"] = None
            results.append(d)
        return results

    def hessian(self, numbers, positions, charge=0.0, uhf=None, start=None, step=0.005, hessian_accuracy=0.01, **kw):
        """Semi-numerical Hessian (axtbhip_gfn2_hessian): energy, gradient [N,3], hessian [3N,3N],
        dipole_derivatives [3N,3], max_asymmetry."""
        opts = make_opts(**kw)
        mols, res, keep, outs = self._marshal([(numbers, positions, charge, uhf)])
        gs = self._starts(mols, [start], keep) if start is not None else None
        nat = mols[0].nat
        g = np.zeros((nat, 3))
        H = np.zeros((3 * nat, 3 * nat))
        dd = np.zeros((3 * nat, 3))
        out = HessianOut(0.0, g.ctypes.data, H.ctypes.data, dd.ctypes.data, 0.0, 0, 0)
        rc = lib().axtbhip_gfn2_hessian(self._h, mols, ctypes.byref(opts), gs, float(step), float(hessian_accuracy),
                                        ctypes.byref(out))
        if rc:
            raise AxtbError(rc, self.last_error())
        return {"energy": out.energy, "gradient": g, "hessian": H, "dipole_derivatives": dd,
                "max_asymmetry": out.max_asymmetry, "displacements": out.displacements}

    def engrad(self, numbers, positions, charge=0.0, uhf=None, start=None, **kw):
        """One molecule (axtbhip_gfn2_engrad_guess); start: None or a given SCF start (a dict with
        "shell_charges", "dipoles", "quadrupoles")."""
        opts = make_opts(**kw)
        mols, res, keep, outs = self._marshal([(numbers, positions, charge, uhf)])
        gs = self._starts(mols, [start], keep) if start is not None else None
        rc = lib().axtbhip_gfn2_engrad_guess(self._h, mols, ctypes.byref(opts), gs, res)
        r = self._unmarshal(rc, res, outs, opts, False)[0]
        if r["status"]:
            raise AxtbError(r["status"], self.last_error())
        return r


def engrad_noalloc(ctx, numbers, positions, charge=0.0, uhf=None, **kw):
    """One molecule through axtbhip_gfn2_engrad_noalloc (no device allocation during the call)."""
    opts = make_opts(**kw)
    mols, res, keep, outs = ctx._marshal([(numbers, positions, charge, uhf)])
    rc = lib().axtbhip_gfn2_engrad_noalloc(ctx._h, mols, ctypes.byref(opts), res)
    return ctx._unmarshal(rc, res, outs, opts, False)[0]


def basis_size(numbers):
    """(nsh, nao) of a molecule's GFN2 basis (axtb::basis::Basis, crates/axtb/src/basis.rs:165-229)."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    m = Mol(len(z), z.ctypes.data, None, 0.0, -1)
    nsh, nao = ctypes.c_int(), ctypes.c_int()
    rc = lib().axtbhip_basis_size(ctypes.byref(m), ctypes.byref(nsh), ctypes.byref(nao))
    if rc:
        raise AxtbError(rc, "unsupported element")
    return nsh.value, nao.value
