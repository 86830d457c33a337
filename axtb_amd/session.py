"""Many geometries of one molecule by lineage: the warm-start session of axtb's Python engine.

``Session.compute`` / ``compute_many`` follow ``Engine.compute`` / ``Engine.compute_many``
(crates/axtb-py/src/engine.rs:403-470, 836-843): every request names a lineage (a key) and
optional seeds; it starts from the moments its lineage's latest committed request left (earlier
in the same call, else the session's history), else from the first of its seeds that has such
moments, else from the initial guess; with ``commit`` its converged moments become its lineage's
history. A lineage's requests run in order, and a request of a lineage without history waits for
its seeds' earlier requests, so the results are those of running the requests one at a time.

On the GPU the requests whose start is settled run together: each round is one batch
(``axtbhip_gfn2_engrad_batch_guess``) of the next request of every lineage that can run, with
their starts, so many trajectories advance side by side. A warm start that does not converge is
computed once more from the initial guess (crates/base/src/method/retry.rs:150-158).
"""
from typing import Dict, List, Optional

import numpy as np

from . import _lib

MOMENTS = ("shell_charges", "dipoles", "quadrupoles")


class Session:
    """GFN2-xTB of one molecule (numbers, charge, uhf fixed) at many geometries, by lineage."""

    def __init__(self, numbers, charge=0.0, uhf=None, options=None, context=None):
        from .calculator import Options
        self.numbers = np.ascontiguousarray(numbers, dtype=np.int32)
        self.charge = float(charge)
        self.uhf = uhf
        self.options = options or Options()
        self._ctx = context
        self.history: Dict[str, dict] = {}  # lineage -> the moments of its latest commit

    def compute(self, positions, lineage="main", seeds=(), commit=True, accuracy=None, gradient=True):
        """One request (engine.rs `compute`): the response dict, or raises its error."""
        r = self.compute_many([dict(positions=positions, lineage=lineage, seeds=seeds, commit=commit,
                                    accuracy=accuracy, gradient=gradient)])[0]
        if isinstance(r, Exception):
            raise r
        return r

    def compute_many(self, requests: List[dict]):
        """Run `requests` (dicts with positions and optional lineage, seeds, commit, accuracy,
        gradient) as one at a time in their order would; returns, in order, each response dict
        (the engine's result with "from": the lineage whose moments it started from, or None) or
        its ``AxtbError``."""
        n = len(requests)
        lin = [str(q.get("lineage", "main")) for q in requests]
        seeds = [[str(s) for s in q.get("seeds", ())] for q in requests]
        order: Dict[str, List[int]] = {}
        for i, l in enumerate(lin):
            order.setdefault(l, []).append(i)
        done = [False] * n
        commits: List[Optional[dict]] = [None] * n
        out: List[object] = [None] * n

        def before(l, i):
            return [j for j in order.get(l, []) if j < i]

        def latest(l, i):
            for j in reversed(before(l, i)):
                if commits[j] is not None:
                    return commits[j]
            return self.history.get(l)

        def start(i):
            """(settled, lineage, moments): where request i starts, once its lineage's (and, if
            needed, its seeds') earlier requests have run (engine.rs:440-466)."""
            if not all(done[j] for j in before(lin[i], i)):
                return False, None, None
            m = latest(lin[i], i)
            if m is not None:
                return True, lin[i], m
            for s in seeds[i]:
                if not all(done[j] for j in before(s, i)):
                    return False, None, None
                m = latest(s, i)
                if m is not None:
                    return True, s, m
            return True, None, None

        while not all(done):
            ready = []
            for i in range(n):
                if not done[i]:
                    ok, src, m = start(i)
                    if ok:
                        ready.append((i, src, m))
            # ready is never empty: the first pending request of the lowest index has all its
            # earlier requests (own lineage and seeds alike) done
            groups: Dict[tuple, list] = {}
            for i, src, m in ready:
                q = requests[i]
                key = (q.get("accuracy") or self.options.accuracy, bool(q.get("gradient", True)))
                groups.setdefault(key, []).append((i, src, m))
            for (acc, grad), items in groups.items():
                res = self._run([requests[i]["positions"] for i, _, _ in items], [m for _, _, m in items], acc, grad)
                for (i, src, _), r in zip(items, res):
                    done[i] = True
                    if isinstance(r, Exception):
                        out[i] = r
                        continue
                    r["from"] = src
                    out[i] = r
                    if requests[i].get("commit", True):
                        commits[i] = {k: np.array(r[k]) for k in MOMENTS}
        for l, idx in order.items():
            for j in reversed(idx):
                if commits[j] is not None:
                    self.history[l] = commits[j]
                    break
        return out

    def _run(self, positions, starts, accuracy, gradient):
        """One batch of geometries with their starts on the GPU: results or errors, in order."""
        if self._ctx is None:
            self._ctx = _lib.Context(self.options.device)
        o = self.options
        opts = _lib.make_opts(accuracy=accuracy, etemp=o.etemp, max_iter=o.max_iter, damping=o.damping,
                              guess=o.guess, gradient=gradient, precision=o.precision, dispersion=o.dispersion,
                              solvation=o.solvation, solvent=o.solvent, state=o.solvation_state)
        mols = [(self.numbers, np.asarray(x, dtype=np.float64).reshape(-1, 3), self.charge, self.uhf)
                for x in positions]
        res = self._ctx.engrad_batch(mols, opts, starts=list(starts))
        err = self._ctx.last_error()
        return [r if r["status"] == 0 else _lib.AxtbError(r["status"], err) for r in res]
