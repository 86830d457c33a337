"""Autotuner of the engine's launch shapes per molecule size (the batched path, n <= 200 basis
functions): the Jacobi eigensolver's threads per matrix and whether its eigenvectors stay in LDS,
and which kernel takes the small per-molecule products.

Phase 1 proposes: for every size bucket (n / 8) a batch of molecules of that size (water clusters
and n-alkanes of the benchmark generator with 0.05 A noise), the kernels a knob controls timed
(HIP events per launch, the engine's profile) for each candidate value in each precision; a value
that beats the built-in rule by more than `--gain` (2 %) is a proposal. Phase 2 decides on the
target workload (the benchmark batch by default): the proposals are tried one at a time, in
order, and each is kept only if the workload's E+grad time (median of `--reps2` passes) drops by
more than `--gain2` (0.5 %): in a batch of uneven molecules, launches of several size classes
share the GPU (the eigensolver's HBM-resident launches run beside the LDS-resident ones), so a
shape that is faster alone need not be faster there. The result is a JSON cache,

    {"device": ..., "entries": [[knob, variant, bucket, value], ...], "measured": {...}}

which ``_lib.Context`` applies at creation (``axtbhip_tuning_set``; the committed
``tuning_gfx950.json`` by default). Run on the GPU it is for:

    python -m axtb_amd.autotune [--out axtb_amd/tuning_gfx950.json] [--molecules 256]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = os.path.join(HERE, "tuning_gfx950.json")
NB = 26


def bucket_molecules(b, count, seed=0):
    """Molecules whose basis size n satisfies n // 8 == b: (H2O)w (n = 6 w) and CkH2k+2
    (n = 6 k + 2), w, k in 3..33."""
    from . import structures as S
    rng = np.random.default_rng(seed + b)
    kinds = [("w", w) for w in range(3, 34) if (6 * w) // 8 == b] + [("k", k) for k in range(3, 34)
                                                                     if (6 * k + 2) // 8 == b]
    if not kinds:
        return []
    mols = []
    for i in range(count):
        kind, v = kinds[i % len(kinds)]
        if kind == "w":
            mols.append(S.water_cluster(v, math.ceil(v ** (1.0 / 3.0) - 1e-9), rng=rng, noise=0.05))
        else:
            mols.append(S.alkane(v, rng=rng, noise=0.05))
    return mols


def kernel_ms(ctx, mols, opts, prefixes, reps):
    """Best of `reps` profiled passes: total ms of the kernels whose names start with a prefix."""
    best = None
    for _ in range(reps):
        ctx.profile(True)
        ctx.engrad_batch(mols, opts)
        rep = ctx.profile_report()
        ctx.profile(False)
        t = sum(v[1] for k, v in rep.items() if k.startswith(prefixes))
        best = t if best is None else min(best, t)
    return best


def step_ms(ctx, mols, opts, reps):
    """Median wall time (ms) of `reps` E+grad passes over the batch resident on the device (as the
    benchmark times them: prepared once), one untimed pass first."""
    ctx.prepare(mols, opts)
    ctx.run()
    ts = []
    for _ in range(reps):
        ctx.synchronize()
        t = time.perf_counter()
        ctx.run()
        ctx.synchronize()
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))


def tune(out=DEFAULT, molecules=256, reps=2, gain=0.02, device=0, log=print, target=None, reps2=5, gain2=0.005):
    from . import _lib
    from . import structures as S
    ctx = _lib.Context(device, tuning=None)
    entries, measured = [], {}
    t0 = time.time()
    for prec in ("fp64", "mixed"):
        fp32 = 1 if prec == "mixed" else 0
        opts = _lib.make_opts(accuracy=1e-2, precision=prec)
        for b in range(NB):
            mols = bucket_molecules(b, molecules)
            if not mols:
                continue
            # the knobs of this bucket: Jacobi mode and threads (the Jacobi kernels), small products
            # (the GEMM kernels)
            trials = [("jacobi_mode", fp32, 1)]  # eigenvectors in HBM where both fit in LDS
            for mode in (1, 2):  # a value for the mode the launch does not use changes nothing
                for thr in (256, 512, 1024):
                    trials.append(("jacobi_threads", 3 * fp32 + mode, thr))
            if fp32 == 0:
                trials += [("gemm_small", 0, 1), ("gemm_small", 0, 2)]
            for knob in ("jacobi_mode", "jacobi_threads", "gemm_small"):
                cands = [t for t in trials if t[0] == knob]
                if not cands:
                    continue
                prefixes = ("k_gemm",) if knob == "gemm_small" else ("k_jacobi",)
                base = kernel_ms(ctx, mols, opts, prefixes, reps)
                best, best_t = None, base
                for k, variant, value in cands:
                    if ctx.tune(k, variant, b, value) != 0:
                        continue
                    t = kernel_ms(ctx, mols, opts, prefixes, reps)
                    ctx.tune(k, variant, b, 0)
                    measured[f"{prec}/{knob}/{variant}/{b}/{value}"] = round(t, 4)
                    if t < best_t * (1.0 - gain):
                        best, best_t = (k, variant, value), t
                measured[f"{prec}/{knob}/{b}/rule"] = round(base, 4)
                if best:
                    ctx.tune(best[0], best[1], b, best[2])  # keep it for the next knob of the bucket
                    entries.append([prec, best[0], best[1], b, best[2]])
                log(f"[autotune] {prec} bucket {b} (n {8 * b}-{8 * b + 7}) {knob}: rule {base:.3f} ms, "
                    f"best {best_t:.3f} ms {best or ''}")
        for e in entries:  # back to the rule before the next precision and phase 2
            ctx.tune(e[1], e[2], e[3], 0)
    # phase 2: the proposals on the target workload, one at a time
    mols = target if target is not None else S.batch(4096)
    kept, decisions = [], []
    for prec in ("fp64", "mixed"):
        opts = _lib.make_opts(accuracy=1e-2, precision=prec)
        cur = step_ms(ctx, mols, opts, reps2)
        log(f"[autotune] target {len(mols)} molecules {prec}: rule {cur:.2f} ms")
        start = cur
        for e in [e for e in entries if e[0] == prec] + [e for e in entries if e[0] != prec]:
            _, knob, variant, b, value = e
            prev = _current(kept, knob, variant, b)
            if prev == value:
                continue
            ctx.tune(knob, variant, b, value)
            t = step_ms(ctx, mols, opts, reps2)
            keep = t < cur * (1.0 - gain2)
            decisions.append([prec, knob, variant, b, value, round(t, 3), keep])
            log(f"[autotune] {prec} {knob} {variant} bucket {b} = {value}: {t:.2f} ms ({'kept' if keep else 'dropped'})")
            if keep:
                cur = t
                kept = [k for k in kept if k[:3] != [knob, variant, b]] + [[knob, variant, b, value]]
            else:
                ctx.tune(knob, variant, b, prev)
        measured[f"target/{prec}/rule"] = round(start, 3)
        measured[f"target/{prec}/tuned"] = round(cur, 3)
    res = {"device": _device_name(), "molecules_per_bucket": molecules, "gain": gain, "gain_target": gain2,
           "target_molecules": len(mols), "entries": kept, "proposals": entries, "decisions": decisions,
           "measured": measured, "seconds": round(time.time() - t0, 1)}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    ctx.close()
    return res


def _current(kept, knob, variant, b):
    for k in kept:
        if k[:3] == [knob, variant, b]:
            return k[3]
    return 0


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return None


def apply(ctx, path=DEFAULT):
    """Apply a tuning cache to a context; returns the number of entries applied."""
    with open(path) as f:
        entries = json.load(f)["entries"]
    n = 0
    for knob, variant, b, value in entries:
        if ctx.tune(knob, int(variant), int(b), int(value)) == 0:
            n += 1
    return n


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=DEFAULT)
    ap.add_argument("--molecules", type=int, default=256)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--gain", type=float, default=0.02)
    a = ap.parse_args()
    res = tune(a.out, a.molecules, a.reps, a.gain, log=lambda s: print(s, flush=True))
    print(json.dumps({"entries": res["entries"], "seconds": res["seconds"]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
