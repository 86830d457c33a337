#!/usr/bin/env python3
"""Throughput of GFN2-xTB energy + analytic gradient on one MI355X.

Workload (default): the uneven 4096-molecule batch of the benchmark configuration: per molecule,
at random from the seed, a water cluster of 3-20 waters on a cubic lattice (2.9 A, random
orientations) or an all-trans n-alkane C3..C20, with 0.05 A Gaussian coordinate noise; axtb's
default conventions, SCC energy convergence 1e-8 Eh (accuracy 1e-2), 300 K, D4, EEQ start.
A step is one energy + gradient pass over the whole batch (setup, SCC with per-molecule
convergence masks, gradient), inputs resident on the device. Other workloads: --workload
water1000 / water200 / c100h202 / water20 / c30h62 (one molecule per step), conformers2048
(2048 C30H62 conformers, 0.1 A noise, one batch per step). --warm: a trajectory step of the
workload (every coordinate moved by 0.02 bohr, each molecule started from the converged moments
of its previous geometry).

Precision (default both): the workload is run in plain FP64 and in mixed precision (FP32 Jacobi
eigensolves and FP64 Ogita-Aishima refinement for the batched eigenproblems, n <= 200; larger
molecules on the FP64 dense solver in both modes), each prepared once and timed separately. `value` is the mixed
E+grad/s; `config` holds the FP64 E+grad/s, the speedup and the largest differences of the
mixed results from the FP64 results over all molecules, and both against the CPU oracle on the
first molecules.

Prints one JSON line: E+grad/s (value), the roofline of the dominant kernel measured with HIP
events on the engine's stream during the timed steps, and the CPU oracle's E+grad/s on a
bounded sample of the same workload as the reported baseline.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12   # MI355X dense FP64 matrix (flop/s)
PEAK_FP32_MFMA = 157.3e12  # MI355X dense FP32 matrix (flop/s); FP32 kernels are priced against it
PEAK_HBM = 8.0e12          # MI355X HBM3E (byte/s)
STREAMING = {"k_fock", "k_mulliken_cols", "k_integrals", "k_trd_symv"}


def workload(name, seed=0):
    from axtb_amd import structures as S
    if name == "batch4096":
        return S.batch(4096, seed=seed)
    if name.startswith("batch"):
        return S.batch(int(name[5:]), seed=seed)
    if name.startswith("conformers"):  # C30H62 conformers, 0.1 A noise (conformers2048)
        return S.conformers(int(name[10:]), seed=seed)
    return [S.config(name, seed=seed)]


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter summary
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), or None when it was not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get(kernel, {}).get(workload)
    return None if rec is None else rec["bytes_per_launch"]


def cpu_baseline(mols, budget_s, acc):
    """The CPU oracle (C++ restatement of axtb's FP64 path, one thread) on the first molecules of
    the workload until `budget_s` seconds of CPU work are spent."""
    import oracle
    t0 = time.perf_counter()
    n = 0
    for z, x in mols:
        oracle.engrad(z, x, acc=acc)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return n / dt, n, dt


def measure(mols, precision, steps, warmup, acc, starts=None):
    """Prepare the batch once, run `warmup` untimed and `steps` timed passes; returns the wall time of
    the timed passes, the HIP-event kernel profile of those passes and the results."""
    from axtb_amd import _lib
    opts = _lib.make_opts(accuracy=acc, etemp=300.0, precision=precision, gradient=True)
    ctx = _lib.Context(0)
    ctx.prepare(mols, opts, starts=starts)
    for i in range(warmup):
        ctx.run()
        print(f"[bench] {precision} warmup {i + 1}/{warmup} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_report()  # drop the warmup records
    t0 = time.perf_counter()
    for i in range(steps):
        ctx.run()
        print(f"[bench] {precision} step {i + 1}/{steps} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    prof = ctx.profile_report()
    ctx.profile(False)
    res = ctx.results()
    ctx.close()
    return dt, prof, res


def max_diffs(a, b):
    """max |dE|, |dg|, |dq|, |d eps| between two result sets (eps compared sorted)"""
    d = [0.0] * 4
    for r, o in zip(a, b):
        if r["status"] or o["status"]:
            continue
        d[0] = max(d[0], abs(r["energy"] - o["energy"]))
        d[1] = max(d[1], float(np.abs(r["gradient"] - o["gradient"]).max()))
        d[2] = max(d[2], float(np.abs(r["charges"] - o["charges"]).max()))
        d[3] = max(d[3], float(np.abs(np.sort(r["orbital_energies"]) - np.sort(o["orbital_energies"])).max()))
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch4096")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="both (default): FP64 and mixed on the same inputs, value = mixed, with the FP64 "
                         "rate and the mixed-vs-FP64 differences; or one of fp64 / mixed alone")
    ap.add_argument("--warm", action="store_true",
                    help="a trajectory step: every molecule moved by 0.02 bohr (seeded) and started from the "
                         "converged moments of its previous geometry (computed untimed)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=8, help="molecules checked against the oracle after timing")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (a kernel of known traffic for the HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": nbytes, "bytes_written": nbytes}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the batch shards by molecule with no exchange between GPUs; "
              "multi-GPU runs are not part of this build (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    acc = 1e-2
    mols = workload(args.workload)
    print(f"[bench] workload {args.workload}: {len(mols)} molecules, {sum(len(m[0]) for m in mols)} atoms",
          file=sys.stderr, flush=True)
    starts = None
    if args.warm:
        from axtb_amd import _lib
        c0 = _lib.Context(0)
        r0 = c0.engrad_batch(mols, _lib.make_opts(accuracy=acc), raise_on_error=True)
        c0.close()
        starts = [{k: r[k] for k in ("shell_charges", "dipoles", "quadrupoles")} for r in r0]
        rng = np.random.default_rng(1)
        mols = [(m[0], np.asarray(m[1]) + rng.normal(scale=0.02, size=np.shape(m[1]))) + tuple(m[2:]) for m in mols]
        print("[bench] trajectory step: previous geometry computed, starts taken", file=sys.stderr, flush=True)
    runs = {}
    for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        runs[prec] = measure(mols, prec, args.steps, args.warmup, acc, starts)
    main_prec = "mixed" if "mixed" in runs else "fp64"
    dt, prof, res = runs[main_prec]
    bad = [i for i, r in enumerate(res) if r["status"] != 0]
    iters = [r["iterations"] for r in res if r["status"] == 0]

    # parity of a few molecules against the CPU oracle (not timed)
    import oracle
    checked = {p: [0.0] * 4 for p in runs}
    for i in range(min(args.check, len(mols))):
        st = None if starts is None else {"qsh": starts[i]["shell_charges"], "dipoles": starts[i]["dipoles"],
                                          "quadrupoles": starts[i]["quadrupoles"]}
        o = oracle.engrad(mols[i][0], mols[i][1], acc=acc, start=st)
        ref = {"status": 0, "energy": o["energy"], "gradient": o["gradient"], "charges": o["qat"],
               "orbital_energies": o["emo"]}
        for p, (_, _, rs) in runs.items():
            checked[p] = [max(a, b) for a, b in zip(checked[p], max_diffs([rs[i]], [ref]))]

    # host-side counters (Jacobi sweeps per solve) are reported separately from the kernels
    counters = {k: v for k, v in prof.items() if not k.startswith("k_")}
    prof = {k: v for k, v in prof.items() if k.startswith("k_")}
    sweeps = {k: round(v[2] / v[0], 3) for k, v in counters.items() if k.startswith("sweeps_") and v[0]}
    refinement = {k: round(v[2] / v[0], 3) if k.startswith("refine_steps") else v[0] for k, v in counters.items()
                  if k.startswith("refine") and v[0]}
    # dominant kernel and its roofline from the live HIP-event timings
    total_gpu_ms = sum(v[1] for v in prof.values())
    name, (nl, ms, work) = max(prof.items(), key=lambda kv: kv[1][1])
    avg_s = ms / nl / 1e3
    per_launch = work / nl
    if any(name.startswith(s) for s in STREAMING):
        bound, peak, unit, achieved = "hbm", PEAK_HBM / 1e9, "GB/s", per_launch / avg_s / 1e9
    else:
        pk = PEAK_FP32_MFMA if name.endswith("_f32") else PEAK_FP64_MFMA
        bound, peak, unit, achieved = "mfma", pk / 1e12, "TFLOP/s", per_launch / avg_s / 1e12

    traffic = None if args.warm else measured_traffic(name, args.workload)  # counters exist for fresh starts only
    value = len(mols) * args.steps / dt
    names = ("dE", "dg", "dq", "deps")
    cfg_extra = {}
    for p in runs:
        cfg_extra[f"{p}_vs_oracle_max_abs"] = dict(zip(names, checked[p]))
    if len(runs) == 2:
        d = max_diffs(runs["mixed"][2], runs["fp64"][2])
        fp64_rate = len(mols) * args.steps / runs["fp64"][0]
        cfg_extra.update({"fp64_value": fp64_rate, "mixed_speedup_vs_fp64": value / fp64_rate,
                          "mixed_vs_fp64_max_abs_all_molecules": dict(zip(names, d)),
                          "mixed_vs_fp64_max_abs_efermi": max((abs(r["efermi"] - o["efermi"])
                                                               for r, o in zip(runs["mixed"][2], runs["fp64"][2])
                                                               if not r["status"] and not o["status"]), default=0.0),
                          "mixed_within_1e-6_Eh": d[0] <= 1e-6})
    line = {
        "metric": "E+grad/s (GFN2-xTB energy + analytic gradient)",
        "value": value,
        "unit": "E+grad/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64",
        "data": "synthetic (seeded molecules of the benchmark configuration)",
        "config": {"workload": args.workload + (" trajectory step (0.02 bohr, warm start)" if args.warm else ""),
                   "molecules": len(mols), "precision": main_prec,
                   "accuracy": acc, "etemp_K": 300.0, "mean_scf_iterations": float(np.mean(iters)) if iters else None,
                   "failed": len(bad), "parity_checked": min(args.check, len(mols)), **cfg_extra,
                   "mean_jacobi_sweeps": sweeps, "refinement": refinement},
        "roofline": {"bound": bound, "kernel": name, "achieved": achieved, "peak": peak, "unit": unit,
                     "frac": achieved / peak, "traffic": traffic, "avg_launch_ms": avg_s * 1e3, "launches": nl,
                     "kernel_share_of_gpu_time": ms / total_gpu_ms if total_gpu_ms else None},
        "kernels_ms": {k: round(v[1], 3) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])[:12]},
    }
    if not args.no_cpu:
        rate, n, cdt = cpu_baseline(mols, args.cpu_budget, acc)
        line["cpu_baseline"] = {"value": rate, "unit": "E+grad/s", "cores": 1, "kind": "port",
                                "sample": f"first {n} of {len(mols)} molecules of the workload, {cdt:.1f} s, "
                                          f"C++ FP64 oracle on 1 thread ({platform.processor() or platform.machine()})"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
