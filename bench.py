#!/usr/bin/env python3
"""Throughput of GFN2-xTB energy + analytic gradient on one MI355X: mixed precision against
plain FP64 on the 4096-molecule batch and on (H2O)1000.

Workloads (default: both):
- batch4096: 4096 independent molecules, per molecule at random from the seed a water cluster of
  3-20 waters on a cubic lattice (2.9 A, random orientations) or an all-trans n-alkane C3..C20,
  0.05 A Gaussian coordinate noise; one step = energy + gradient of the whole batch;
- water1000: (H2O)1000 on a 10x10x10 lattice (3000 atoms, 6000 basis functions); one step =
  one energy + gradient.
Settings: axtb's default conventions, SCC energy convergence 1e-8 Eh (accuracy 1e-2), 300 K, D4,
EEQ start. Inputs are resident on the device when the timed region starts (batch_prepare once,
batch_run per step). Each workload runs in plain FP64, in mixed precision and in FP64-refine (the
mixed path's algorithm, the previous orbitals refined, with every step in FP64, so that mixed /
FP64-refine is the gain of the precision alone) on the same inputs:
`value` is the mixed E+grad/s of the batch; `config` holds every rate, the mixed/FP64 ratios and
the differences of each precision from the FP64 CPU oracle (the batch: its first molecules,
computed after timing; (H2O)1000: the oracle's committed result, tests/golden/oracle_water1000.npz)
and of mixed from FP64 over all molecules. Fields that were not checked are null.

Timing: `steps` timed passes after `warmup` untimed ones, profiling off; then `--profile-steps`
passes with per-kernel HIP events on the engine stream (not part of `value`), from which the
dominant kernel's roofline is computed (its algorithmic work per launch / its average launch
time). The whole-path roofline follows the benchmark's work model: per SCF iteration
16/3 n^3 + 6 n^2 k flops and 184 n^2 bytes, per E+grad I times that plus 2 n^2 k flops and
104 n^2 bytes (n basis functions, k occupied orbitals, I the molecule's SCF iterations), priced at
the FP64 (plain) or FP32 (mixed) matrix peak and the HBM peak, divided by the measured time.

The CPU baseline is the C++ FP64 oracle on the host: the batch's first molecules, one molecule per
core (OpenMP over molecules), for about `--cpu-budget` seconds, plus a one-thread sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12   # MI355X dense FP64 matrix (flop/s)
PEAK_FP32_MFMA = 157.3e12  # MI355X dense FP32 matrix (flop/s)
PEAK_HBM = 8.0e12          # MI355X HBM3E (byte/s)
STREAMING = {"k_fock", "k_mulliken_cols", "k_integrals", "k_trd_symv"}
FP32_KERNELS = ("_f32", "_upd32")
VALENCE = {1: 1, 6: 4, 7: 5, 8: 6}  # GFN2 reference occupations of the workloads' elements
METRIC = "E+grad/s: (H2O)1000 & 4096-mol batch, mixed vs FP64 GPU at |ΔE|≤1e-6 Eh"


def workload(name, seed=0):
    from axtb_amd import structures as S
    if name == "batch4096":
        return S.batch(4096, seed=seed)
    if name.startswith("batch"):
        return S.batch(int(name[5:]), seed=seed)
    if name.startswith("conformers"):  # C30H62 conformers, 0.1 A noise (conformers2048)
        return S.conformers(int(name[10:]), seed=seed)
    return [S.config(name, seed=seed)]


def measured_traffic(kernel, wl):
    """HBM bytes per launch of `kernel` on workload `wl` from the committed counter summary
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), or None when it was not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get(kernel, {}).get(wl)
    return None if rec is None else rec["bytes_per_launch"]


def path_work(mols, results):
    """Algorithmic flops and HBM bytes of one E+grad pass over `mols` (the work model above)."""
    from axtb_amd import _lib
    wf = wb = 0.0
    for (z, _x, *rest), r in zip(mols, results):
        n = _lib.basis_size(z)[1]
        k = (sum(VALENCE[int(a)] for a in z) - (rest[0] if rest else 0.0)) / 2.0
        it = r["iterations"]
        wf += it * (16.0 / 3.0 * n ** 3 + 6.0 * n * n * k) + 2.0 * n * n * k
        wb += it * 184.0 * n * n + 104.0 * n * n
    return wf, wb


def measure(mols, precision, steps, warmup, profile_steps, acc, starts=None):
    """Prepare once, `warmup` untimed and `steps` timed passes (profiling off), then
    `profile_steps` passes with kernel events; returns (wall time of the timed passes, kernel
    profile, results, the ms covered by the union of each kernel's launch intervals)."""
    from axtb_amd import _lib
    opts = _lib.make_opts(accuracy=acc, etemp=300.0, precision=precision, gradient=True)
    ctx = _lib.Context(0)
    ctx.prepare(mols, opts, starts=starts)
    for i in range(warmup):
        ctx.run()
        print(f"[bench] {precision} warmup {i + 1}/{warmup} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ctx.run()
        print(f"[bench] {precision} step {i + 1}/{steps} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    prof, spans = {}, {}
    if profile_steps:
        ctx.profile(True)
        ctx.profile_report()
        for i in range(profile_steps):
            ctx.run()
        ctx.synchronize()
        prof = ctx.profile_report()
        spans = ctx.last_spans
        ctx.profile(False)
        print(f"[bench] {precision} profiled pass done", file=sys.stderr, flush=True)
    res = ctx.results()
    ctx.close()
    return dt, prof, res, spans


def max_diffs(a, b):
    """max |dE|, |dg|, |dq|, |d eps| over the molecule pairs that both succeeded (eps sorted);
    None where nothing was compared"""
    d, seen = [0.0] * 4, False
    for r, o in zip(a, b):
        if r["status"] or o["status"]:
            continue
        seen = True
        d[0] = max(d[0], abs(r["energy"] - o["energy"]))
        d[1] = max(d[1], float(np.abs(np.asarray(r["gradient"]) - np.asarray(o["gradient"])).max()))
        d[2] = max(d[2], float(np.abs(np.asarray(r["charges"]) - np.asarray(o["charges"])).max()))
        d[3] = max(d[3], float(np.abs(np.sort(r["orbital_energies"]) - np.sort(o["orbital_energies"])).max()))
    return dict(zip(("dE", "dg", "dq", "deps"), d if seen else [None] * 4))


def kernel_roofline(prof, wl, spans=None):
    """The dominant kernel (largest total time) of a profile and its roofline. `frac` divides the
    work of a launch by the launches' mean duration; where launches of the kernel overlap on two
    streams, `frac_span` divides the total work by the time their union covers instead."""
    prof = {k: v for k, v in prof.items() if k.startswith(("k_", "rocblas_"))}
    if not prof:
        return None
    total = sum(v[1] for v in prof.values())
    name, (nl, ms, work) = max(prof.items(), key=lambda kv: kv[1][1])
    avg_s = ms / nl / 1e3
    per_launch = work / nl
    if any(name.startswith(s) for s in STREAMING):
        bound, peak, unit, achieved = "hbm", PEAK_HBM / 1e9, "GB/s", per_launch / avg_s / 1e9
    else:
        pk = PEAK_FP32_MFMA if name.endswith(FP32_KERNELS) else PEAK_FP64_MFMA
        bound, peak, unit, achieved = "mfma", pk / 1e12, "TFLOP/s", per_launch / avg_s / 1e12
    span_ms = (spans or {}).get(name)
    frac_span = work / (span_ms / 1e3) / (peak * (1e9 if bound == "hbm" else 1e12)) if span_ms else None
    return {"bound": bound, "kernel": name, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "span_ms": span_ms, "frac_span": frac_span,
            "traffic": measured_traffic(name, wl), "avg_launch_ms": avg_s * 1e3,
            "launches": nl, "work_per_launch": per_launch, "kernel_share_of_gpu_time": ms / total if total else None,
            "kernels_ms": {k: round(v[1], 3) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])[:10]}}


def counters(prof):
    c = {k: v for k, v in prof.items() if not k.startswith(("k_", "rocblas_"))}
    return {"mean_jacobi_sweeps": {k: round(v[2] / v[0], 3) for k, v in c.items() if k.startswith("sweeps_") and v[0]},
            "refinement": {k: (round(v[2] / v[0], 3) if k.startswith("refine_steps") else v[0])
                           for k, v in c.items() if k.startswith("refine") and v[0]}}


def run_workload(wl, args, acc):
    mols = workload(wl)
    print(f"[bench] workload {wl}: {len(mols)} molecules, {sum(len(m[0]) for m in mols)} atoms",
          file=sys.stderr, flush=True)
    big = wl == "water1000"
    steps = min(args.steps, args.water1000_steps) if big else args.steps
    warmup = min(args.warmup, 1) if big else args.warmup
    starts = None
    if args.warm:
        from axtb_amd import _lib
        c0 = _lib.Context(0)
        r0 = c0.engrad_batch(mols, _lib.make_opts(accuracy=acc), raise_on_error=True)
        c0.close()
        starts = [{k: r[k] for k in ("shell_charges", "dipoles", "quadrupoles")} for r in r0]
        rng = np.random.default_rng(1)
        mols = [(m[0], np.asarray(m[1]) + rng.normal(scale=0.02, size=np.shape(m[1]))) + tuple(m[2:]) for m in mols]
    runs = {}
    precs = {"all": ["fp64", "mixed", "fp64-refine"], "both": ["fp64", "mixed"]}.get(args.precision, [args.precision])
    for prec in precs:
        runs[prec] = measure(mols, prec, steps, warmup, args.profile_steps, acc, starts)
    out = {"molecules": len(mols), "steps": steps, "warmup": warmup,
           "warm_start": bool(args.warm)}
    for prec, (dt, prof, res, spans) in runs.items():
        rate = len(mols) * steps / dt
        wf, wbytes = path_work(mols, res)
        peak = PEAK_FP32_MFMA if prec == "mixed" else PEAK_FP64_MFMA
        t_roof = wf / peak + wbytes / PEAK_HBM
        ok = [r for r in res if r["status"] == 0]
        out[prec] = {"value": rate, "ms_per_step": dt / steps * 1e3,
                     "mean_scf_iterations": float(np.mean([r["iterations"] for r in ok])) if ok else None,
                     "failed": len(res) - len(ok),
                     "whole_path_roofline": {"flop": wf, "bytes": wbytes, "roofline_ms": t_roof * 1e3,
                                             "frac": t_roof / (dt / steps), "peak_matrix_TFLOPs": peak / 1e12},
                     "dominant_kernel": kernel_roofline(prof, wl, spans), **counters(prof)}
    # parity against the FP64 CPU oracle (not timed)
    check = {}
    if big:
        f = np.load(os.path.join(ROOT, "tests", "golden", "oracle_water1000.npz"))
        same = np.array_equal(f["positions"], np.asarray(mols[0][1])) and not args.warm
        ref = [{"status": 0, "energy": float(f["energy"]), "gradient": f["gradient"], "charges": f["charges"],
                "orbital_energies": f["orbital_energies"]}]
        for p in runs:
            check[p] = max_diffs(runs[p][2], ref) if same else dict.fromkeys(("dE", "dg", "dq", "deps"))
        out["oracle"] = "committed oracle result (tests/golden/oracle_water1000.npz)" if same else None
    else:
        import oracle
        nchk = min(args.check, len(mols))
        refs = []
        for i in range(nchk):
            st = None if starts is None else {"qsh": starts[i]["shell_charges"], "dipoles": starts[i]["dipoles"],
                                              "quadrupoles": starts[i]["quadrupoles"]}
            o = oracle.engrad(mols[i][0], mols[i][1], acc=acc, start=st)
            refs.append({"status": 0, "energy": o["energy"], "gradient": o["gradient"], "charges": o["qat"],
                         "orbital_energies": o["emo"]})
        for p in runs:
            check[p] = max_diffs(runs[p][2][:nchk], refs) if nchk else dict.fromkeys(("dE", "dg", "dq", "deps"))
        out["oracle"] = f"C++ FP64 oracle on the first {nchk} molecules" if nchk else None
    for p in runs:
        out[p]["vs_oracle_max_abs"] = check[p]
    if "mixed" in runs and "fp64" in runs:
        d = max_diffs(runs["mixed"][2], runs["fp64"][2])
        out["mixed_speedup_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
        out["mixed_vs_fp64_max_abs_all_molecules"] = d
        out["mixed_within_1e-6_Eh_of_oracle"] = (check["mixed"]["dE"] is not None and check["mixed"]["dE"] <= 1e-6)
    if "mixed" in runs and "fp64-refine" in runs:
        # the same algorithm in both: the ratio of the precisions alone
        out["mixed_speedup_vs_fp64_refine"] = out["mixed"]["value"] / out["fp64-refine"]["value"]
        out["fp64_refine_vs_fp64_max_abs_all_molecules"] = (max_diffs(runs["fp64-refine"][2], runs["fp64"][2])
                                                            if "fp64" in runs else None)
    return out, mols


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(mols, budget_s, acc):
    """The CPU oracle (C++ restatement of axtb's FP64 path) on the first molecules of the batch: one
    molecule per core over all cores (at most 16, the host share of one GPU), about budget_s
    seconds, then one thread for budget_s / 3 seconds."""
    import oracle
    cores = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(cores)
    t0 = time.perf_counter()
    n = 0
    while n < len(mols) and time.perf_counter() - t0 < budget_s:
        chunk = mols[n:n + 4 * cores]
        oracle.engrad_many(chunk, acc=acc)
        n += len(chunk)
    dt = time.perf_counter() - t0
    oracle.set_threads(1)
    t1 = time.perf_counter()
    n1 = 0
    while n1 < len(mols) and time.perf_counter() - t1 < budget_s / 3:
        oracle.engrad_many(mols[n1:n1 + 4], acc=acc)
        n1 += len(mols[n1:n1 + 4])
    dt1 = time.perf_counter() - t1
    oracle.set_threads(cores)
    return {"value": n / dt, "unit": "E+grad/s", "cores": cores, "kind": "port",
            "sample": f"first {n} of {len(mols)} molecules of the batch in {dt:.1f} s, one molecule per core on "
                      f"{cores} threads (C++ FP64 oracle, OpenMP over molecules; {cpu_model()}, "
                      f"nproc {os.cpu_count()})",
            "single_thread_value": n1 / dt1, "single_thread_sample": f"first {n1} molecules in {dt1:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--water1000-steps", type=int, default=2,
                    help="timed steps of (H2O)1000 (at most --steps; its warmup is at most one pass)")
    ap.add_argument("--profile-steps", type=int, default=1, help="profiled passes after the timed ones")
    ap.add_argument("--workload", default="default",
                    help="default (batch4096 and water1000), or one of batch4096 / water1000 / water200 / "
                         "c100h202 / water20 / c30h62 / conformers2048 / batchN")
    ap.add_argument("--precision", default="all", choices=["all", "both", "fp64", "mixed", "fp64-refine"],
                    help="all: plain FP64, mixed and FP64-refine (the mixed algorithm with every step in FP64); "
                         "both: plain FP64 and mixed")
    ap.add_argument("--warm", action="store_true",
                    help="a trajectory step: every molecule moved by 0.02 bohr (seeded) and started from the "
                         "converged moments of its previous geometry (computed untimed)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=8, help="batch molecules checked against the oracle after timing")
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
    names = ["batch4096", "water1000"] if args.workload == "default" else [args.workload]
    results, mols0 = {}, None
    for wl in names:
        results[wl], mols = run_workload(wl, args, acc)
        if mols0 is None:
            mols0 = mols
    main_wl = names[0]
    main = results[main_wl]
    main_prec = next(p for p in ("mixed", "fp64", "fp64-refine") if p in main)
    dom = main[main_prec]["dominant_kernel"] or {}
    roof = {k: dom.get(k) for k in ("bound", "kernel", "achieved", "peak", "unit", "frac", "traffic",
                                     "avg_launch_ms", "launches", "work_per_launch", "kernel_share_of_gpu_time",
                                     "span_ms", "frac_span")}
    roof["whole_path"] = {wl: {p: r[p]["whole_path_roofline"]["frac"] for p in ("fp64", "mixed", "fp64-refine")
                               if p in r}
                          for wl, r in results.items()}
    line = {
        "metric": METRIC,
        "value": main[main_prec]["value"],
        "unit": "E+grad/s",
        "n_gpus": 1,
        "steps": main["steps"],
        "warmup": main["warmup"],
        "ms_per_step": main[main_prec]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_prec == "mixed" else "f64",
        "data": "synthetic (seeded molecules of the benchmark configuration)",
        "config": {"workload": main_wl + (" trajectory step (0.02 bohr, warm start)" if args.warm else ""),
                   "value_is": f"{main_wl} {main_prec}", "accuracy": acc, "etemp_K": 300.0, **results},
        "roofline": roof,
    }
    if not args.no_cpu and main_wl.startswith("batch"):
        line["cpu_baseline"] = cpu_baseline(mols0, args.cpu_budget, acc)
        if args.warm:
            line["cpu_baseline"]["sample"] += " (fresh EEQ starts: the GPU steps start from the previous moments)"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
