#!/usr/bin/env python3
"""Benchmark: Lennard-Jones pair interactions/s on one MI355X.

A step is one full pass of the hot path over the workload's particles already
resident in HBM: cell keys, scan, radix sort, gather, candidate count, the
27-cell force/energy kernel and the FP64 reductions (the reference's init +
compute phases, Cuda-version/cudaLJ.cu:613-653, plus forces).  value = accepted
ordered pairs / step time, over steps run without instrumentation; the phase split
(config.ms_bin, config.ms_force) comes from as many steps again with the engine's
phase events, whose time is reported as config.ms_per_step_with_phase_events.
The JSON line also carries:
  roofline      the force kernel's FP32-vector roofline fraction: 25 flops per
                accepted ordered pair over its hipEvent-timed duration
  accuracy      max / RMS relative and conditioned force error, energy error and
                non-finite counts of the last timed step against the FP64 oracle on a
                65,536-particle subsample (the oracle as the checker)
  cpu_baseline  the oracle's FP32 cell-grid restatement of the reference path on the
                whole workload with the host threads this job may use (plus a
                one-thread figure on a sample), with the CPU model
  batched       the 4,096-system batch: one call over all systems (eager and
                replayed from a captured graph) against one call per system
  sorted_mode   the same step through lj_compute_sorted, which leaves the particles in
                cell order (the reference's data flow, cudaLJ.cu:441-471) so every
                output store is consecutive: with the shuffled input copied back before
                each step inside the timed region, and in the steady state of a
                simulation whose particles stay in the order the last step left them
Before the timed steps the kernel shape (force-grid refinement, particles per
workgroup, LDS queue depth) is picked by the engine's autotuner on the same input,
outside the timed region; the chosen shape is reported under config.shape.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "particle-interaction-with-optix_amd"))

FLOPS_PER_PAIR = 25  # distance 8, 1/r^2 1, s^2 1, s^6 2, t 3, w 1, force 6, energy 3
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
METRIC = "LJ pair interactions/s (1 GPU) at force rel-err<=1e-5 vs FP64; batched/serial"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(torch, fn, steps):
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def time_with_phases(torch, eng, fn, steps):
    """The step time over `steps` steps without instrumentation, then the engine's
    per-phase times (hipEvents on the compute stream around binning and the force
    kernel) over as many more steps: the phase events themselves cost about 25 us per
    C3 step, so they stay out of the headline time.  Returns (ms, ms_instrumented,
    phases)."""
    eng.profile(False)
    ms = time_steps(torch, fn, steps)
    eng.profile(True)
    ms_ev = time_steps(torch, fn, steps)
    phases = eng.profile_read()
    eng.profile(False)
    return ms, ms_ev, phases


def sorted_mode(torch, ljamd, eng, dx, dy, dz, p, pairs, steps, warmup):
    """lj_compute_sorted on the workload of the headline step (same shape).  'restored':
    the shuffled positions are copied back before every step, inside the timed region,
    so each step sorts the same input as the headline; 'steady': the positions stay in
    the cell order the previous step left them."""
    sx, sy, sz = dx.clone(), dy.clone(), dz.clone()
    out = eng.compute_sorted(sx, sy, sz, p)

    def restored():
        sx.copy_(dx)
        sy.copy_(dy)
        sz.copy_(dz)
        eng.compute_sorted(sx, sy, sz, p, out=out)

    def steady():
        eng.compute_sorted(sx, sy, sz, p, out=out)

    res = {}
    for name, fn in (("restored", restored), ("steady", steady)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ms, _, ph = time_with_phases(torch, eng, fn, steps)
        r = ljamd.read_results(out["result"])[0]
        if int(r["n_pairs"]) != pairs:
            raise RuntimeError(f"sorted mode found {int(r['n_pairs'])} pairs, the headline step {pairs}")
        res[name] = {"ms_per_step": ms, "value": pairs / (ms / 1e3),
                     "ms_force": ph["force"][0] / max(ph["force"][1], 1),
                     "ms_bin": ph["bin"][0] / max(ph["bin"][1], 1)}
    del out, sx, sy, sz
    return res


def traffic_of(kernel_prefix: str, workload: str, p):
    """HBM bytes per dispatch of the force kernel in the shape being timed, from the
    committed PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), or None when that shape was not profiled."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        shape = f"{p.refine},{p.refine_x},{p.block},{p.depth}"
        v = d.get(workload, {}).get(kernel_prefix)
        return v.get(shape) if isinstance(v, dict) else None
    except Exception:
        return None


def _cpu_info():
    """Host CPU description: logical CPUs the machine shows, the CPUs this process may
    run on, and the model name."""
    model = None
    try:
        for ln in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity": affinity, "model": model}


def oracle_leg(ljamd, workload: str, w, out, max_threads: int):
    """The CPU side of the benchmark, where the oracle is the checker and the timed
    baseline:
      accuracy      the GPU forces of the last timed step against the FP64 oracle on a
                    fixed subsample (every 256th particle of C3: 65,536 rows)
      cpu_baseline  the oracle's FP32 cell-grid restatement of the reference path, timed
                    on the full workload with every host thread this job may use, plus a
                    one-thread run on a bounded sample."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po

    info = _cpu_info()
    threads = max(1, min(max_threads, info["affinity"] or 1))
    q = po.from_lj(w.params)
    n = w.x.size
    # ---- accuracy of the GPU result
    t0 = time.perf_counter()
    step = max(1, n // 65536)
    sub = np.arange(0, n, step, dtype=np.int64)
    ref = po.forces_f64(q, w.x, w.y, w.z, subset=sub, nthreads=threads)
    f = [out[k][:n].cpu().numpy()[sub].astype(np.float64) for k in ("fx", "fy", "fz")]
    e = out["e"][:n].cpu().numpy()[sub].astype(np.float64)
    g = [ref["fx"], ref["fy"], ref["fz"]]
    flt_max = float(np.finfo(np.float32).max)
    fin64 = (np.abs(g[0]) <= flt_max) & (np.abs(g[1]) <= flt_max) & (np.abs(g[2]) <= flt_max)
    fin32 = np.isfinite(f[0]) & np.isfinite(f[1]) & np.isfinite(f[2])
    ok = fin64 & fin32
    diff = np.sqrt(sum((a[ok] - b[ok]) ** 2 for a, b in zip(f, g)))
    mag = np.sqrt(sum(b[ok] ** 2 for b in g))
    rel = np.where(mag > 0, diff / np.where(mag > 0, mag, 1.0), diff)
    cond = np.where(ref["cond"][ok] > 0, diff / np.where(ref["cond"][ok] > 0, ref["cond"][ok], 1.0), diff)
    efin = np.isfinite(e) & (np.abs(ref["e"]) <= flt_max)
    erel = np.abs(e[efin] - ref["e"][efin]) / np.maximum(np.abs(ref["e"][efin]), 1e-300)
    accuracy = {
        "reference": "FP64 oracle (orc_forces_f64) on the same FP32 positions",
        "rows": int(sub.size),
        "sample": f"every {step}th particle",
        "max_rel_force_err": float(rel.max()) if rel.size else 0.0,
        "rms_rel_force_err": float(np.sqrt(np.mean(rel**2))) if rel.size else 0.0,
        "max_cond_force_err": float(cond.max()) if cond.size else 0.0,
        "rms_cond_force_err": float(np.sqrt(np.mean(cond**2))) if cond.size else 0.0,
        "max_rel_energy_err": float(erel.max()) if erel.size else 0.0,
        "nonfinite_fp64_forces": int((~fin64).sum()),
        "gpu_nonfinite_where_fp64_finite": int((fin64 & ~fin32).sum()),
        "within_1e-5": bool(rel.size == 0 or rel.max() <= 1e-5),
        "seconds": round(time.perf_counter() - t0, 1),
    }
    log(f"accuracy {accuracy}")
    # ---- CPU baseline: full workload on `threads` threads, then one thread on a sample
    po.forces_f32(q, w.x[:1000], w.y[:1000], w.z[:1000], nthreads=threads)  # warm the pool
    reps, t_total, pairs = 0, 0.0, 0
    while (t_total < 6.0 and reps < 3) or reps == 0:
        t0 = time.perf_counter()
        r = po.forces_f32(q, w.x, w.y, w.z, nthreads=threads)
        t_total += time.perf_counter() - t0
        pairs += r["n_pairs"]
        reps += 1
    full = pairs / t_total
    if workload == "c3":
        # one thread: 1/8 of the unit box at C3's density (2,097,152 particles in [0, 0.5)^3)
        ns, hi = 2_097_152, 0.5
        x1, y1, z1 = ljamd.generate("uniform", 1236, ns, [0.0, hi])
        q1 = po.params(0.01, box_hi=(hi, hi, hi))
        s1 = f"N={ns} uniform in [0,0.5)^3 (C3 density)"
    else:
        ns = min(n, 262_144)
        x1, y1, z1, q1 = w.x[:ns], w.y[:ns], w.z[:ns], q
        s1 = f"first {ns} particles of the workload"
    t0 = time.perf_counter()
    r1 = po.forces_f32(q1, x1, y1, z1, nthreads=1)
    one = r1["n_pairs"] / (time.perf_counter() - t0)
    return accuracy, {
        "value": full,
        "unit": "pair interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{w.desc}: the whole workload, bin + scan + sort + 27-cell FP32 force, {reps} reps, "
                  f"{t_total:.1f} s, OpenMP on {threads} threads",
        "cpu": info,
        "one_core": {"value": one, "sample": s1 + ", 1 thread"},
    }


def batched(torch, ljamd, steps, autotune=True):
    from ljamd import workloads

    t0 = time.perf_counter()
    systems, X, Y, Z = workloads.batch_c4()
    log(f"c4: {len(systems)} systems, {X.size} particles generated in {time.perf_counter() - t0:.1f} s")
    eng = ljamd.Engine(0)
    dx, dy, dz = eng.to_device(X, Y, Z)
    tuning = None
    if autotune:
        t0 = time.perf_counter()
        tuning = eng.autotune_batched(systems, dx, dy, dz, reps=2)
        log(f"c4 autotune {tuning} in {time.perf_counter() - t0:.1f} s")
    out = eng.compute_batched(systems, dx, dy, dz)
    torch.cuda.synchronize()
    t_batch = time_steps(torch, lambda: eng.compute_batched(systems, dx, dy, dz, out=out), steps)
    res = ljamd.read_results(out["result"], len(systems))
    pairs = int(res["n_pairs"].sum())
    # graph: the same call captured once and replayed
    t_graph = None
    try:
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            eng.compute_batched(systems, dx, dy, dz, out=out)
        torch.cuda.current_stream().wait_stream(s)
        with torch.cuda.graph(g):
            eng.compute_batched(systems, dx, dy, dz, out=out)
        g.replay()
        torch.cuda.synchronize()
        res_g = ljamd.read_results(out["result"], len(systems))
        same = bool((res_g["n_pairs"] == res["n_pairs"]).all() and (res_g["e_total"] == res["e_total"]).all())
        t_graph = time_steps(torch, g.replay, steps)
        if not same:
            log("graph replay disagrees with the eager result; graph time not reported")
            t_graph = None
    except Exception as e:  # graph capture is optional for the metric
        log(f"graph capture failed: {e}")
    # serial: one lj_compute call per system on views of the same arrays
    outs = []
    for sy in systems:
        a, b = sy.first, sy.first + sy.n
        outs.append((dx[a:b], dy[a:b], dz[a:b], sy.params, {
            "fx": out["fx"][a:b], "fy": out["fy"][a:b], "fz": out["fz"][a:b], "e": out["e"][a:b],
            "result": eng.empty_result()}))

    def serial():
        for (x, y, z, p, o) in outs:
            eng.compute(x, y, z, p, out=o)

    serial()
    t_serial = time_steps(torch, serial, max(1, steps // 5))
    serial_pairs = sum(int(ljamd.read_results(o["result"])[0]["n_pairs"]) for *_, o in outs)
    nsys = len(systems)
    best = min(t for t in (t_batch, t_graph) if t is not None)
    eng.close()
    return {
        "systems": nsys,
        "particles": int(X.size),
        "pairs": pairs,
        "pairs_match_serial": pairs == serial_pairs,
        "ms_batched_eager": t_batch,
        "ms_batched_graph": t_graph,
        "ms_serial": t_serial,
        "systems_per_s_batched": nsys / (best / 1e3),
        "systems_per_s_serial": nsys / (t_serial / 1e3),
        "speedup": t_serial / best,
        "shape": {"refine": systems[0].params.refine, "refine_x": systems[0].params.refine_x,
                  "block": systems[0].params.block, "depth": systems[0].params.depth},
        "autotune": tuning,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c5"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32", "fp32_comp"])
    ap.add_argument("--lbvh", action="store_true", help="time the LBVH path instead of the cell grid")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU leg (the oracle's accuracy check and the timed CPU baseline)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16") or 16),
                    help="host threads for the CPU leg (default: OMP_NUM_THREADS, else 16)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="use the default kernel shape")
    ap.add_argument("--shape", default=None, metavar="REFINE,REFINE_X,BLOCK,DEPTH",
                    help="fixed kernel shape for the single-system steps instead of autotuning")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="also run one 1 GiB device-to-device copy (reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import ljamd
    from ljamd import workloads

    t0 = time.perf_counter()
    w = workloads.single(args.workload, precision=args.precision)
    log(f"{w.desc}: generated in {time.perf_counter() - t0:.1f} s")
    eng = ljamd.Engine(0)
    dx, dy, dz = eng.to_device(w.x, w.y, w.z)
    n = w.x.size
    tuning = None
    if args.shape:
        w.params.refine, w.params.refine_x, w.params.block, w.params.depth = (int(v) for v in args.shape.split(","))
    elif not args.no_autotune and not args.lbvh:
        t0 = time.perf_counter()
        tuning = eng.autotune(dx, dy, dz, w.params, reps=3)
        log(f"autotune {tuning} in {time.perf_counter() - t0:.1f} s")
    out = eng.compute(dx, dy, dz, w.params, lbvh=args.lbvh)

    def step():
        eng.compute(dx, dy, dz, w.params, out=out, lbvh=args.lbvh)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ms, ms_events, phases = time_with_phases(torch, eng, step, args.steps)
    res = ljamd.read_results(out["result"])[0]
    pairs = int(res["n_pairs"])
    ms_force = phases["force"][0] / max(phases["force"][1], 1)
    ms_bin = phases["bin"][0] / max(phases["bin"][1], 1)
    achieved = FLOPS_PER_PAIR * pairs / (ms_force / 1e3) / 1e12
    kernel = "k_lbvh_force" if args.lbvh else "k_force"
    line = {
        "metric": METRIC,
        "value": pairs / (ms / 1e3),
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 predicate + f64 pair chain" if args.precision == "mixed" else "f32",
        "data": "synthetic",
        "config": {
            "workload": w.desc + (" (LBVH)" if args.lbvh else ""),
            "n": n,
            "pairs": pairs,
            "candidates": int(res["n_candidates"]),
            "ms_bin": ms_bin,
            "ms_force": ms_force,
            "ms_per_step_with_phase_events": ms_events,
            "pairs_per_s_force_kernel": pairs / (ms_force / 1e3),
            "e_total": float(res["e_total"]),
            "shape": {"refine": w.params.refine, "refine_x": w.params.refine_x, "block": w.params.block,
                      "depth": w.params.depth},
            "autotune": tuning,
        },
        "roofline": {
            "bound": "valu",
            "note": "vector-ALU bound, no matrix instructions: 25 flops per accepted ordered pair over the "
                    "kernel's hipEvent time, against the 157.3 TFLOP/s FP32 vector peak (which assumes packed "
                    "FP32; this build issues unpacked instructions only, so 50% of it is the ceiling). "
                    "frac_fp64_peak: the same flops against the 78.6 TFLOP/s FP64 vector peak, the precision "
                    "of the accepted-pair arithmetic in the mixed mode",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "frac_fp64_peak": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic_of(kernel, args.workload, w.params),
        },
    }
    if args.calibration_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        line["calibration_copy_bytes"] = 2 * src.numel() * 4
        del src, dst
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        line["accuracy"], line["cpu_baseline"] = oracle_leg(ljamd, args.workload, w, out, args.cpu_threads)
        log(f"cpu baseline {time.perf_counter() - t0:.1f} s")
    if not args.lbvh:
        line["sorted_mode"] = sorted_mode(torch, ljamd, eng, dx, dy, dz, w.params, pairs, args.steps, args.warmup)
    del out
    eng.close()
    if not args.no_batched:
        line["batched"] = batched(torch, ljamd, args.steps, autotune=not args.no_autotune)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
