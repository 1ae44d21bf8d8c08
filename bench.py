"""Benchmark: 2D DIM helium shock-bubble cell-updates/s on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload helium_fp64|helium_fp32] [--gpus 1]

A "step" is one dim::advance_one_step (CFL dt, x sweep, ring BCs, y sweep,
ring BCs) over the whole grid, state resident in HBM.  Prints one JSON line.

value = cells x steps / wall seconds of the timed region (the reference's
metric, tools/quant/report_builders.hpp:298-300).  roofline: the sweep
kernel's algorithmic bytes per launch (each directional sweep reads and
writes the nv conserved fields once: 2 * nv * bytes/value per cell) over its
mean device duration from HIP events on the library's stream.  cpu_baseline:
the reference's own OpenMP solver (oracle/_ref/libdimref.so, built from the
reference's sources) timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (grid, precision, value bytes)
    "helium_fp64": ((2048, 1024), "fp64", 8),
    "helium_fp32": ((4096, 2048), "fp32_mixed", 4),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="helium_fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--fp32-steps", type=int, default=2000,
                    help="steps of the FP32-mixed 4096x2048 companion measurement (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=1000,
                    help="steps of the batched 64-instance companion measurement (0: skip)")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="N",
                    help="only run a device copy of N doubles (calibrates HBM byte counters) and exit")
    return ap.parse_args()


def traffic_for(workload):
    """HBM bytes per sweep launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_sweep_launch")


def cpu_baseline(case, nsteps):
    """The reference's OpenMP solver on this host's cores (bounded sample)."""
    from oracle.pyoracle import Checker, make_params, ref_available

    if not ref_available():
        return None
    threads = min(16, os.cpu_count() or 1)
    ref = Checker("ref")
    p = make_params(case.dim, case.materials)
    _, _, secs = ref.advance(p, case.N, case.dx, case.bc_lo, case.bc_hi, case.U, case.cfl, 1.0,
                             nsteps=nsteps, threads=threads)
    return {
        "value": case.ncells * nsteps / secs,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"{nsteps} FP64 steps of helium {case.N[0]}x{case.N[1]} from the initial state, "
                  f"reference dim::advance_one_step, OpenMP guided,1, {secs:.1f} s",
    }


def sweep_bytes(case, precision):
    """Algorithmic bytes of one sweep launch: every stored field read and written once."""
    if precision == "fp64":
        per_cell = 8 * case.nv
    else:  # FP64 partial masses + FP32 momentum, energy, volume fractions
        per_cell = 8 * case.nmat + 4 * (case.nv - case.nmat)
    return 2 * per_cell * case.ncells


def measure(dimhip, workloads, grid, precision, steps, warmup):
    case = workloads.helium_shock_bubble_2d(*grid)
    t_end = 1.0  # far beyond the steps taken: dt is CFL-limited throughout
    solver = dimhip.Solver.from_case(case, precision=precision)
    t, n = solver.advance(warmup, case.cfl, t_end, 0.0, use_graph=False)
    solver.synchronize()
    solver.set_timing(True)
    t0 = time.perf_counter()
    t, n = solver.advance(steps, case.cfl, t_end, t, use_graph=False)
    solver.synchronize()
    wall = time.perf_counter() - t0
    ms, counts = solver.kernel_times()
    solver.set_timing(False)
    solver.close()
    if n != steps:
        raise RuntimeError(f"took {n} steps, expected {steps}")
    return {"case": case, "value": case.ncells * steps / wall, "wall": wall, "ms": ms, "counts": counts}


def measure_batch(dimhip, workloads, count, steps):
    """Seeded uneven shock-interface instances, FP32-mixed, advanced together
    (autotuned tiles and schedule, hipGraph-captured steps)."""
    cases = workloads.batch_instances(count=count)
    cells = sum(c.ncells for c in cases)
    with dimhip.Batch(cases) as b:
        tune = b.autotune(0.8, trial_steps=3)
        b.advance(10, 0.8, use_graph=True)
        t0 = time.perf_counter()
        _, nsteps = b.advance(steps, 0.8, use_graph=True)
        wall = time.perf_counter() - t0
        b.set_timing(True)
        b.advance(20, 0.8, use_graph=False)
        ms, counts = b.kernel_times()
        b.set_timing(False)
    if min(nsteps) != steps or max(nsteps) != steps:
        raise RuntimeError("batched instances did not all take the requested steps")
    per_cell = 2 * (8 * 2 + 4 * 4)  # FP64 masses + FP32 momentum, energy, fraction, read and written
    sweep_bytes_avg = per_cell * cells
    achieved = sweep_bytes_avg / (ms[0] * 1e-3) / 1e9
    return {
        "workload": f"{count} seeded shock-interface instances, 256..1024 x 256..512, FP32-mixed, "
                    f"{steps} steps each, one launch per stage, hipGraph-captured steps",
        "value": cells * steps / wall,
        "unit": "cell-updates/s",
        "cells": cells,
        "ms_per_step": wall / steps * 1e3,
        "autotuned": tune["best"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "mean_ms": {"sweep": ms[0], "boundary": ms[1], "cfl": ms[2]},
                     "bytes_per_launch": sweep_bytes_avg},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from diffuse_vs_gfm_mms_amd import dimhip, workloads

    if args.copy_probe:
        import numpy as np
        src = np.arange(args.copy_probe, dtype=np.float64)
        out, _ = dimhip.kat(dimhip.KAT_COPY, 1, [("ideal_gas", 1.4, 0.0)], args.copy_probe, src)
        print(json.dumps({"copy_probe_doubles": args.copy_probe, "bytes_read": 8 * args.copy_probe,
                          "bytes_written": 8 * args.copy_probe, "ok": bool(np.array_equal(out, src))}))
        return

    grid, precision, _ = WORKLOADS[args.workload]
    main_run = measure(dimhip, workloads, grid, precision, args.steps, args.warmup)
    case = main_run["case"]
    cells = case.ncells
    value = main_run["value"]
    wall = main_run["wall"]
    ms = main_run["ms"]
    counts = main_run["counts"]
    sweep_ms = ms[0]
    bytes_per_launch = sweep_bytes(case, precision)
    achieved = bytes_per_launch / (sweep_ms * 1e-3) / 1e9
    out = {
        "metric": "cell-updates/s, 2D DIM MUSCL-Hancock+HLLC, FP32-mixed vs FP64; % HBM peak",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"helium shock-bubble {grid[0]}x{grid[1]}, {precision}, split x->y, CFL 0.8",
            "grid": list(grid),
            "precision": precision,
            "cells": cells,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_for(args.workload),
            "kernel": "sweep (x and y launches averaged)",
            "mean_ms": {"sweep": ms[0], "boundary": ms[1], "cfl": ms[2]},
            "launches": {"sweep": counts[0], "boundary": counts[1], "cfl": counts[2]},
            "bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": None,
    }
    if args.fp32_steps > 0 and precision == "fp64":
        g32 = WORKLOADS["helium_fp32"][0]
        r32 = measure(dimhip, workloads, g32, "fp32_mixed", args.fp32_steps, args.warmup)
        b32 = sweep_bytes(r32["case"], "fp32_mixed")
        a32 = b32 / (r32["ms"][0] * 1e-3) / 1e9
        out["fp32_mixed"] = {
            "workload": f"helium shock-bubble {g32[0]}x{g32[1]}, fp32_mixed, split x->y, CFL 0.8",
            "value": r32["value"],
            "unit": "cell-updates/s",
            "steps": args.fp32_steps,
            "ms_per_step": r32["wall"] / args.fp32_steps * 1e3,
            "speedup_vs_fp64_cell_rate": r32["value"] / value,
            "roofline": {"bound": "hbm", "achieved": a32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": a32 / HBM_PEAK_GBS, "traffic": traffic_for("helium_fp32"),
                         "mean_ms": {"sweep": r32["ms"][0], "boundary": r32["ms"][1], "cfl": r32["ms"][2]},
                         "bytes_per_launch": b32},
        }
    if args.batch_steps > 0 and precision == "fp64":
        out["batched"] = measure_batch(dimhip, workloads, args.batch_count, args.batch_steps)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(workloads.helium_shock_bubble_2d(2048, 1024), args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
