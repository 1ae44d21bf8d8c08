"""Benchmark: 2D DIM helium shock-bubble cell-updates/s on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload helium_fp64|helium_fp32] [--gpus 1]

A "step" is one dim::advance_one_step (CFL dt, x sweep, ring BCs, y sweep,
ring BCs) over the whole grid, state resident in HBM.  Prints one JSON line.

value = cells x steps / wall seconds of the timed region (the reference's
metric, tools/quant/report_builders.hpp:298-300).  roofline: the sweep
kernel's algorithmic bytes per launch (each directional sweep reads and
writes the nv conserved fields once: 2 * nv * bytes/value per cell) over its
mean device duration from HIP events on the library's stream, recorded
around the kernels of one step in TIMING_EVERY of the timed region (events on
every step cost 4-8 % of the throughput they would measure).  cpu_baseline:
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
# event timing samples one step in this many inside the timed region
TIMING_EVERY = 10

WORKLOADS = {
    # name: (grid, precision, value bytes)
    "helium_fp64": ((2048, 1024), "fp64", 8),
    "helium_fp32": ((4096, 2048), "fp32_mixed", 4),
}


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="helium_fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--fp32-steps", type=int, default=2000,
                    help="steps of the FP32-mixed 4096x2048 companion measurement (0: skip)")
    ap.add_argument("--fp64-c3-steps", type=int, default=200,
                    help="steps of the FP64 form on the FP32-mixed grid (4096x2048), for the speed-up on one grid "
                         "(0: skip)")
    ap.add_argument("--batch-steps", type=int, default=1000,
                    help="steps of the batched 64-instance companion measurement (0: skip)")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-tuning", default=os.path.join(ROOT, "diffuse_vs_gfm_mms_amd", "batch_tuning.json"),
                    help="JSON file of stored autotuner choices (read, and written when the instance set has none)")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="steps per run of the reference CPU solver on helium 2048x1024 (0: the sample's default, "
                         "2; 1 warm-up + 3 timed runs, about 15 s)")
    ap.add_argument("--cpu-all", action="store_true",
                    help="also time the reference CPU solver on the 4096x2048 grid, the 64 batched instances and "
                         "the 3D bubble collapse at 128^3 (about 2 minutes)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--accuracy-steps", type=int, default=-1,
                    help="steps of the FP32-mixed vs FP64 comparison at 4096x2048 (-1: as --fp32-steps, 0: skip "
                         "every FP32-mixed vs FP64 comparison)")
    ap.add_argument("--c5-steps", type=int, default=500,
                    help="steps of the 3D 512^3 FP32-mixed bubble-collapse measurement (0: skip)")
    ap.add_argument("--c5-n", type=int, default=512)
    ap.add_argument("--c5-accuracy-n", type=int, default=256,
                    help="grid of the 3D FP32-mixed vs FP64 comparison (0: skip)")
    ap.add_argument("--c5-accuracy-steps", type=int, default=100)
    ap.add_argument("--next-steps", type=int, default=100,
                    help="steps of each widened-row measurement (noble-Abel, unsplit, lambda model, "
                         "nmat 3; 0: skip)")
    ap.add_argument("--tube", action=argparse.BooleanOptionalAction, default=True,
                    help="the 1D water/air tube full run on the GPU and in the reference's CPU solver")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-probe", choices=["fp64", "fp32_mixed"], default=None,
                    help="profiling aid: 5 steps of the helium workload, then 30 single unfused x sweeps and "
                         "30 single unfused y sweeps (the fused y sweep runs inside the steps)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="N",
                    help="only run a device copy of N doubles (calibrates HBM byte counters) and exit")
    return ap.parse_args()


def traffic_for(workload, axis):
    """HBM bytes per launch of the sweep along `axis` from the committed
    rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(f"{axis}_sweep_bytes")


def valu_for(workload):
    """The SQ-counter instruction mix of the workload's sweeps from the
    committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"valu_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


# the reference's benchmark environment (tools/quant/runner.hpp:323-333)
CPU_OMP_ENV = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores", "OMP_DYNAMIC": "FALSE", "OMP_SCHEDULE": "guided,1"}

CPU_SAMPLES = {
    # name: (description, steps per run)
    "helium_2048x1024": ("FP64 helium shock-bubble 2048x1024", 2),
    "helium_4096x2048": ("FP64 helium shock-bubble 4096x2048 (the FP32-mixed grid)", 1),
    "batch_64": ("the 64 batched instances, FP64, one after another", 1),
    "bubble_collapse_128": ("FP64 3D bubble collapse 128^3", 1),
}


def cpu_cases(name):
    from diffuse_vs_gfm_mms_amd import workloads
    if name == "helium_2048x1024":
        return [workloads.helium_shock_bubble_2d(2048, 1024)]
    if name == "helium_4096x2048":
        return [workloads.helium_shock_bubble_2d(4096, 2048)]
    if name == "batch_64":
        return workloads.batch_instances(count=64)
    if name == "bubble_collapse_128":
        return [workloads.bubble_collapse_3d(128)]
    raise ValueError(name)


def physical_cores():
    """Physical cores among the CPUs this process may run on (lscpu's core
    ids of the affinity mask), the reference's "all physical cores"."""
    import subprocess
    cpus = sorted(os.sched_getaffinity(0))
    try:
        text = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        core_of = {}
        for line in text.splitlines():
            if line and not line.startswith("#"):
                cpu, core, sock = line.split(",")[:3]
                core_of[int(cpu)] = (sock, core)
        cores = {core_of[c] for c in cpus if c in core_of}
        if cores:
            return len(cores), len(cpus)
    except (OSError, ValueError):
        pass
    return len(cpus), len(cpus)


def cpu_worker(spec):
    """Child process: the reference's solver (oracle/_ref/libdimref.so, built
    from its sources) on the sample's cases, `warmup` untimed passes then
    `repeats` timed ones, each pass from the initial states; prints JSON."""
    from oracle.pyoracle import Checker, make_params
    cases = cpu_cases(spec["name"])
    ref = Checker("ref")

    def one_pass():
        cells = secs = 0.0
        for c in cases:
            p = make_params(c.dim, c.materials)
            _, _, t = ref.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, c.U, c.cfl, 1.0, nsteps=spec["steps"],
                                  threads=spec["threads"])
            cells += c.ncells * spec["steps"]
            secs += t
        return cells, secs

    for _ in range(spec["warmup"]):
        one_pass()
    runs = [one_pass() for _ in range(spec["repeats"])]
    print(json.dumps({"cell_updates": runs[0][0], "seconds": [r[1] for r in runs]}))


def cpu_baseline(name, steps=None, repeats=3, warmup=1):
    """The reference's OpenMP solver on this host: all physical cores of the
    affinity mask, the reference's OpenMP environment, 1 warm-up pass and the
    median of 3 (scripts/run_report2_local.sh:13-14), cells x steps / time of
    the time loop (report_builders.hpp:298-300), in a child process so the
    environment reaches the OpenMP runtime."""
    import statistics
    import subprocess
    from oracle.pyoracle import ref_available

    if not ref_available():
        return None
    desc, default_steps = CPU_SAMPLES[name]
    steps = steps or default_steps
    cores, cpus = physical_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), **CPU_OMP_ENV)
    spec = {"name": name, "steps": steps, "repeats": repeats, "warmup": warmup, "threads": cores}
    res = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", json.dumps(spec)],
                         capture_output=True, text=True, env=env, timeout=3000)
    if res.returncode != 0:
        raise RuntimeError("CPU baseline worker failed: " + res.stderr[-2000:])
    r = json.loads(res.stdout.strip().splitlines()[-1])
    rates = [r["cell_updates"] / t for t in r["seconds"]]
    return {
        "value": statistics.median(rates),
        "unit": "cell-updates/s",
        "cores": cores,
        "kind": "reference",
        "sample": f"{desc}: {steps} step(s) per run from the initial state, reference dim::advance_one_step, "
                  f"{warmup} warm-up + median of {repeats}, {cores} threads on {cores} physical cores "
                  f"({cpus} CPUs in the affinity mask), OMP_PROC_BIND=close OMP_PLACES=cores OMP_DYNAMIC=FALSE "
                  f"OMP_SCHEDULE=guided,1",
        "runs": rates,
        "seconds": r["seconds"],
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
    solver.set_timing(True, every=TIMING_EVERY)
    t0 = time.perf_counter()
    t, n = solver.advance(steps, case.cfl, t_end, t, use_graph=False)
    solver.synchronize()
    wall = time.perf_counter() - t0
    ms, counts = solver.kernel_times()
    dir_ms, dir_counts = solver.sweep_times()
    solver.set_timing(False)
    solver.close()
    if n != steps:
        raise RuntimeError(f"took {n} steps, expected {steps}")
    return {"case": case, "value": case.ncells * steps / wall, "wall": wall, "ms": ms, "counts": counts,
            "dir_ms": dir_ms, "dir_counts": dir_counts}


AXES = "xyz"


def sweep_roofline(case, precision, ms, dir_ms, traffic=None):
    """Roofline of the dominant sweep kernel: the slowest of the per-axis
    sweep launches (they move the same algorithmic bytes), with every axis's
    figure beside it."""
    b = sweep_bytes(case, precision)
    per_axis = {}
    for d in range(case.dim):
        a = b / (dir_ms[d] * 1e-3) / 1e9
        per_axis[AXES[d]] = {"mean_ms": dir_ms[d], "achieved": a, "frac": a / HBM_PEAK_GBS}
    dom = max(range(case.dim), key=lambda d: dir_ms[d])
    out = {"bound": "hbm", "achieved": per_axis[AXES[dom]]["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": per_axis[AXES[dom]]["frac"], "traffic": traffic,
           "kernel": f"{AXES[dom]} sweep (the slowest sweep launch; the last sweep of a step carries the fused CFL "
                     f"reduction)",
           "per_axis": per_axis, "mean_ms": {"sweep": ms[0], "boundary": ms[1], "cfl": ms[2]},
           "bytes_per_launch": b}
    return out


def field_l1(a, b):
    """Normalised L1, sum|a-b| / sum|b| (tools/quant/reference_error_metrics.hpp:123-137)."""
    import numpy as np
    return float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))


def mass_balance(case, c0, c1, flux_integral):
    """Relative mass change and the reference's boundary-flux-corrected mass
    balance (conservation_report.hpp:506-520, 596-607, 625-635) from the
    device's FP64 totals at the start (c0) and end (c1) and its time integral
    of the boundary flux; totals layout [mass_k (nmat), mass, mom_d, E]."""
    k = case.nmat
    m0, m1, flux = c0[k], c1[k], flux_integral[k]
    scale = max(abs(m0), abs(m1), 1e-12)
    residual = (m1 - m0) + flux
    return {"relative_mass_change": (m1 - m0) / m0,
            "balance": {"raw_drift": m1 - m0, "boundary_flux": flux, "residual": residual, "scale": scale,
                        "normalized_residual": abs(residual) / scale}}


def accuracy_vs_fp64(dimhip, case, steps, label):
    """The FP32-mixed form against the FP64 form (bit-identical to the
    reference's solver) at the same physical time: the FP64 run takes `steps`
    steps to time T, the FP32-mixed run advances from the same initial state
    to T (its own dt sequence, the last step clamped onto T, as at the
    reference's output times).  Normalised L1 of rho, p and alpha_0 from the
    device's FP64 CSV row values, and each form's relative change of total
    mass.  The flow's own sensitivity is reported beside it: the FP64 run
    repeated from the initial state with 1e-7 relative noise, compared at T
    with the same metric."""
    import numpy as np

    rows = {}
    out = {}
    with dimhip.Solver.from_case(case, precision="fp64") as s:
        U0 = s.download()
        s.track_boundary_flux(True)
        c0 = s.conservation()
        T, n64 = s.advance(steps, case.cfl, 1.0, 0.0, use_graph=True)
        rows["fp64"] = s.output_rows()
        out["fp64"] = dict(steps=n64, **mass_balance(case, c0, s.conservation(), s.boundary_flux_integral()))
    noisy = U0 * (1.0 + 1e-7 * np.random.default_rng(12345).standard_normal(U0.shape))
    for tag, prec, U in (("fp32_mixed", "fp32_mixed", U0), ("fp64_noise", "fp64", noisy)):
        with dimhip.Solver(case.dim, case.N, case.dx, case.materials, case.bc_lo, case.bc_hi,
                           precision=prec) as s:
            s.upload(U)
            s.track_boundary_flux(True)
            c0 = s.conservation()
            t, n = s.advance(4 * steps + 100, case.cfl, T, 0.0, use_graph=True)
            if abs(t - T) > 1e-12 * T:
                raise RuntimeError(f"{tag} run ended at t={t}, not at {T}")
            rows[tag] = s.output_rows()
            out[tag] = dict(steps=n, **mass_balance(case, c0, s.conservation(), s.boundary_flux_integral()))
        del U
    D = case.dim
    b = rows["fp64"]

    def l1s(a):
        return {"rho": field_l1(a[:, 0], b[:, 0]), "p": field_l1(a[:, 1 + D], b[:, 1 + D]),
                "alpha0": field_l1(a[:, 3 + D], b[:, 3 + D])}

    e32, e_noise = l1s(rows["fp32_mixed"]), l1s(rows["fp64_noise"])
    return {
        "workload": f"{label}, to the time of {steps} FP64 steps, both forms from the same state",
        "time": T,
        "steps_fp64": out["fp64"]["steps"],
        "steps_fp32_mixed": out["fp32_mixed"]["steps"],
        "l1_rho": e32["rho"],
        "l1_p": e32["p"],
        "l1_alpha0": e32["alpha0"],
        "fp64_noise_1e-7_l1": e_noise,
        "relative_mass_change_fp64": out["fp64"]["relative_mass_change"],
        "relative_mass_change_fp32_mixed": out["fp32_mixed"]["relative_mass_change"],
        "mass_change_difference": out["fp32_mixed"]["relative_mass_change"] - out["fp64"]["relative_mass_change"],
        "mass_balance": {
            "definition": "(M(T) - M(0)) + integral of the boundary mass flux of each step's new state, dt-weighted "
                          "(accumulate_fluxes, conservation_report.hpp:506-520), normalised by max(|M(0)|, "
                          "max |M|, 1e-12) (:596-607, 625-635)",
            "fp64": out["fp64"]["balance"],
            "fp32_mixed": out["fp32_mixed"]["balance"],
            "fp64_noise_1e-7": out["fp64_noise"]["balance"],
            "normalized_residual_difference": out["fp32_mixed"]["balance"]["residual"] / out["fp64"]["balance"]["scale"]
            - out["fp64"]["balance"]["residual"] / out["fp64"]["balance"]["scale"],
        },
    }


def accuracy_fp32(dimhip, workloads, grid, steps):
    r = accuracy_vs_fp64(dimhip, workloads.helium_shock_bubble_2d(*grid), steps,
                         f"helium shock-bubble {grid[0]}x{grid[1]}")
    r["note"] = ("open boundaries: mass changes by the boundary flux; the forms' difference is the FP32 form's "
                 "extra drift (closed-box drift: tests/test_gpu_fp32_mixed.py::test_closed_box_mass_drift)")
    return r


def accuracy_batch(dimhip, workloads, count, steps, instances):
    """The first batched instances, each run on its own in both forms (a
    batched instance equals its single-grid FP32-mixed run bit for bit:
    tests/test_gpu_batch.py::test_batch_equals_single_instances)."""
    cases = workloads.batch_instances(count=count)[:instances]
    per = []
    for i, c in enumerate(cases):
        r = accuracy_vs_fp64(dimhip, c, steps, f"instance {i} ({c.N[0]}x{c.N[1]})")
        per.append({"instance": i, "grid": list(c.N), "time": r["time"], "l1_rho": r["l1_rho"], "l1_p": r["l1_p"],
                    "l1_alpha0": r["l1_alpha0"], "fp64_noise_1e-7_l1": r["fp64_noise_1e-7_l1"],
                    "mass_change_difference": r["mass_change_difference"]})
    return {
        "workload": f"instances 0..{instances - 1} of the batch, to the time of {steps} FP64 steps",
        "instances": per,
    }


def measure_3d(dimhip, workloads, n, steps, warmup):
    """3D shock-induced air-bubble collapse in water, n^3 cells, FP32-mixed,
    initial state generated on the device."""
    case = workloads.bubble_collapse_3d(n, build_state=False)
    with dimhip.Solver.from_case(case, precision="fp32_mixed") as s:
        t, _ = s.advance(warmup, case.cfl, 1.0, 0.0, use_graph=False)
        s.synchronize()
        s.set_timing(True, every=TIMING_EVERY)
        t0 = time.perf_counter()
        t, nt = s.advance(steps, case.cfl, 1.0, t, use_graph=False)
        s.synchronize()
        wall = time.perf_counter() - t0
        ms, counts = s.kernel_times()
        dir_ms, _ = s.sweep_times()
    if nt != steps:
        raise RuntimeError(f"took {nt} steps, expected {steps}")
    return {
        "workload": f"3D bubble collapse {n}^3 (water/air, 1e9 Pa shock), fp32_mixed, split x->y->z, CFL 0.8, "
                    f"{steps} steps, device-generated initial state",
        "value": case.ncells * steps / wall,
        "unit": "cell-updates/s",
        "ms_per_step": wall / steps * 1e3,
        "roofline": sweep_roofline(case, "fp32_mixed", ms, dir_ms),
    }


def measure_next_rows(dimhip, workloads, steps):
    """The path's widened rows, FP64, each timed over `steps` steps after 5
    warm-up steps (no events): the noble-Abel water/air bubble (Gorsse TC9),
    the unsplit update, the equal-velocity lambda model, a third material
    slot, and the device-side output rows and initial state."""
    def run(case, label, **kw):
        with dimhip.Solver.from_case(case, precision="fp64", **kw) as s:
            t, _ = s.advance(5, case.cfl, 1.0, 0.0, use_graph=False)
            s.synchronize()
            t0 = time.perf_counter()
            t, n = s.advance(steps, case.cfl, 1.0, t, use_graph=False)
            s.synchronize()
            wall = time.perf_counter() - t0
        if n != steps:
            raise RuntimeError(f"{label}: took {n} steps, expected {steps}")
        return {"workload": label, "steps": n, "ms_per_step": wall / n * 1e3,
                "value": case.ncells * n / wall, "unit": "cell-updates/s"}

    out = {}
    c = workloads.gorsse_tc9_2d(960, 800, build_state=False)
    out["noble_abel_tc9_2d"] = run(c, "Gorsse TC9 water (stiffened) / air (noble-Abel) bubble 960x800, "
                                      "device initial state, CFL 0.45")
    c = workloads.helium_shock_bubble_2d(2048, 1024)
    out["unsplit_2d"] = run(c, "helium shock-bubble 2048x1024, unsplit update", time_update="unsplit")
    c.lambda_model = "equal_velocity"
    out["equal_velocity_2d"] = run(c, "helium shock-bubble 2048x1024, equal-velocity lambda model")
    c3 = workloads.helium_shock_bubble_2d(2048, 1024)
    c3.materials = c3.materials + [workloads.WATER]
    c3.U = None  # three-material state from the device initialiser (water slot empty)
    out["three_material_2d"] = run(c3, "helium shock-bubble 2048x1024 with a third (water) material "
                                       "slot, empty, nmat = 3 kernels")
    with dimhip.Solver.from_case(c, precision="fp64") as s:
        s.advance(5, c.cfl, 1.0, 0.0, use_graph=False)
        s.synchronize()
        t0 = time.perf_counter()
        s.output_rows()
        wall = time.perf_counter() - t0
    out["output_rows_2d"] = {"workload": "CSV row values of helium 2048x1024 computed on the device "
                                         "(FP64 primitives, energies, per-material fields) and copied "
                                         "to the host", "seconds": wall, "rows_per_s": c.ncells / wall}
    c = workloads.bubble_collapse_3d(512, build_state=False)
    with dimhip.Solver(c.dim, c.N, c.dx, c.materials, c.bc_lo, c.bc_hi, c.lambda_model,
                       c.alpha_source_floor, "split", "fp32_mixed") as s:
        s.synchronize()
        t0 = time.perf_counter()
        s.initialise(c.ic)
        s.synchronize()
        wall = time.perf_counter() - t0
    out["device_ic_3d"] = {"workload": "3D bubble collapse 512^3 initial state built on the device",
                           "seconds": wall, "cells_per_s": c.ncells / wall}
    return out


def measure_tube(dimhip, workloads, cpu):
    """The 1D water/air tube (1000 cells) run to t = 2.4e-4 s in the FP64 form
    with graph-captured steps; the reference's solver takes the same number of
    steps on one host thread."""
    case = workloads.water_air_tube_1d(1000)
    with dimhip.Solver.from_case(case, precision="fp64") as s:
        s.advance(1, case.cfl, 2.4e-4, 0.0, use_graph=True)  # graph capture outside the timing
        s.upload(case.U)
        t0 = time.perf_counter()
        t, n = s.advance(1 << 20, case.cfl, 2.4e-4, 0.0, use_graph=True)
        wall = time.perf_counter() - t0
    out = {"workload": "1D water/air shock tube, 1000 cells, FP64, to t = 2.4e-4 s",
           "steps": n, "final_time": t, "gpu_seconds": wall, "gpu_value": case.ncells * n / wall}
    if cpu:
        from oracle.pyoracle import Checker, make_params, ref_available
        if ref_available():
            ref = Checker("ref")
            p = make_params(1, case.materials)
            _, _, secs = ref.advance(p, case.N, case.dx, case.bc_lo, case.bc_hi, case.U, case.cfl, 1.0,
                                     nsteps=n, threads=1)
            out.update({"cpu_seconds": secs, "cpu_value": case.ncells * n / secs, "cpu_cores": 1,
                        "cpu_kind": "reference"})
    return out


def measure_batch(dimhip, workloads, count, steps, tuning):
    """Seeded uneven shock-interface instances, FP32-mixed, advanced together
    (autotuned tiles and schedule, hipGraph-captured steps).  The autotuner's
    choice for this instance set is read from the JSON file `tuning` when it
    holds one; otherwise every configuration is timed twice over 10 steps and
    the choice stored there."""
    cases = workloads.batch_instances(count=count)
    cells = sum(c.ncells for c in cases)
    with dimhip.Batch(cases) as b:
        tune = b.autotune(0.8, trial_steps=10, repeats=2, cache=tuning)
        b.advance(10, 0.8, use_graph=True)
        t0 = time.perf_counter()
        _, nsteps = b.advance(steps, 0.8, use_graph=True)
        wall = time.perf_counter() - t0
        b.set_timing(True)
        b.advance(20, 0.8, use_graph=False)
        ms, counts = b.kernel_times()
        dir_ms, _ = b.sweep_times()
        b.set_timing(False)
    if min(nsteps) != steps or max(nsteps) != steps:
        raise RuntimeError("batched instances did not all take the requested steps")
    per_cell = 2 * (8 * 2 + 4 * 4)  # FP64 masses + FP32 momentum, energy, fraction, read and written
    sweep_bytes_avg = per_cell * cells
    dom = 0 if dir_ms[0] >= dir_ms[1] else 1
    achieved = sweep_bytes_avg / (dir_ms[dom] * 1e-3) / 1e9
    return {
        "workload": f"{count} seeded shock-interface instances, 256..1024 x 256..512, FP32-mixed, "
                    f"{steps} steps each, one launch per stage, hipGraph-captured steps",
        "value": cells * steps / wall,
        "unit": "cell-updates/s",
        "cells": cells,
        "ms_per_step": wall / steps * 1e3,
        "autotuned": {"best": tune["best"], "cached": tune["cached"], "key": tune["key"],
                      "trial_steps": tune["trial_steps"], "repeats": tune["repeats"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "kernel": f"batched {AXES[dom]} sweep (the slower of the two persistent sweep launches)",
                     "per_axis_ms": {"x": dir_ms[0], "y": dir_ms[1]},
                     "mean_ms": {"sweep": ms[0], "boundary": ms[1], "cfl": ms[2]},
                     "bytes_per_launch": sweep_bytes_avg},
    }


def main():
    args = parse()
    if args.cpu_worker:
        cpu_worker(json.loads(args.cpu_worker))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from diffuse_vs_gfm_mms_amd import dimhip, workloads

    if args.sweep_probe:
        grid = WORKLOADS["helium_fp64" if args.sweep_probe == "fp64" else "helium_fp32"][0]
        case = workloads.helium_shock_bubble_2d(*grid)
        with dimhip.Solver.from_case(case, precision=args.sweep_probe) as s:
            t, _ = s.advance(5, case.cfl, 1.0, 0.0, use_graph=False)
            dt = s.compute_dt(case.cfl, 1.0) * 1e-3  # a small step: the state barely moves
            for d in (0, 1):
                for _ in range(30):
                    s.sweep(d, dt)
            s.synchronize()
        print(json.dumps({"sweep_probe": args.sweep_probe, "grid": list(grid), "steps": 5,
                          "unfused_sweeps_per_axis": 30}))
        return
    if args.copy_probe:
        import numpy as np
        src = np.arange(args.copy_probe, dtype=np.float64)
        out, _ = dimhip.kat(dimhip.KAT_COPY, 1, [("ideal_gas", 1.4, 0.0)], args.copy_probe, src)
        print(json.dumps({"copy_probe_doubles": args.copy_probe, "bytes_read": 8 * args.copy_probe,
                          "bytes_written": 8 * args.copy_probe, "ok": bool(np.array_equal(out, src))}))
        return

    grid, precision, _ = WORKLOADS[args.workload]
    progress(f"{args.workload}: {args.warmup} + {args.steps} steps")
    main_run = measure(dimhip, workloads, grid, precision, args.steps, args.warmup)
    case = main_run["case"]
    cells = case.ncells
    value = main_run["value"]
    wall = main_run["wall"]
    ms = main_run["ms"]
    counts = main_run["counts"]
    roof = sweep_roofline(case, precision, ms, main_run["dir_ms"])
    dom = roof["kernel"][0]
    roof["traffic"] = traffic_for(args.workload, dom)
    roof["launches"] = {"sweep": counts[0], "boundary": counts[1], "cfl": counts[2]}
    roof["timed_steps"] = f"1 in {TIMING_EVERY}"
    valu = valu_for(args.workload)
    if valu:
        roof["limiter"] = valu.get("limiter")
        roof["valu"] = valu.get(f"{dom}_sweep")
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
        "roofline": roof,
        "cpu_baseline": None,
    }
    if args.fp32_steps > 0 and precision == "fp64":
        g32 = WORKLOADS["helium_fp32"][0]
        progress(f"fp32_mixed {g32[0]}x{g32[1]}: {args.fp32_steps} steps")
        r32 = measure(dimhip, workloads, g32, "fp32_mixed", args.fp32_steps, args.warmup)
        roof32 = sweep_roofline(r32["case"], "fp32_mixed", r32["ms"], r32["dir_ms"])
        roof32["traffic"] = traffic_for("helium_fp32", roof32["kernel"][0])
        v32 = valu_for("helium_fp32")
        if v32:
            roof32["limiter"] = v32.get("limiter")
            roof32["valu"] = v32.get(f"{roof32['kernel'][0]}_sweep")
        out["fp32_mixed"] = {
            "workload": f"helium shock-bubble {g32[0]}x{g32[1]}, fp32_mixed, split x->y, CFL 0.8",
            "value": r32["value"],
            "unit": "cell-updates/s",
            "steps": args.fp32_steps,
            "ms_per_step": r32["wall"] / args.fp32_steps * 1e3,
            "roofline": roof32,
        }
        if args.fp64_c3_steps > 0:
            progress(f"fp64 {g32[0]}x{g32[1]}: {args.fp64_c3_steps} steps (the FP32-mixed ratio on one grid)")
            r64 = measure(dimhip, workloads, g32, "fp64", args.fp64_c3_steps, args.warmup)
            out["fp32_mixed"]["fp64_same_grid"] = {
                "workload": f"helium shock-bubble {g32[0]}x{g32[1]}, fp64, split x->y, CFL 0.8",
                "value": r64["value"], "steps": args.fp64_c3_steps,
                "ms_per_step": r64["wall"] / args.fp64_c3_steps * 1e3,
                "roofline": sweep_roofline(r64["case"], "fp64", r64["ms"], r64["dir_ms"]),
            }
            out["fp32_mixed"]["speedup_vs_fp64_same_grid"] = r32["value"] / r64["value"]
        acc_steps = args.fp32_steps if args.accuracy_steps < 0 else args.accuracy_steps
        if acc_steps > 0:
            progress(f"fp32_mixed vs fp64 accuracy: {acc_steps} steps of each")
            out["fp32_mixed"]["accuracy_vs_fp64"] = accuracy_fp32(dimhip, workloads, g32, acc_steps)
    if args.batch_steps > 0 and precision == "fp64":
        progress(f"batched {args.batch_count} instances: {args.batch_steps} steps")
        out["batched"] = measure_batch(dimhip, workloads, args.batch_count, args.batch_steps, args.batch_tuning)
        if args.accuracy_steps != 0:
            progress(f"batched instances 0..3 fp32_mixed vs fp64 accuracy: {args.batch_steps} steps")
            out["batched"]["accuracy_vs_fp64"] = accuracy_batch(dimhip, workloads, args.batch_count,
                                                                args.batch_steps, 4)
    if args.c5_steps > 0 and precision == "fp64":
        progress(f"3D bubble collapse {args.c5_n}^3: {args.c5_steps} steps")
        out["bubble_collapse_3d"] = measure_3d(dimhip, workloads, args.c5_n, args.c5_steps, min(args.warmup, 5))
        if args.accuracy_steps != 0 and args.c5_accuracy_n > 0:
            n3 = args.c5_accuracy_n
            progress(f"3D bubble collapse {n3}^3 fp32_mixed vs fp64 accuracy: {args.c5_accuracy_steps} steps")
            out["bubble_collapse_3d"]["accuracy_vs_fp64"] = accuracy_vs_fp64(
                dimhip, workloads.bubble_collapse_3d(n3, build_state=False), args.c5_accuracy_steps,
                f"3D bubble collapse {n3}^3")
    if args.next_steps > 0 and precision == "fp64":
        progress(f"widened rows: {args.next_steps} FP64 steps each")
        out["next_rows"] = measure_next_rows(dimhip, workloads, args.next_steps)
    if args.tube and precision == "fp64":
        progress("1D water/air tube full run")
        out["water_air_tube_1d"] = measure_tube(dimhip, workloads, not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        progress("reference CPU solver: helium 2048x1024")
        out["cpu_baseline"] = cpu_baseline("helium_2048x1024", args.cpu_steps)
        if args.cpu_all and precision == "fp64":
            others = {}
            for name in ("helium_4096x2048", "batch_64", "bubble_collapse_128"):
                progress(f"reference CPU solver: {name}")
                others[name] = cpu_baseline(name)
            out["cpu_baseline"]["other_configs"] = others
    print(json.dumps(out))


if __name__ == "__main__":
    main()
