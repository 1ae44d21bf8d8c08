"""The case driver (dim_main_<D>d) against the reference's own driver
(oracle/_ref/mm_main_<D>d, built from the reference's sources) on the same
case files: every solution CSV must be byte-identical, the runtime reports
must agree on steps and final time, and the conservation reports
(SOLVER_CONSERVATION=1) must agree to FP64 summation-order noise.

Also checks the I/O entry points on their own: dimhip_output_rows against the
oracle's cons_to_prim, dimhip_upload_primitive against the oracle's
prim_to_cons, and dimhip_boundary_flux against the flux through the faces
recomputed from oracle primitives (conservation_report.hpp:253-305).
"""
import csv
import os
import subprocess

import numpy as np
import pytest

from oracle.pyoracle import make_params, ref_available, ref_initial_state
from diffuse_vs_gfm_mms_amd import dimhip, workloads as W
from diffuse_vs_gfm_mms_amd.build import driver_path

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = os.path.join(ROOT, "tests", "cases")
REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "mm_main_{}d")

CASE_FILES = {
    "tube_1d_diffuse.txt": 1,
    "water_air_1d_sharp.txt": 1,
    "helium_bubble_2d.txt": 2,
    "oblique_2d.txt": 2,
    "coated_bubble_2d.txt": 2,
    "box_regions_2d.txt": 2,
    "bubble_collapse_3d.txt": 3,
    "gorsse_tc9_2d.txt": 2,
    "three_material_1d.txt": 1,
}


def write_case(tmp_path, name, outdir):
    text = open(os.path.join(CASES, name)).read().replace("OUTDIR", str(outdir))
    p = tmp_path / f"{outdir.name}_{name}"
    p.write_text(text)
    return str(p)


def run(exe, case, env_extra=None, cwd=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    print("running", exe, case, flush=True)
    res = subprocess.run([exe, case], capture_output=True, text=True, env=env, timeout=300, cwd=cwd)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def files_under(d):
    out = {}
    for base, _, names in os.walk(d):
        for n in names:
            out[os.path.relpath(os.path.join(base, n), d)] = os.path.join(base, n)
    return out


def report(path):
    kv = {}
    for line in open(path):
        if " = " in line:
            k, v = line.rstrip("\n").split(" = ", 1)
            kv[k] = v
    return kv


needs_ref = pytest.mark.skipif(not all(os.path.exists(REF_MAIN.format(d)) for d in (1, 2, 3)),
                               reason="reference drivers (oracle/_ref) not built")


@needs_ref
@pytest.mark.parametrize("name", sorted(CASE_FILES))
def test_driver_csv_matches_reference_driver(tmp_path, name):
    dim = CASE_FILES[name]
    ours, theirs = tmp_path / "ours", tmp_path / "ref"
    out_a = run(driver_path(dim), write_case(tmp_path, name, ours))
    out_b = run(REF_MAIN.format(dim), write_case(tmp_path, name, theirs))
    fa, fb = files_under(ours), files_under(theirs)
    assert sorted(fa) == sorted(fb)
    csvs = [k for k in fa if k.endswith(".csv")]
    assert csvs
    for k in csvs:
        a, b = open(fa[k], "rb").read(), open(fb[k], "rb").read()
        assert a == b, f"{name}: {k} differs"
    runtime = [k for k in fa if k.endswith("_runtime.txt")]
    assert runtime
    for k in runtime:
        ra, rb = report(fa[k]), report(fb[k])
        for key in ("interface_method", "dimension", "N", "cells", "steps", "final_time"):
            assert ra[key] == rb[key], (name, key, ra[key], rb[key])
    # console lines other than timings are the reference's
    keep = lambda s: [l.replace(str(ours), "D").replace(str(theirs), "D") for l in s.splitlines()]
    assert keep(out_a) == keep(out_b)


@needs_ref
@pytest.mark.parametrize("name,interval", [("helium_bubble_2d.txt", "1"), ("coated_bubble_2d.txt", "3"),
                                           ("tube_1d_diffuse.txt", "5")])
def test_conservation_report_matches_reference_driver(tmp_path, name, interval):
    dim = CASE_FILES[name]
    env = {"SOLVER_CONSERVATION": "1", "SOLVER_CONSERVATION_INTERVAL": interval}
    ours, theirs = tmp_path / "ours", tmp_path / "ref"
    run(driver_path(dim), write_case(tmp_path, name, ours), env)
    run(REF_MAIN.format(dim), write_case(tmp_path, name, theirs), env)
    fa, fb = files_under(ours), files_under(theirs)
    keys = [k for k in fa if k.endswith("_conservation.csv")]
    assert keys and sorted(keys) == sorted(k for k in fb if k.endswith("_conservation.csv"))
    for key in keys:
        compare_conservation(name, dim, fa[key], fb[key])


def compare_conservation(name, dim, path_a, path_b):
    ra = list(csv.reader(open(path_a)))
    rb = list(csv.reader(open(path_b)))
    assert ra[0] == rb[0] and len(ra) == len(rb)
    A = np.array(ra[1:], dtype=np.float64)
    B = np.array(rb[1:], dtype=np.float64)
    assert np.array_equal(A[:, :2], B[:, :2])  # step, time
    # totals and integrals are FP64 sums in another order: compare each column
    # against the size of the totals of its quantity (all momentum components
    # share one scale: a component whose total cancels to zero still sums
    # terms of the other components' size)
    header = ra[0]
    col = {h: j for j, h in enumerate(header)}
    mom = max(np.max(np.abs(B[:, col[f"momentum{d}"]])) for d in range(dim))
    base = {"mass": np.max(np.abs(B[:, col["mass"]])), "energy": np.max(np.abs(B[:, col["energy"]]))}
    for j, h in enumerate(header[2:], start=2):
        if h.startswith("eps_") or h.startswith("normalized_") or h.startswith("near_zero"):
            continue
        q = h.split("_")[-1] if "_" in h else h
        S = mom if q.startswith("momentum") else base["energy"] if q == "energy" else base["mass"]
        # the files hold 6 significant digits: one unit of the last digit is
        # up to 1e-5 of the value
        tol = 1e-12 * S + 1e-5 * np.abs(B[:, j])
        assert np.all(np.abs(A[:, j] - B[:, j]) <= tol), (name, h)


def test_output_rows_and_primitive_upload_match_oracle(oracle):
    c = W.helium_shock_bubble_2d(96, 40)
    p = make_params(2, c.materials)
    with dimhip.Solver.from_case(c, precision="fp64") as s:
        s.advance(5, c.cfl, 1.0)
        U = s.download()
        rows = s.output_rows()
        P = oracle.cons_to_prim(p, U)
        nm = c.nmat
        rho = U[:, :nm].sum(1)
        v2 = P[:, 2 * nm] * P[:, 2 * nm] + P[:, 2 * nm + 1] * P[:, 2 * nm + 1]
        e = np.maximum((U[:, nm + 2] - 0.5 * rho * v2) / rho, 0.0)
        expect = np.column_stack([rho, P[:, 2 * nm:2 * nm + 2], P[:, -1], e, P[:, nm:2 * nm], P[:, :nm], U[:, :nm]])
        assert np.array_equal(rows.view(np.uint64), expect.view(np.uint64))
        # primitive upload = oracle prim_to_cons
        s.upload_primitive(P)
        U2 = s.download()
        assert np.array_equal(U2.view(np.uint64), oracle.prim_to_cons(p, P).view(np.uint64))


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_boundary_flux_and_its_integral(oracle, precision):
    c = W.helium_shock_bubble_2d(64, 32)
    p = make_params(2, c.materials)
    nx, ny = c.N
    dx, dy = c.dx
    with dimhip.Solver.from_case(c, precision=precision) as s:
        s.advance(3, c.cfl, 1.0)
        U = s.download()
        got = s.boundary_flux()
        P = oracle.cons_to_prim(p, U)
        g = U.reshape(ny, nx, -1)
        q = P.reshape(ny, nx, -1)
        expect = np.zeros(6)

        def add(cells, prims, d, area):
            un = prims[:, 4 + d]
            rho = cells[:, :2].sum(1)
            expect[0:2] += (area * cells[:, :2] * un[:, None]).sum(0)
            expect[2] += (area * cells[:, :2] * un[:, None]).sum()
            for e in range(2):
                expect[3 + e] += (area * (rho * prims[:, 4 + e] * un + (prims[:, -1] if e == d else 0.0))).sum()
            expect[5] += (area * (cells[:, 4] + prims[:, -1]) * un).sum()

        add(g[:, -1], q[:, -1], 0, dy)
        add(g[:, 0], q[:, 0], 0, -dy)
        add(g[-1, :], q[-1, :], 1, dx)
        add(g[0, :], q[0, :], 1, -dx)
        scale = np.abs(expect).max()
        assert np.allclose(got, expect, rtol=1e-12 if precision == "fp64" else 1e-6, atol=1e-12 * scale)
        # the integral accumulates dt * flux of each new state
        s.track_boundary_flux(True)
        assert np.all(s.boundary_flux_integral() == 0.0)
        acc = np.zeros(6)
        t = 0.0
        for _ in range(3):
            dt = s.step(c.cfl, 1.0)
            acc += dt * s.boundary_flux()
        assert np.allclose(s.boundary_flux_integral(), acc, rtol=1e-13, atol=1e-14 * np.abs(acc).max())
        # device loop and graph replays accumulate the same way
        s.track_boundary_flux(True)
        before = s.download()
        t, n = s.advance(4, c.cfl, 1.0, 0.0, use_graph=True)
        integ = s.boundary_flux_integral()
        s.upload(before)
        s.track_boundary_flux(True)
        acc = np.zeros(6)
        for _ in range(4):
            dt = s.step(c.cfl, 1.0)
            acc += dt * s.boundary_flux()
        assert np.allclose(integ, acc, rtol=1e-13, atol=1e-14 * np.abs(acc).max())


def read_csv_numbers(path):
    rows = list(csv.reader(open(path)))
    return rows[0], np.array(rows[1:], dtype=np.float64)


@pytest.mark.parametrize("name", sorted(CASE_FILES))
def test_device_initial_conditions_match_host(tmp_path, name):
    """--device-ic (dimhip_initialise) against the host-built initial state,
    whose bits are the reference's, written at t = 0: equal to the CSV's six
    digits (the device and host tanh may differ in the last bit in diffuse
    cells; later states are not compared, since a last-bit difference can flip
    a limiter choice and grow)."""
    dim = CASE_FILES[name]
    text = open(os.path.join(CASES, name)).read()
    lines = [l for l in text.splitlines() if not l.strip().startswith("output_times")]
    lines.append("output_times = [0.0]")
    src = tmp_path / name
    src.write_text("\n".join(lines) + "\n")
    outs = {}
    for tag, extra in (("host", []), ("dev", ["--device-ic"])):
        out = tmp_path / tag
        case = tmp_path / f"{tag}_{name}"
        case.write_text(src.read_text().replace("OUTDIR", str(out)))
        print("running", tag, case, flush=True)
        res = subprocess.run([driver_path(dim), str(case)] + extra, capture_output=True, text=True, timeout=300)
        assert res.returncode == 0, res.stdout + res.stderr
        outs[tag] = files_under(out)
    fa, fb = outs["host"], outs["dev"]
    csvs = sorted(k for k in fa if "_t0p000000ep00" in k)
    assert csvs and csvs == sorted(k for k in fb if "_t0p000000ep00" in k)
    for k in csvs:
        ha, A = read_csv_numbers(fa[k])
        hb, B = read_csv_numbers(fb[k])
        assert ha == hb and A.shape == B.shape
        scale = np.maximum(np.abs(A).max(axis=0), 1e-300)
        tol = 1e-5 * np.abs(A) + 1e-14 * scale
        bad = np.argwhere(np.abs(A - B) > tol)
        detail = [(ha[j], int(i), A[i, j], B[i, j]) for i, j in bad[:8]]
        assert len(bad) == 0, (name, k, len(bad), detail)


# case files whose interfaces are sharp (interface_thickness = 0): no tanh
SHARP = {"water_air_1d_sharp.txt"}


def dump_initial(tmp_path, name, extra=()):
    """The driver's packed FP64 initial state (DIMHIP_DUMP_INITIAL) of the
    first resolution of a case file, taking no step (DIM_MAX_STEPS=0)."""
    dim = CASE_FILES[name]
    out = tmp_path / "out"
    case = write_case(tmp_path, name, out)
    dump = tmp_path / "ic"
    env = dict(os.environ, DIMHIP_DUMP_INITIAL=str(dump), DIM_MAX_STEPS="0")
    print("running", driver_path(dim), case, *extra, flush=True)
    res = subprocess.run([driver_path(dim), case, *extra], capture_output=True, text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return case, np.fromfile(str(dump) + ".0", dtype=np.float64)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", sorted(CASE_FILES))
def test_device_initial_state_matches_reference_fp64(tmp_path, name):
    """dimhip_initialise (--device-ic) against the reference's own initial
    state (initialise_dim_from_config, dim_initial_conditions.hpp:16-628,
    from oracle/_ref), packed FP64: bit for bit where interfaces are sharp;
    where they are tanh-blended the device tanh may differ from the host C
    library's in its last bit, and every field stays within 1e-15 of its
    scale (the largest |value| of the field) -- a bound the host-built rows
    (dimhip_upload_primitive) meet with equality of bits."""
    dim = CASE_FILES[name]
    case, U = dump_initial(tmp_path, name, ["--device-ic"])
    U_ref = ref_initial_state(case, dim, 0)
    U = U.reshape(U_ref.shape)
    host = tmp_path / "host"
    host.mkdir()
    _, U_host = dump_initial(host, name)
    assert np.array_equal(U_host.reshape(U_ref.shape).view(np.uint64), U_ref.view(np.uint64))
    same = np.all(U.view(np.uint64) == U_ref.view(np.uint64), axis=1)
    scale = np.maximum(np.abs(U_ref).max(axis=0), 1e-300)
    rel = (np.abs(U - U_ref) / scale).max(axis=0)
    print(f"{name}: {int((~same).sum())} of {len(same)} cells differ in any bit; max |diff|/scale per field {rel}")
    if name in SHARP:
        assert same.all()
    assert np.all(rel <= 1e-15), rel


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
def test_step_cap_matches_reference_run_to_the_same_time(tmp_path):
    """DIM_MAX_STEPS=K (tfinal far beyond the K steps, no output times): the
    driver stops after K steps and writes its end-of-run CSV.  The reference's
    driver, which has no step cap, is run to the time of those K steps
    (tfinal = t_K to 17 digits): same step count, and the same CSV values to
    the six digits both print (its last step may be clamped by an ulp of t)."""
    K = 37
    name = "helium_bubble_2d.txt"
    text = open(os.path.join(CASES, name)).read()
    base = [l for l in text.splitlines() if not l.strip().startswith(("output_times", "tfinal"))]

    def case_with(tag, tfinal):
        p = tmp_path / f"{tag}.txt"
        p.write_text("\n".join(base + [f"tfinal = {tfinal!r}", "output_times = []"]).replace(
            "OUTDIR", str(tmp_path / tag)) + "\n")
        return str(p)

    # t_K: the FP64 path from the driver's initial state (the reference's bits)
    c0 = case_with("probe", 1.0)
    env = dict(os.environ, DIMHIP_DUMP_INITIAL=str(tmp_path / "ic"), DIM_MAX_STEPS="0")
    res = subprocess.run([driver_path(2), c0], capture_output=True, text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    U0 = np.fromfile(str(tmp_path / "ic") + ".0", dtype=np.float64).reshape(-1, 6)
    mats = [("ideal_gas", 1.4, 0.0), ("ideal_gas", 1.648, 0.0)]
    with dimhip.Solver(2, (128, 36), (0.325 / 128, 0.089 / 36), mats, ("transmissive", "reflective"),
                       ("transmissive", "reflective")) as s:
        s.upload(U0)
        tK, n = s.advance(K, 0.8, 1.0, 0.0, use_graph=False)
    assert n == K
    ours = case_with("ours", 1.0)
    env = dict(os.environ, DIM_MAX_STEPS=str(K))
    res = subprocess.run([driver_path(2), ours], capture_output=True, text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    run(REF_MAIN.format(2), case_with("ref", tK))
    fa, fb = files_under(tmp_path / "ours"), files_under(tmp_path / "ref")
    assert sorted(fa) == sorted(fb)
    for k in fa:
        if k.endswith("_runtime.txt"):
            ra = dict(l.split(" = ", 1) for l in open(fa[k]).read().splitlines() if " = " in l)
            rb = dict(l.split(" = ", 1) for l in open(fb[k]).read().splitlines() if " = " in l)
            assert ra["steps"] == rb["steps"] == str(K)
        elif k.endswith(".csv"):
            ha, A = read_csv_numbers(fa[k])
            hb, B = read_csv_numbers(fb[k])
            assert ha == hb and A.shape == B.shape
            assert np.allclose(A, B, rtol=2e-6, atol=1e-12 * np.abs(B).max()), k


def test_device_initial_condition_matches_numpy_builder():
    """The helium shock-bubble generated on the device against the numpy
    workload builder, a looser restatement of the same geometry: the builder
    omits prim_to_cons's 1e-12 volume-fraction floor (primitives.hpp:166-212),
    so pure cells differ by that floor's mass and energy, ~1e-12 of each
    field's scale.  A geometry or blending mistake would be of order one."""
    c = W.helium_shock_bubble_2d(256, 128)
    with dimhip.Solver(c.dim, c.N, c.dx, c.materials, c.bc_lo, c.bc_hi) as s:
        s.initialise(c.ic)
        U = s.download()
    assert U.shape == c.U.shape
    scale = np.abs(c.U).max(axis=0)
    assert np.all(np.abs(U - c.U) <= 1e-11 * scale)


def test_device_initial_condition_errors():
    c = W.helium_shock_bubble_2d(64, 32)
    with dimhip.Solver(c.dim, c.N, c.dx, c.materials, c.bc_lo, c.bc_hi) as s:
        # a region set that leaves cells uncovered, sharp interface
        ic = dimhip.make_ic("regions", (0.0, 0.0), (0.325, 0.089), 0.0,
                            regions=[((0.0, 0.0), (0.1, 0.089), 1.0, (0.0, 0.0), 1e5, 0)])
        with pytest.raises(RuntimeError, match="point not covered by any region"):
            s.initialise(ic)
        bad = dict(c.ic)
        bad["bubble"] = (0.2, (0.0, 0.0), 1e5, 5)
        with pytest.raises(RuntimeError, match="material id not found"):
            s.initialise(bad)
