"""Noble-Abel materials (NobleAbelGasEOS, src/euler/eos.hpp:663-738) on the
FP64 HIP path, bit for bit against the reference's golden fixtures and the
oracle, through the C ABI."""
import numpy as np
import pytest

from _states import MATSETS, bits_equal, mismatch_report, random_line
from oracle.pyoracle import make_params
from diffuse_vs_gfm_mms_amd import dimhip, workloads as W

pytestmark = pytest.mark.gpu
LAMBDAS = ("kapila", "allaire", "equal_velocity")
NA_SETS = ("water_na_air", "water_na_air_helium", "na_air")


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("ms", NA_SETS)
@pytest.mark.parametrize("lm", LAMBDAS)
def test_device_functions_match_golden(golden, dim, ms, lm):
    g = golden["noble_abel"]
    key = f"d{dim}_{ms}_{lm}"
    mats = MATSETS[ms]
    U, V = g[key + "_U"], g[key + "_V"]
    n = U.shape[0]
    P, _ = dimhip.kat(dimhip.KAT_CONS_TO_PRIM, dim, mats, n, U, lambda_model=lm)
    P = P.reshape(g[key + "_c2p"].shape)
    assert bits_equal(P, g[key + "_c2p"]), mismatch_report(P, g[key + "_c2p"])
    C, _ = dimhip.kat(dimhip.KAT_PRIM_TO_CONS, dim, mats, n, g[key + "_c2p"], lambda_model=lm)
    assert bits_equal(C.reshape(U.shape), g[key + "_p2c"])
    R, _ = dimhip.kat(dimhip.KAT_REPAIR, dim, mats, n, U, lambda_model=lm)
    assert bits_equal(R.reshape(U.shape), g[key + "_repair"]), mismatch_report(R.reshape(U.shape), g[key + "_repair"])
    rhs, _ = dimhip.kat(dimhip.KAT_ALPHA_RHS, dim, mats, n, g[key + "_c2p"], lambda_model=lm)
    assert bits_equal(rhs.reshape(g[key + "_rhs"].shape), g[key + "_rhs"])
    for d in range(dim):
        F, fv = dimhip.kat(dimhip.KAT_HLLC, dim, mats, n, U, V, d=d, lambda_model=lm)
        ref_F = g[key + f"_hllc{d}_F"]
        assert bits_equal(F.reshape(ref_F.shape), ref_F), mismatch_report(F.reshape(ref_F.shape), ref_F)
        assert bits_equal(fv, g[key + f"_hllc{d}_fv"])


def _line_solver(dim, d, n, mats):
    N = [1] * dim
    N[d] = n
    dx = [1.0] * dim
    dx[d] = 1e-3
    return dimhip.Solver(dim, N, dx, mats, ["transmissive"] * dim, ["transmissive"] * dim)


def test_sweeps_match_golden_lines(golden):
    g = golden["noble_abel"]
    keys = sorted({k[:-2] for k in g.files if k.startswith("line_") and k.endswith("_U")})
    assert len(keys) == 18
    for key in keys:
        dim = int(key[6])
        n = int(key.split("_n")[1].split("_")[0])
        d = int(key.split("dir")[1])
        with _line_solver(dim, d, n, MATSETS["water_na_air"]) as s:
            s.upload(g[key + "_U"])
            s.sweep(d, 2e-8)
            out = s.download()
        assert bits_equal(out, g[key + "_out"]), (key, mismatch_report(out, g[key + "_out"]))


def test_sweeps_match_oracle_random_lines(oracle):
    rng = np.random.default_rng(43)
    for dim in (1, 2, 3):
        for ms in NA_SETS:
            mats = MATSETS[ms]
            p = make_params(dim, mats)
            for n in (3, 9, 123, 300):
                d = int(rng.integers(0, dim))
                U = random_line(mats, dim, n, rng)
                ref = oracle.solve_line(p, d, U, 1e-3, 3e-8)
                with _line_solver(dim, d, n, mats) as s:
                    s.upload(U)
                    s.sweep(d, 3e-8)
                    out = s.download()
                assert bits_equal(out, ref), (dim, ms, n, d, mismatch_report(out, ref))


def test_steps_match_golden(golden):
    """The coarse Gorsse TC9 water/noble-Abel-air bubble (5 steps) and the 1D
    water/noble-Abel-air tube (20 steps) against the reference's states."""
    g = golden["noble_abel"]
    c = W.gorsse_tc9_2d(48, 40)
    with dimhip.Solver.from_case(c) as s:
        s.upload(g["gorsse_U0"])
        dts = np.array([s.step(c.cfl, 1.0) for _ in range(5)])
        out = s.download()
    assert bits_equal(out, g["gorsse_U5"]), mismatch_report(out, g["gorsse_U5"])
    assert bits_equal(dts, g["gorsse_dt"])
    c = W.water_na_air_tube_1d(200)
    with dimhip.Solver.from_case(c) as s:
        s.upload(g["tube_U0"])
        dts = np.array([s.step(c.cfl, 1.0) for _ in range(20)])
        out = s.download()
    assert bits_equal(out, g["tube_U20"]), mismatch_report(out, g["tube_U20"])
    assert bits_equal(dts, g["tube_dt"])


@pytest.mark.parametrize("variant", ["split", "unsplit", "nonreflective", "three_material"])
def test_gorsse_steps_match_oracle(oracle, variant):
    """Twenty device-loop steps of a 120x100 Gorsse TC9 bubble (graph-captured)
    against the oracle: split, unsplit, nonreflective ends, and with helium as
    a third material in the bubble."""
    c = W.gorsse_tc9_2d(120, 100)
    lo, hi, tu = c.bc_lo, c.bc_hi, "split"
    if variant == "unsplit":
        tu = "unsplit"
    if variant == "nonreflective":
        lo, hi = ("nonreflective", "reflective"), ("nonreflective", "reflective")
    U = c.U
    mats = c.materials
    if variant == "three_material":
        mats = [W.WATER, W.NA_AIR, W.HELIUM]
        # split the air partial mass: the lower half of the domain holds helium
        # packed 2D nmat-3 cells: m0 m1 m2 mom0 mom1 E a0 a1
        U = np.zeros((c.ncells, 8))
        lower = np.arange(c.ncells) // c.N[0] < c.N[1] // 2
        U[:, 0] = c.U[:, 0]
        U[:, 1] = np.where(lower, 0.0, c.U[:, 1])
        U[:, 2] = np.where(lower, c.U[:, 1] * 0.214 / 1.2, 0.0)
        U[:, 3:5] = c.U[:, 2:4]
        U[:, 5] = c.U[:, 4]
        U[:, 6] = c.U[:, 5]
        U[:, 7] = np.where(lower, 0.0, 1.0 - c.U[:, 5])
        # E is kept: the helium cells then hold a different pressure, a valid state
    p = make_params(2, mats)
    ref, ref_dts, _ = oracle.advance(p, c.N, c.dx, lo, hi, U, c.cfl, 1.0, nsteps=20, unsplit=(tu == "unsplit"))
    with dimhip.Solver(2, c.N, c.dx, mats, lo, hi, time_update=tu) as s:
        s.upload(U)
        t, steps = s.advance(20, c.cfl, t_end=1.0, time=0.0, use_graph=True)
        out = s.download()
    t_ref = 0.0
    for dt in ref_dts:  # the device accumulates time += dt step by step
        t_ref += float(dt)
    assert steps == 20
    assert t == t_ref
    assert bits_equal(out, ref), mismatch_report(out, ref)


def test_fp32_mixed_rejects_noble_abel():
    c = W.gorsse_tc9_2d(16, 16)
    with pytest.raises(RuntimeError, match="FP32-mixed form supports ideal and stiffened gases only"):
        dimhip.Solver(2, c.N, c.dx, c.materials, c.bc_lo, c.bc_hi, precision="fp32_mixed")
