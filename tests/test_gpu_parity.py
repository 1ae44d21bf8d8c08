"""FP64 HIP path against the reference's golden fixtures and the oracle, bit
for bit, through the C ABI (include/dimhip.h)."""
import numpy as np
import pytest

from _states import MATSETS, bits_equal, mismatch_report, random_line, random_states, rel_linf
from oracle.pyoracle import Checker, make_params, ref_available
from diffuse_vs_gfm_mms_amd import dimhip, workloads as W

pytestmark = pytest.mark.gpu
LAMBDAS = ("kapila", "allaire", "equal_velocity")


def test_ieee_division_and_sqrt_are_correctly_rounded():
    """The reference's x86-64 build rounds / and sqrt correctly; so must the device."""
    rng = np.random.default_rng(1)
    n = 200000
    a = rng.uniform(0.5, 1.0, n) * 2.0 ** rng.integers(-1060, 1000, n)
    b = rng.uniform(0.5, 1.0, n) * 2.0 ** rng.integers(-1000, 1000, n)
    a[:1000] = rng.uniform(0, 1, 1000) * 2.0 ** -1030  # subnormals
    pairs = np.stack([np.abs(a), b], 1).reshape(-1)
    q, s = dimhip.kat(dimhip.KAT_IEEE, 1, MATSETS["air"], n, pairs)
    assert bits_equal(q, np.abs(a) / b), mismatch_report(q, np.abs(a) / b)
    assert bits_equal(s, np.sqrt(np.abs(a))), mismatch_report(s, np.sqrt(np.abs(a)))


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("ms", ["water_air", "air_helium", "water_air_helium"])
@pytest.mark.parametrize("lm", LAMBDAS)
def test_device_functions_match_golden(golden, dim, ms, lm):
    g = golden["functions"]
    key = f"d{dim}_{ms}_{lm}"
    mats = MATSETS[ms]
    U, V = g[key + "_U"], g[key + "_V"]
    n = U.shape[0]
    P, _ = dimhip.kat(dimhip.KAT_CONS_TO_PRIM, dim, mats, n, U, lambda_model=lm)
    assert bits_equal(P.reshape(g[key + "_c2p"].shape), g[key + "_c2p"]), mismatch_report(
        P.reshape(g[key + "_c2p"].shape), g[key + "_c2p"])
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
    """One directional sweep over a single line equals dim::solve_line."""
    g = golden["lines"]
    keys = sorted({k[:-2] for k in g.files if k.endswith("_U")})
    for key in keys:
        dim = int(key[1])
        ms = "water_air" if "water_air" in key else "air_helium"
        n = int(key.split("_n")[1].split("_")[0])
        d = int(key.split("dir")[1])
        with _line_solver(dim, d, n, MATSETS[ms]) as s:
            s.upload(g[key + "_U"])
            s.sweep(d, 2e-8)
            out = s.download()
        assert bits_equal(out, g[key + "_out"]), (key, mismatch_report(out, g[key + "_out"]))


def test_sweeps_match_oracle_random_lines(oracle):
    rng = np.random.default_rng(42)
    for dim in (1, 2, 3):
        for ms in ("water_air", "air_helium", "air", "water_air_helium"):
            mats = MATSETS[ms]
            p = make_params(dim, mats)
            for n in (3, 5, 9, 120, 123, 250, 300):
                d = int(rng.integers(0, dim))
                U = random_line(mats, dim, n, rng)
                ref = oracle.solve_line(p, d, U, 1e-3, 3e-8)
                with _line_solver(dim, d, n, mats) as s:
                    s.upload(U)
                    s.sweep(d, 3e-8)
                    out = s.download()
                assert bits_equal(out, ref), (dim, ms, n, d, mismatch_report(out, ref))


def _run_steps(case, nsteps, bc_lo=None, bc_hi=None, time_update="split", dt_max=1.0):
    with dimhip.Solver(case.dim, case.N, case.dx, case.materials, bc_lo or case.bc_lo, bc_hi or case.bc_hi,
                       time_update=time_update) as s:
        s.upload(case.U)
        dts = [s.step(case.cfl, dt_max) for _ in range(nsteps)]
        return s.download(), np.array(dts)


def test_steps_match_golden(golden):
    g = golden["steps"]
    c = W.water_air_tube_1d(200)
    with dimhip.Solver.from_case(c) as s:
        done, dts = 0, []
        for k in (1, 10, 50):
            dts += [s.step(c.cfl, 1.0) for _ in range(k - done)]
            done = k
            out = s.download()
            assert bits_equal(out, g[f"tube1d_U{k}"]), (k, mismatch_report(out, g[f"tube1d_U{k}"]))
        assert bits_equal(np.array(dts), np.concatenate([g["tube1d_dt1"], g["tube1d_dt10"], g["tube1d_dt50"]]))
    c = W.helium_shock_bubble_2d(40, 20)
    combos = {
        "tr_refl": (c.bc_lo, c.bc_hi, "split"),
        "mixed": (("reflective", "reflective"), ("nonreflective", "reflection"), "split"),
        "unsplit": (c.bc_lo, c.bc_hi, "unsplit"),
    }
    for name, (lo, hi, tu) in combos.items():
        out, dts = _run_steps(c, 5, lo, hi, tu)
        assert bits_equal(out, g[f"bubble2d_{name}_U5"]), (name, mismatch_report(out, g[f"bubble2d_{name}_U5"]))
        assert bits_equal(dts, g[f"bubble2d_{name}_dt"]), name
    c = W.bubble_collapse_3d(10)
    out, dts = _run_steps(c, 2)
    assert bits_equal(out, g["collapse3d_U2"]), mismatch_report(out, g["collapse3d_U2"])
    assert bits_equal(dts, g["collapse3d_dt"])


@pytest.mark.parametrize("lm", LAMBDAS)
@pytest.mark.parametrize("bcs", ["tr_refl", "all_refl", "nonrefl", "reflection"])
def test_2d_steps_match_oracle(oracle, lm, bcs):
    c = W.helium_shock_bubble_2d(96, 40)
    c.lambda_model = lm
    lo, hi = {"tr_refl": (c.bc_lo, c.bc_hi),
              "all_refl": (("reflective",) * 2, ("reflective",) * 2),
              "nonrefl": (("nonreflective",) * 2, ("transmissive", "nonreflective")),
              "reflection": (("reflection",) * 2, ("reflective", "reflection"))}[bcs]
    p = make_params(2, c.materials, lm)
    ref, ref_dt, _ = oracle.advance(p, c.N, c.dx, lo, hi, c.U, c.cfl, 1.0, nsteps=4)
    with dimhip.Solver(2, c.N, c.dx, c.materials, lo, hi, lambda_model=lm) as s:
        s.upload(c.U)
        dts = [s.step(c.cfl, 1.0) for _ in range(4)]
        out = s.download()
    assert bits_equal(out, ref), mismatch_report(out, ref)
    assert bits_equal(np.array(dts), ref_dt)


def test_3d_steps_match_oracle(oracle):
    c = W.bubble_collapse_3d(18)
    lo = ("transmissive", "reflective", "nonreflective")
    hi = ("reflective", "transmissive", "transmissive")
    p = make_params(3, c.materials)
    for tu in ("split", "unsplit"):
        ref, ref_dt, _ = oracle.advance(p, c.N, c.dx, lo, hi, c.U, c.cfl, 1.0, nsteps=2, unsplit=tu == "unsplit")
        out, dts = _run_steps(c, 2, lo, hi, tu)
        assert bits_equal(out, ref), (tu, mismatch_report(out, ref))
        assert bits_equal(dts, ref_dt)


def test_dt_max_clamps_step(oracle):
    c = W.water_air_tube_1d(100)
    out, dts = _run_steps(c, 3, dt_max=1e-8)
    assert np.all(dts == 1e-8)
    p = make_params(1, c.materials)
    ref, _, _ = oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, c.U, c.cfl, 1e-8, nsteps=3)
    assert bits_equal(out, ref)


def test_device_loop_and_graph_match_host_steps():
    """dimhip_advance (device-resident dt/time, with and without a hipGraph)
    reproduces run_dim_case's loop of dimhip_step calls bit for bit."""
    c = W.helium_shock_bubble_2d(128, 64)
    t_end = 1.0
    with dimhip.Solver.from_case(c) as s:
        t = 0.0
        for _ in range(9):
            t += s.step(c.cfl, t_end - t)
        host = s.download()
    for use_graph in (False, True):
        with dimhip.Solver.from_case(c) as s:
            t2, n = s.advance(9, c.cfl, t_end, 0.0, use_graph=use_graph)
            dev = s.download()
        assert n == 9
        assert t2 == t
        assert bits_equal(dev, host), (use_graph, mismatch_report(dev, host))


@pytest.mark.parametrize("every", [1, 3, 4])
def test_sampled_kernel_timing(every):
    """dimhip_set_timing(ctx, k) times the kernels of one step in k and leaves
    the state unchanged: 9 steps (after the first, whose CFL pass runs on its
    own) give ceil(8 / k) timed steps of 2 sweeps and 1 boundary pass (the
    x sweep writes the y-face ring rows the y sweep reads, so the step's one
    ring pass follows the y sweep)."""
    c = W.helium_shock_bubble_2d(128, 64)
    with dimhip.Solver.from_case(c) as s:
        t, _ = s.advance(1, c.cfl, 1.0, 0.0, use_graph=False)
        s.set_timing(True, every=every)
        t, n = s.advance(8, c.cfl, 1.0, t, use_graph=False)
        ms, counts = s.kernel_times()
        timed = s.download()
        s.set_timing(False)
    with dimhip.Solver.from_case(c) as s:
        s.advance(9, c.cfl, 1.0, 0.0, use_graph=False)
        plain = s.download()
    steps = -(-8 // every)
    assert n == 8
    assert counts == [2 * steps, steps, 0]
    assert ms[0] > 0.0 and ms[1] > 0.0
    assert bits_equal(timed, plain), mismatch_report(timed, plain)


def test_device_loop_stops_at_t_end(oracle):
    """The loop ends once time >= t_end - 1e-14 and the last dt lands on t_end."""
    c = W.water_air_tube_1d(120)
    p = make_params(1, c.materials)
    U = c.U.copy()
    t, t_end, steps = 0.0, 2.0e-6, 0
    while t < t_end - 1e-14:
        U, dts, _ = oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, U, c.cfl, t_end - t, nsteps=1)
        t += dts[0]
        steps += 1
    with dimhip.Solver.from_case(c) as s:
        t2, n = s.advance(steps + 5, c.cfl, t_end, 0.0)
        out = s.download()
    assert n == steps and t2 == t
    assert bits_equal(out, U)


@pytest.mark.parametrize("case", ["helium_2d", "collapse_3d"])
def test_advance_through_output_times_matches_oracle(oracle, case):
    """run_dim_case's loop over two output times (run_dim.hpp:88-111): each
    advance call overshoots its step count, so the steps past the output
    time run as pass-through steps on the device (an odd number of sweeps in
    3D moves the state to the other buffer); the next call continues from
    there with the graph.  Every state bit for bit against the oracle stepped
    with dt_max = t_out - t."""
    c = W.helium_shock_bubble_2d(48, 24) if case == "helium_2d" else W.bubble_collapse_3d(12)
    p = make_params(c.dim, c.materials)
    U = c.U.copy()
    t = 0.0
    outs = []
    for k, t_out in enumerate((3.0e-5, 6.5e-5) if case == "helium_2d" else (1.0e-4, 2.3e-4)):
        n = 0
        while t < t_out - 1e-14:
            U, dts, _ = oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, U, c.cfl, t_out - t, nsteps=1)
            t += dts[0]
            n += 1
        outs.append((t_out, n, t, U.copy()))
    with dimhip.Solver.from_case(c) as s:
        time = 0.0
        for t_out, n, t_ref, U_ref in outs:
            time, steps = s.advance(n + 7, c.cfl, t_out, time, use_graph=True)
            assert steps == n and time == t_ref
            out = s.download()
            assert bits_equal(out, U_ref), (t_out, mismatch_report(out, U_ref))


def test_conservation_totals():
    c = W.helium_shock_bubble_2d(300, 130)
    with dimhip.Solver.from_case(c) as s:
        s.advance(3, c.cfl, 1.0)
        tot = s.conservation()
        U = s.download()
    vol = c.dx[0] * c.dx[1]
    expect = np.concatenate([U[:, :2].sum(0) * vol, [U[:, :2].sum() * vol], U[:, 2:4].sum(0) * vol, [U[:, 4].sum() * vol]])
    scale = np.abs(expect).max()
    assert np.all(np.abs(tot - expect) <= 1e-12 * np.maximum(np.abs(expect), 1e-6 * scale))


def test_water_air_tube_full_run_matches_oracle(oracle):
    """Config 1: 1000-cell water-air tube to t=2.4e-4 s, CFL 0.8, FP64, with the
    run_dim_case loop (dt_max = t_end - time), against the oracle: <= 1e-12
    relative L-inf in every conserved field (observed: bit-identical)."""
    c = W.water_air_tube_1d(1000)
    p = make_params(1, c.materials)
    t_end = 2.4e-4
    with dimhip.Solver.from_case(c) as s:
        t_gpu, n = s.advance(100000, c.cfl, t_end, 0.0)
        out = s.download()
    U, t, steps = c.U.copy(), 0.0, 0
    while t < t_end - 1e-14:
        U, dts, _ = oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, U, c.cfl, t_end - t, nsteps=1)
        t += dts[0]
        steps += 1
    assert n == steps and t_gpu == t
    assert np.all(rel_linf(out, U) <= 1e-12)
    assert bits_equal(out, U), mismatch_report(out, U)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
def test_helium_2048x1024_step_matches_reference():
    """The helium shock-bubble at its full FP64 size, 2 steps, against the
    reference's own OpenMP solver compiled from its sources."""
    c = W.helium_shock_bubble_2d(2048, 1024)
    ref = Checker("ref")
    p = make_params(2, c.materials)
    U, dts, _ = ref.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, c.U, c.cfl, 1.0, nsteps=2, threads=0)
    out, gdts = _run_steps(c, 2)
    assert bits_equal(gdts, dts)
    assert np.all(rel_linf(out, U) <= 1e-12)
    assert bits_equal(out, U), mismatch_report(out, U)


@pytest.mark.parametrize("ms", ["water_air", "air_helium", "water_air_helium", "air", "water_na_air",
                                "water_na_air_helium", "na_air"])
def test_pressure_bisection_matches_oracle(oracle, ms):
    """The device pressure bisection equals the oracle's, bit for bit, over
    wide-range and adversarial inputs: tiny/huge/negative energies, vanishing
    fractions, floored densities, non-finite values."""
    mats = MATSETS[ms]
    nm = len(mats)
    p = make_params(1, mats)
    rng = np.random.default_rng(2024)
    n = 60000
    rt = 10.0 ** rng.uniform(-13, 4, n)
    e = 10.0 ** rng.uniform(-14, 9, n) * rng.choice([1.0, 1.0, 1.0, -1.0], n)
    alpha = rng.uniform(0, 1, (n, nm))
    alpha[rng.random(n) < 0.2, 0] = 0.0
    alpha[rng.random(n) < 0.1, -1] = 1e-13
    alpha /= alpha.sum(1, keepdims=True)
    rho = 10.0 ** rng.uniform(-13, 4, (n, nm))
    special = rng.integers(0, 40, n)
    e[special == 0] = 0.0
    e[special == 1] = np.nan
    rho[special == 2, 0] = np.inf
    alpha[special == 3, 0] = np.nan
    e[special == 4] = 1e300
    rt[special == 5] = -1.0
    packed = np.concatenate([rt[:, None], e[:, None], alpha, rho], axis=1)
    got, _ = dimhip.kat(dimhip.KAT_PRESSURE, 1, mats, n, packed)
    ref = oracle.pressure(p, rt, e, alpha, rho)
    assert bits_equal(got, ref), mismatch_report(got, ref)


@pytest.mark.parametrize("case", range(4))
def test_hllc_error_text_matches_oracle(oracle, case):
    """A face whose density sum overflows fails the device HLLC with the
    reference's message for its first failing check, and of several failing
    faces the first in order names the error (hllc.hpp:88-91); the oracle's
    messages are the reference's (tests/test_oracle.py)."""
    from test_oracle import hllc_error_cases
    name, UL, UR = hllc_error_cases()[case]
    mats = [W.AIR, W.HELIUM]
    with pytest.raises(RuntimeError) as e:
        oracle.hllc(make_params(2, mats), 0, UL, UR)
    with pytest.raises(RuntimeError) as d:
        dimhip.kat(dimhip.KAT_HLLC, 2, mats, UL.shape[0], UL.reshape(-1), UR.reshape(-1), d=0)
    assert str(d.value) == str(e.value), name


@pytest.mark.parametrize("cells", [(0,), (63,), (0, 63)])
def test_step_error_text_matches_oracle(oracle, cells):
    """A 1D step over a line whose end cells' masses overflow the density sum
    (an end cell's unpredicted state is its HLLC input): the device step fails
    with the oracle's message -- the first failing face of the line, its
    first failing check, as the reference's exception (whose what() the
    reference's own OpenMP line loop ends the program with) -- and the
    context keeps working after a fresh upload."""
    c = W.water_air_tube_1d(64)
    U = c.U.copy()
    for j in cells:
        U[j, :2] = 1.5e308
    p = make_params(1, c.materials)
    with pytest.raises(RuntimeError) as e:
        oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, U, c.cfl, 1.0)
    with dimhip.Solver.from_case(c) as s:
        s.upload(U)
        with pytest.raises(RuntimeError) as d:
            s.step(c.cfl, 1.0)
        assert str(d.value) == str(e.value)
        s.upload(c.U)
        s.step(c.cfl, 1.0)
