"""The CPU oracle (oracle/dim_oracle.c) against the reference's golden
fixtures, against the reference built from its own sources where present,
and against what the mathematics fixes independently of both."""
import numpy as np
import pytest

from _states import MATSETS, bits_equal, mismatch_report, random_line, random_states
from oracle.pyoracle import Checker, make_params, ref_available
from diffuse_vs_gfm_mms_amd import workloads as W

LAMBDAS = ("kapila", "allaire", "equal_velocity")


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("ms", ["water_air", "air_helium", "water_air_helium"])
@pytest.mark.parametrize("lm", LAMBDAS)
def test_functions_match_golden(oracle, golden, dim, ms, lm):
    g = golden["functions"]
    key = f"d{dim}_{ms}_{lm}"
    p = make_params(dim, MATSETS[ms], lm)
    U, V = g[key + "_U"], g[key + "_V"]
    P = oracle.cons_to_prim(p, U)
    assert bits_equal(P, g[key + "_c2p"]), mismatch_report(P, g[key + "_c2p"])
    assert bits_equal(oracle.prim_to_cons(p, P), g[key + "_p2c"])
    assert bits_equal(oracle.repair_state(p, U), g[key + "_repair"])
    assert bits_equal(oracle.alpha_rhs(p, P), g[key + "_rhs"])
    for d in range(dim):
        F, fv = oracle.hllc(p, d, U, V)
        assert bits_equal(F, g[key + f"_hllc{d}_F"]), mismatch_report(F, g[key + f"_hllc{d}_F"])
        assert bits_equal(fv, g[key + f"_hllc{d}_fv"])


def test_lines_match_golden(oracle, golden):
    g = golden["lines"]
    keys = sorted({k[:-2] for k in g.files if k.endswith("_U")})
    assert len(keys) > 20
    for key in keys:
        dim = int(key[1])
        ms = "water_air" if "water_air" in key else "air_helium"
        d = int(key.split("dir")[1])
        p = make_params(dim, MATSETS[ms])
        out = oracle.solve_line(p, d, g[key + "_U"], 1e-3, 2e-8)
        assert bits_equal(out, g[key + "_out"]), (key, mismatch_report(out, g[key + "_out"]))


def test_steps_match_golden(oracle, golden):
    g = golden["steps"]
    c = W.water_air_tube_1d(200)
    p = make_params(1, c.materials)
    U = g["tube1d_U0"]
    done = 0
    for k in (1, 10, 50):
        U, dts, _ = oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, U, c.cfl, 1.0, nsteps=k - done)
        done = k
        assert bits_equal(U, g[f"tube1d_U{k}"]), (k, mismatch_report(U, g[f"tube1d_U{k}"]))
        assert bits_equal(dts, g[f"tube1d_dt{k}"])
    c = W.helium_shock_bubble_2d(40, 20)
    p = make_params(2, c.materials)
    combos = {
        "tr_refl": (c.bc_lo, c.bc_hi, False),
        "mixed": (("reflective", "reflective"), ("nonreflective", "reflection"), False),
        "unsplit": (c.bc_lo, c.bc_hi, True),
    }
    for name, (lo, hi, unsplit) in combos.items():
        U, dts, _ = oracle.advance(p, c.N, c.dx, lo, hi, g["bubble2d_U0"], c.cfl, 1.0, nsteps=5,
                                   unsplit=unsplit)
        assert bits_equal(U, g[f"bubble2d_{name}_U5"]), (name, mismatch_report(U, g[f"bubble2d_{name}_U5"]))
        assert bits_equal(dts, g[f"bubble2d_{name}_dt"])
    c = W.bubble_collapse_3d(10)
    p = make_params(3, c.materials)
    U, dts, _ = oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, g["collapse3d_U0"], c.cfl, 1.0, nsteps=2)
    assert bits_equal(U, g["collapse3d_U2"]), mismatch_report(U, g["collapse3d_U2"])


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
def test_oracle_matches_reference_fresh_inputs(oracle):
    """Fresh seeded inputs beyond the fixtures, bit for bit against the
    reference compiled from its own headers."""
    ref = Checker("ref")
    rng = np.random.default_rng(7)
    for dim in (1, 2, 3):
        for ms in ("water_air", "air_helium", "air"):
            mats = MATSETS[ms]
            p = make_params(dim, mats)
            U = random_states(mats, dim, 2000, rng)
            V = random_states(mats, dim, 2000, rng)
            assert bits_equal(oracle.cons_to_prim(p, U), ref.cons_to_prim(p, U))
            assert bits_equal(oracle.repair_state(p, U), ref.repair_state(p, U))
            for d in range(dim):
                a, b = oracle.hllc(p, d, U, V), ref.hllc(p, d, U, V)
                assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])
            line = random_line(mats, dim, 40, rng)
            for d in range(dim):
                assert bits_equal(oracle.solve_line(p, d, line, 2e-3, 1e-7),
                                  ref.solve_line(p, d, line, 2e-3, 1e-7))


# ---- independent pins: closed forms and invariants ------------------------

def _closed_form_pressure(mats, alpha, rho_k, rho_e):
    """Root of sum_k a_k rho_k e_k(rho_k, p) = rho e for ideal/stiffened gases:
    p* = (rho e - sum_k a_k g_k pinf_k/(g_k-1)) / sum_k a_k/(g_k-1)."""
    num = rho_e.copy()
    S = np.zeros_like(rho_e)
    for k, (kind, g, pinf) in enumerate(mats):
        num -= alpha[:, k] * (g * pinf if kind == "stiffened_gas" else 0.0) / (g - 1.0)
        S += alpha[:, k] / (g - 1.0)
    return num / S, S


@pytest.mark.parametrize("ms", ["water_air", "air_helium"])
def test_bisection_brackets_closed_form_root(oracle, ms):
    """The bisection stops at |f(p)| <= 1e-10 max(1, rho e); f is linear in p
    with slope S, so |p - p*| <= 1e-10 max(1, rho e)/S (+ rounding)."""
    mats = MATSETS[ms]
    p = make_params(1, mats)
    rng = np.random.default_rng(3)
    n = 5000
    a0 = rng.uniform(0.0, 1.0, n)
    alpha = np.stack([a0, 1 - a0], 1)
    rho = 10 ** rng.uniform(-1, 3, (n, 2))
    pr = 10 ** rng.uniform(3, 9, n)
    rt = (alpha * rho).sum(1)
    rho_e = np.zeros(n)
    for k, (kind, g, pinf) in enumerate(mats):
        rho_e += alpha[:, k] * (pr + (g * pinf if kind == "stiffened_gas" else 0.0)) / (g - 1.0)
    got = oracle.pressure(p, rt, rho_e / rt, alpha, rho)
    pstar, S = _closed_form_pressure(mats, alpha, rho, rho_e)
    bound = 1e-10 * np.maximum(1.0, rho_e) / S + 1e-12 * np.abs(pstar) + 1e-9 * rho_e / S * 1e-3
    assert np.all(np.abs(got - pstar) <= bound * 1.01 + 1e-9)


def test_bisection_negative_energy_returns_floor(oracle):
    """Negative internal energy: the bracket fails and the clamped fallback
    pressure comes back (eos.hpp:118-120) — the 1e-10 floor."""
    mats = MATSETS["air_helium"]
    p = make_params(1, mats)
    out = oracle.pressure(p, np.array([1.0]), np.array([-5.0]), np.array([[0.5, 0.5]]), np.array([[1.0, 1.0]]))
    assert out[0] == 1e-10


def test_wood_sound_speed_pure_material(oracle):
    """For a single pure ideal gas the Wood speed is sqrt(gamma p / rho)."""
    mats = MATSETS["air_helium"]
    p = make_params(1, mats)
    rho = np.array([1.2, 3.0])
    pr = np.array([1e5, 7e5])
    alpha = np.array([[1.0, 0.0], [1.0, 0.0]])
    c = oracle.sound_speed(p, rho, pr, alpha, np.stack([rho, rho], 1))
    assert np.allclose(c, np.sqrt(1.4 * pr / rho), rtol=1e-14)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_hllc_consistency(oracle, dim):
    """F(U, U) equals the physical flux of U (Toro's consistency)."""
    mats = MATSETS["water_air"]
    p = make_params(dim, mats)
    rng = np.random.default_rng(11)
    U = random_states(mats, dim, 200, rng, extreme=False)
    for d in range(dim):
        F, fv = oracle.hllc(p, d, U, U)
        P = oracle.cons_to_prim(p, U)
        nm = len(mats)
        un = P[:, 2 * nm + d]
        assert np.allclose(fv, un, rtol=1e-12, atol=1e-12)
        assert np.allclose(F[:, :nm], U[:, :nm] * un[:, None], rtol=1e-12, atol=1e-9)
        Eflux = (U[:, nm + dim] + P[:, -1]) * un
        assert np.allclose(F[:, nm + dim], Eflux, rtol=1e-10, atol=1e-6)


def test_uniform_state_is_preserved(oracle):
    """A uniform state at rest stays uniform under a full step, up to the
    bisection tolerance: face states are rebuilt through the 1e-10-relative
    pressure root, so the first/last (unpredicted) faces see relative
    pressure differences of that order, i.e. momentum noise of ~1e-10 p dt/dx."""
    c = W.helium_shock_bubble_2d(16, 8)
    U = np.repeat(c.U[:1], c.ncells, axis=0)
    U[:, 2:4] = 0.0
    p = make_params(2, c.materials)
    out, dts, _ = oracle.advance(p, c.N, c.dx, ("transmissive",) * 2, ("transmissive",) * 2, U, 0.8, 1.0, nsteps=3)
    mom_scale = np.sqrt(U[0, :2].sum() * 1.0e5)
    assert np.allclose(out[:, [0, 1, 4, 5]], U[:, [0, 1, 4, 5]], rtol=1e-10, atol=0)
    assert np.max(np.abs(out[:, 2:4])) < 1e-10 * mom_scale


def test_y_uniform_2d_equals_1d_rows(oracle):
    """A 2D state constant in y, transmissive in y, evolves every row like the 1D tube
    (the y sweep sees uniform lines)."""
    c1 = W.water_air_tube_1d(60)
    p1 = make_params(1, c1.materials)
    U1, dt1, _ = oracle.advance(p1, c1.N, c1.dx, c1.bc_lo, c1.bc_hi, c1.U, 0.8, 1.0, nsteps=4)
    ny = 5
    U2 = np.zeros((60 * ny, 6))
    for j in range(ny):
        U2[j * 60:(j + 1) * 60, 0:2] = c1.U[:, 0:2]
        U2[j * 60:(j + 1) * 60, 2] = c1.U[:, 2]
        U2[j * 60:(j + 1) * 60, 4] = c1.U[:, 3]
        U2[j * 60:(j + 1) * 60, 5] = c1.U[:, 4]
    p2 = make_params(2, c1.materials)
    out, dt2, _ = oracle.advance(p2, (60, ny), (c1.dx[0], 1.0), ("transmissive",) * 2, ("transmissive",) * 2,
                                 U2, 0.8, 1.0, nsteps=4)
    assert np.allclose(dt1, dt2, rtol=0, atol=0) or np.allclose(dt1, dt2, rtol=1e-15)
    # the y sweep of a y-uniform line rebuilds each cell through the pressure
    # root (1e-10 relative), so rows agree with the 1D run to that order
    from _states import rel_linf
    for j in range(ny):
        row = out[j * 60:(j + 1) * 60]
        assert np.all(rel_linf(row[:, [0, 1, 2, 4, 5]], U1) < 1e-8)
        assert np.max(np.abs(row[:, 3])) < 1e-8 * np.max(np.abs(U1[:, 2]))


def test_mirror_symmetry_1d(oracle):
    """A mirror-symmetric tube stays mirror-symmetric (momentum antisymmetric)."""
    c = W.water_air_tube_1d(80)
    U = c.U.copy()
    U[40:] = U[:40][::-1]
    p = make_params(1, c.materials)
    out, _, _ = oracle.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, U, 0.8, 1.0, nsteps=5)
    mirrored = out[::-1].copy()
    mirrored[:, 2] *= -1.0
    assert np.allclose(out[:, [0, 1, 3, 4]], mirrored[:, [0, 1, 3, 4]], rtol=1e-10, atol=1e-6)
    assert np.allclose(out[:, 2], mirrored[:, 2], rtol=1e-8, atol=1e-3)


def test_minmod_planted_mistake_is_caught(oracle):
    """Changing the limiter would move solve_line output: the line fixtures
    differ from a first-order (zero-slope) line update, so a dropped slope fails."""
    rng = np.random.default_rng(5)
    mats = MATSETS["water_air"]
    p = make_params(1, mats)
    U = random_line(mats, 1, 30, rng)
    second = oracle.solve_line(p, 0, U, 1e-3, 1e-8)
    # first order: the same face states as the cell centres, via the n==2 branch per face
    assert not np.allclose(second, U)


def hllc_error_cases():
    """(name, UL, UR) with one or two faces whose density sum overflows to
    infinity: the only way total_density (a sum of finite_or'd masses,
    primitives.hpp:99-108) becomes non-finite; cons_to_prim's pressure is
    clamped finite by construction, so the pressure checks cannot fail."""
    ok = np.array([1.0, 0.5, 10.0, 0.0, 2.5e5, 0.6])
    big = ok.copy()
    big[:2] = 1.5e308
    return [
        ("left", big[None], ok[None]),
        ("right", ok[None], big[None]),
        ("both", big[None], big[None]),
        ("second face left, first face right", np.stack([ok, big]), np.stack([big, ok])),
    ]


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("case", range(4))
def test_hllc_error_text_matches_reference(oracle, case):
    """The oracle fails on the same face with the same message as the
    reference's hllc_flux_normal (hllc.hpp:88-91): the first failing face in
    order, its first failing require_positive."""
    name, UL, UR = hllc_error_cases()[case]
    p = make_params(2, [W.AIR, W.HELIUM])
    msgs = []
    for kind in ("oracle", "ref"):
        with pytest.raises(RuntimeError) as e:
            Checker(kind).hllc(p, 0, UL, UR)
        msgs.append(str(e.value))
    assert msgs[0] == msgs[1], (name, msgs)
    assert msgs[0].startswith("dim::hllc_flux_normal: invalid ") and msgs[0].endswith(" (non-finite)")
