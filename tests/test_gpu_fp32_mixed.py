"""FP32-mixed form against the FP64 path (itself bit-identical to the
reference): normalised L1 error of rho, p and alpha, mass conservation, and
the stiffened-gas tube where FP64 interface-cell pressure recovery matters.

L1 follows the reference's normalised definition
dx*sum|a-b| / (dx*sum|b|) (tools/quant/reference_error_metrics.hpp:123-137).
"""
import numpy as np
import pytest

from oracle.pyoracle import Checker, make_params
from diffuse_vs_gfm_mms_amd import dimhip, workloads as W

pytestmark = pytest.mark.gpu


def l1(a, b):
    return float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))


def fields(oracle, case, U):
    """rho, p, alpha_0 of packed cells (oracle cons_to_prim: FP64, bisection)."""
    p = make_params(case.dim, case.materials)
    P = oracle.cons_to_prim(p, U)
    nm = case.nmat
    return U[:, :nm].sum(1), P[:, -1], P[:, nm]


def run(case, precision, nsteps, bc_lo=None, bc_hi=None):
    with dimhip.Solver(case.dim, case.N, case.dx, case.materials, bc_lo or case.bc_lo, bc_hi or case.bc_hi,
                       precision=precision) as s:
        s.upload(case.U)
        m0 = s.conservation()
        t, n = s.advance(nsteps, case.cfl, 1.0, 0.0)
        m1 = s.conservation()
        return s.download(), t, n, m0, m1


@pytest.mark.parametrize("nsteps", [20, 200])
def test_helium_bubble_l1_vs_fp64(oracle, nsteps):
    """Helium shock-bubble 512x256: rho and alpha L1 <= 1e-5 and p L1 <= 1e-5
    against the FP64 path after the same number of steps."""
    c = W.helium_shock_bubble_2d(512, 256)
    U64, t64, n64, _, _ = run(c, "fp64", nsteps)
    U32, t32, n32, _, _ = run(c, "fp32_mixed", nsteps)
    assert n64 == n32 == nsteps
    assert abs(t32 - t64) <= 1e-5 * t64
    r64, p64, a64 = fields(oracle, c, U64)
    r32, p32, a32 = fields(oracle, c, U32)
    errs = (l1(r32, r64), l1(p32, p64), l1(a32, a64))
    print(f"L1 after {nsteps} steps: rho {errs[0]:.3e} p {errs[1]:.3e} alpha {errs[2]:.3e}")
    assert errs[0] <= 1e-5 and errs[2] <= 1e-5 and errs[1] <= 1e-5, errs


def test_closed_box_mass_drift():
    """Helium bubble at 2 atm in still air, all-reflective box, 40 steps: the
    waves stay clear of the walls, so the ring boundary pass copies still air
    onto still air and the scheme's only mass change is the reference's
    inactive-mass removal (primitives.hpp:392-399).  FP64 partial masses
    updated from FP32 face fluxes telescope like the FP64 path: the
    FP32-mixed relative mass drift stays within 1e-12 of the FP64 path's."""
    rest = (1.29, (0.0, 0.0), 101325.0)
    c = W.shock_bubble((256, 128), (0.0, 0.0), (0.325, 0.1625), [W.AIR, W.HELIUM], 0, 1,
                       pre=rest, post=rest, bubble=(0.3, (0.0, 0.0), 2.0e5), centre=(0.1625, 0.08125),
                       radius=0.02, shock_axis=0, shock_position=1.0, thickness=2 * 0.325 / 256,
                       bc_lo=("reflective", "reflective"), bc_hi=("reflective", "reflective"))
    _, _, _, m0_64, m1_64 = run(c, "fp64", 40)
    U32, _, _, m0_32, m1_32 = run(c, "fp32_mixed", 40)
    M0 = m0_64[2]
    drift64 = (m1_64[2] - m0_64[2]) / M0
    drift32 = (m1_32[2] - m0_32[2]) / M0
    print(f"relative mass drift: fp64 {drift64:.3e} fp32-mixed {drift32:.3e}")
    assert abs(m0_32[2] - M0) <= 1e-15 * M0
    assert abs(drift32 - drift64) <= 1e-12
    # the walls stayed quiet: ring cells still at rest
    ring = np.concatenate([U32[:256], U32[-256:]])
    assert np.max(np.abs(ring[:, 2:4])) < 1e-6


def balance(s, c0, k):
    """The reference's boundary-flux-corrected mass balance
    (conservation_report.hpp:506-520, 596-607, 625-635): residual =
    (M(T) - M(0)) + integral of dt * boundary mass flux of each new state,
    normalised by max(|M(0)|, max |M|, 1e-12)."""
    c1 = s.conservation()
    res = (c1[k] - c0[k]) + s.boundary_flux_integral()[k]
    return abs(res) / max(abs(c0[k]), abs(c1[k]), 1e-12), c1[k] - c0[k]


def test_open_boundary_mass_balance():
    """Helium shock-bubble 512x256, 200 steps (the post-shock gas flows in
    through the right boundary): with the boundary flux integrated on the
    device, the FP64 path's mass balance closes to roundoff (<= 1e-12), as the
    reference's does.  The FP32-mixed form's residual is the FP32 rounding of
    that inflow flux (its FP32 HLLC flux against the report's FP64 physical
    flux of the same state, ~1e-7 relative): <= 5e-9, and the same order as
    the difference of the two forms' mass changes."""
    c = W.helium_shock_bubble_2d(512, 256)
    k = c.nmat
    out = {}
    T = 1.0
    for prec, n in (("fp64", 200), ("fp32_mixed", 1000)):
        with dimhip.Solver.from_case(c, precision=prec) as s:
            s.track_boundary_flux(True)
            c0 = s.conservation()
            # the FP32-mixed run ends at the FP64 run's time (its last step clamped)
            T, _ = s.advance(n, c.cfl, T, 0.0)
            out[prec] = balance(s, c0, k) + (c0[k],)
    (r64, d64, m0), (r32, d32, _) = out["fp64"], out["fp32_mixed"]
    print(f"normalised mass balance residual: fp64 {r64:.2e}, fp32-mixed {r32:.2e}; "
          f"mass change difference {(d32 - d64) / m0:.2e}")
    assert r64 <= 1e-12
    assert r32 <= 5e-9
    assert abs(d32 - d64) / m0 <= 5e-9


@pytest.mark.parametrize("gammas", [(1.4, 1.67), (2.5, 1.4), (1.4, 3.0)])
def test_ideal_gas_kernels_and_gamma_above_two(oracle, gammas):
    """Ideal-gas shock-interface instances, 60 steps, FP32-mixed against the
    FP64 path at the same time.  Contexts whose ideal gases all have gamma <=
    2 take the ideal-gas FP32 kernels, which fold min(1, 1/(gamma-1)) = 1 in
    as a constant; a gamma above 2 makes that bound 1/(gamma-1) < 1, and the
    context takes the general FP32 kernels instead.  Either way rho, p and
    alpha_0 stay within 1e-5 normalised L1 of FP64."""
    c = W.shock_interface(128, 96, gammas[0], gammas[1], 0.5, 1.8, 0.2)
    U64, t64, n64, _, _ = run(c, "fp64", 60)
    with dimhip.Solver.from_case(c, precision="fp32_mixed") as s:
        t32, n32 = s.advance(1 << 20, c.cfl, t64, 0.0)
        U32 = s.download()
    assert abs(t32 - t64) <= 1e-12 * t64
    errs = tuple(l1(a, b) for a, b in zip(fields(oracle, c, U32), fields(oracle, c, U64)))
    print(f"gammas {gammas}: L1 rho {errs[0]:.2e} p {errs[1]:.2e} alpha {errs[2]:.2e}")
    assert max(errs) <= 1e-5, errs


def test_water_air_tube_fp32_mixed(oracle):
    """Stiffened-gas water/air tube (1000 cells, 300 steps).  The FP32-mixed
    form is accurate for gas-gas mixtures; in nearly pure water expanding at
    high speed (p ~ 3e4 Pa << gamma p_inf = 2.6e9 Pa, u ~ 480 m/s) the shifted
    internal energy is mostly kinetic (1 part in 1e4 is internal), so its
    FP32 storage leaves ~1e-4 to 1e-3 relative pressure error there, even with
    the kinetic energy subtracted in FP64 (which the form does in such fast
    liquid cells as in interface cells); the interface then evolves
    measurably differently (FP64 itself moves by 4e-5 under 1e-7 input
    noise).  Measured: rho 5.0e-4, p 3.9e-4, alpha 6.4e-4 normalised L1; the
    test bounds them at 2e-3 and the FP64 form is the one for stiff liquids."""
    c = W.water_air_tube_1d(1000)
    U64, _, _, _, _ = run(c, "fp64", 300)
    U32, _, _, _, _ = run(c, "fp32_mixed", 300)
    r64, p64, a64 = fields(oracle, c, U64)
    r32, p32, a32 = fields(oracle, c, U32)
    errs = (l1(r32, r64), l1(p32, p64), l1(a32, a64))
    print(f"water-air L1: rho {errs[0]:.3e} p {errs[1]:.3e} alpha {errs[2]:.3e}")
    assert max(errs) <= 2e-3, errs


def test_fp32_mixed_3d_and_unsplit_run():
    """3D and unsplit FP32-mixed runs stay finite and close to FP64."""
    c = W.bubble_collapse_3d(24)
    for tu in ("split", "unsplit"):
        outs = []
        for prec in ("fp64", "fp32_mixed"):
            with dimhip.Solver(3, c.N, c.dx, c.materials, c.bc_lo, c.bc_hi, time_update=tu, precision=prec) as s:
                s.upload(c.U)
                s.advance(10, c.cfl, 1.0)
                outs.append(s.download())
        assert np.all(np.isfinite(outs[1]))
        assert l1(outs[1][:, :2].sum(1), outs[0][:, :2].sum(1)) < 1e-4


def test_fp32_mixed_3d_bubble_collapse_l1_vs_fp64(oracle):
    """3D water/air bubble collapse 64^3 (the 512^3 benchmark case on a
    coarser grid), 20 steps: the 1e9 Pa liquid shock runs through
    stiffened-gas water toward the bubble.  rho, p and alpha_0 normalised L1
    against the FP64 path at the same time stay below 1e-5 (measured ~2e-7 to
    4e-7).  Later, when the shock collapses the bubble, the FP32 energy
    storage of fast, low-pressure water costs ~1e-4 to 1e-3 (256^3, 100 steps,
    reported by bench.py with the FP64 path's own sensitivity to 1e-7 input
    noise beside it)."""
    c = W.bubble_collapse_3d(64)
    U64, t64, n64, _, _ = run(c, "fp64", 20)
    with dimhip.Solver(3, c.N, c.dx, c.materials, c.bc_lo, c.bc_hi, precision="fp32_mixed") as s:
        s.upload(c.U)
        t32, n32 = s.advance(100, c.cfl, t64, 0.0)
        U32 = s.download()
    assert abs(t32 - t64) <= 1e-12 * t64
    r64, p64, a64 = fields(oracle, c, U64)
    r32, p32, a32 = fields(oracle, c, U32)
    errs = (l1(r32, r64), l1(p32, p64), l1(a32, a64))
    print(f"3D L1 after {n64} FP64 / {n32} FP32-mixed steps: rho {errs[0]:.3e} p {errs[1]:.3e} alpha {errs[2]:.3e}")
    assert max(errs) <= 1e-5, errs
