"""The batched form's ring boundary treatment fused into its sweeps
(copy-kind boundaries: transmissive, reflection, reflective;
boundary.hpp:165-226): every ring cell receives its source's new state from
the thread that computes the source, with the normal momentum of each
reflective side negated, and the y sweep closes each instance's step.
Instances with small extents (3 cells: one interior cell feeds both ends of
its axis), corners where a source feeds 3 ring cells, several tiles per
line, and tangential velocities that the flips act on, against the same
instances advanced alone with the separate boundary pass, bit for bit.
The single grid's separate ring pass for copy-kind boundaries (one stored
field per thread), in 3D the sweeps' x-face ring cells taken from their
x-neighbour's lane, and in 2D the y-face ring rows written by the x sweep
(no ring pass between the two sweeps), against the oracle, bit for bit."""
import numpy as np
import pytest

from _states import bits_equal, mismatch_report
from oracle.pyoracle import make_params
from diffuse_vs_gfm_mms_amd import dimhip, workloads as W

pytestmark = pytest.mark.gpu

SHAPES = [
    ((3, 7), ("reflective", "reflection"), ("transmissive", "reflective")),
    ((9, 3), ("reflection", "reflective"), ("reflective", "reflective")),
    ((130, 9), ("reflective", "transmissive"), ("reflection", "reflective")),
    ((4, 70), ("transmissive", "reflective"), ("reflective", "reflection")),
    ((61, 66), ("reflective", "reflective"), ("reflective", "reflective")),
]


def _case(N, lo, hi, mats):
    """A shock-bubble with velocity along both axes, so that each reflective
    face flips a non-zero momentum."""
    light = mats[1] != W.AIR
    vel = (350.0, -40.0) if not light else (120.0, -30.0)
    pre = (1200.0, vel, 4.0e8) if not light else (2.0, vel, 2.0e5)
    post = (1000.0, (0.0, 0.0), 1.0e5) if not light else (1.2, (0.0, 0.0), 1.0e5)
    return W.shock_bubble(N, (0.0, 0.0), (1.0, 1.0), list(mats), 0, 1, pre=pre, post=post,
                          bubble=(1.0 if not light else 0.2, (-vel[0], -vel[1]), 1.0e5), centre=(0.55, 0.5),
                          radius=0.3, shock_axis=0, shock_position=0.35, thickness=0.5 / max(N),
                          bc_lo=lo, bc_hi=hi, name="ring_case")


def _alone(case, nsteps):
    with dimhip.Solver.from_case(case, precision="fp32_mixed") as s:
        t, n = s.advance(nsteps, case.cfl, 1e300, 0.0, use_graph=False)
        return s.download(), t, n


@pytest.mark.parametrize("mats", ["water_air", "air_helium"])
@pytest.mark.parametrize("cfg", [(0, 0, 2, 1), (1, 1, 0, 4), (0, 2, 1, 2)])
def test_fused_ring_batch_equals_separate_passes(mats, cfg):
    pair = {"water_air": (W.WATER, W.AIR), "air_helium": (W.AIR, W.HELIUM)}[mats]
    cases = [_case(N, lo, hi, pair) for N, lo, hi in SHAPES]
    refs = [_alone(c, 5) for c in cases]
    with dimhip.Batch(cases) as b:
        b.configure(*cfg)
        times, steps = b.advance(5, 0.8, use_graph=False)
        for i, c in enumerate(cases):
            out = b.download(i)
            assert steps[i] == 5, (mats, cfg, i)
            assert times[i] == refs[i][1], (mats, cfg, i)
            assert bits_equal(out, refs[i][0]), (mats, cfg, c.N, mismatch_report(out, refs[i][0]))


def test_batch_with_a_nonreflective_instance_keeps_boundary_passes():
    """One nonreflective instance turns the fused treatment off for the whole
    batch (its ring needs two interior cells after the sweep)."""
    pair = (W.AIR, W.HELIUM)
    cases = [_case(N, lo, hi, pair) for N, lo, hi in SHAPES[2:]]
    cases.append(_case((40, 24), ("nonreflective", "reflective"), ("transmissive", "nonreflective"), pair))
    refs = [_alone(c, 5) for c in cases]
    with dimhip.Batch(cases) as b:
        b.configure(0, 0, 2, 1)
        times, steps = b.advance(5, 0.8, use_graph=False)
        for i in range(len(cases)):
            out = b.download(i)
            assert times[i] == refs[i][1], i
            assert bits_equal(out, refs[i][0]), (i, mismatch_report(out, refs[i][0]))


# single-grid FP64 steps with copy-kind boundaries (the field-parallel ring
# pass, with the recovered-pressure copies), extents down to 3, 1D to 3D
GRID_CASES = {
    "1d_3": ((3,), ("reflective",), ("reflection",)),
    "1d_4": ((4,), ("transmissive",), ("reflective",)),
    "2d_3x7": ((3, 7), ("reflective", "reflection"), ("transmissive", "reflective")),
    "2d_130x9": ((130, 9), ("reflective", "transmissive"), ("reflection", "reflective")),
    "2d_61x66": ((61, 66), ("reflection", "reflective"), ("reflective", "transmissive")),
    "2d_300x5": ((300, 5), ("transmissive", "reflective"), ("reflective", "reflection")),
    "3d_3x5x4": ((3, 5, 4), ("reflective", "reflection", "transmissive"),
                 ("transmissive", "reflective", "reflective")),
    "3d_7x3x6": ((7, 3, 6), ("reflection", "reflective", "reflective"), ("reflective", "transmissive", "reflection")),
}


def _grid_case(N, lo, hi):
    dim = len(N)
    vel = (350.0, -40.0, 25.0)[:dim]
    return W.shock_bubble(N, (0.0,) * dim, (1.0,) * dim, [W.WATER, W.AIR], 0, 1,
                          pre=(1200.0, vel, 4.0e8), post=(1000.0, (0.0,) * dim, 1.0e5),
                          bubble=(1.0, tuple(-v for v in vel), 1.0e5), centre=(0.55,) * dim, radius=0.3,
                          shock_axis=0, shock_position=0.35, thickness=0.5 / max(N), bc_lo=lo, bc_hi=hi,
                          name="ring_grid_case")


@pytest.mark.parametrize("name", sorted(GRID_CASES))
def test_copy_ring_pass_fp64_matches_oracle(oracle, name):
    N, lo, hi = GRID_CASES[name]
    c = _grid_case(N, lo, hi)
    p = make_params(c.dim, c.materials)
    ref, ref_dt, _ = oracle.advance(p, c.N, c.dx, lo, hi, c.U, c.cfl, 1.0, nsteps=4)
    with dimhip.Solver(c.dim, c.N, c.dx, c.materials, lo, hi) as s:
        s.upload(c.U)
        dts = [s.step(c.cfl, 1.0) for _ in range(4)]
        out = s.download()
    assert bits_equal(out, ref), (name, mismatch_report(out, ref))
    assert bits_equal(np.array(dts), ref_dt), name


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp64"])
@pytest.mark.parametrize("name", ["2d_3x7", "2d_130x9", "2d_61x66", "2d_300x5", "3d_3x5x4", "3d_7x3x6"])
def test_fused_ring_steps_match_separate_passes(name, precision):
    """Steps whose sweeps carry part of the ring treatment (3D: x-face ring
    cells from their x-neighbour's lane; 2D: the y-face ring rows written by
    the x sweep, with no ring pass before the y sweep) against the same steps
    taken as one sweep and one full ring pass at a time (dimhip_compute_dt,
    dimhip_sweep, dimhip_apply_bc), bit for bit."""
    N, lo, hi = GRID_CASES[name]
    c = _grid_case(N, lo, hi)
    with dimhip.Solver(c.dim, c.N, c.dx, c.materials, lo, hi, precision=precision) as a, \
            dimhip.Solver(c.dim, c.N, c.dx, c.materials, lo, hi, precision=precision) as b:
        a.upload(c.U)
        b.upload(c.U)
        for k in range(3):
            dt_a = a.step(c.cfl, 1.0)
            dt_b = b.compute_dt(c.cfl, 1.0)
            assert dt_a == dt_b, (name, k, dt_a, dt_b)
            for d in range(c.dim):
                b.sweep(d, dt_b)
                b.apply_bc()
            ua, ub = a.download(), b.download()
            assert bits_equal(ua, ub), (name, k, mismatch_report(ua, ub))
