"""Batched form: seeded uneven instances advanced together equal the same
instances advanced one by one, bit for bit, for every tile shape and
schedule; the autotuner restores the state it times on."""
import numpy as np
import pytest

from _states import bits_equal, mismatch_report
from diffuse_vs_gfm_mms_amd import dimhip, workloads as W

pytestmark = pytest.mark.gpu


def small_batch():
    return W.batch_instances(count=6, seed=7, nx_range=(40, 130), ny_range=(24, 70))


def alone(case, nsteps):
    with dimhip.Solver.from_case(case, precision="fp32_mixed") as s:
        t, n = s.advance(nsteps, case.cfl, 1e300, 0.0)
        return s.download(), t, n


@pytest.mark.parametrize("cfg", [(0, 0, 1, 4), (1, 1, 0, 2), (0, 2, 0, 8), (1, 0, 1, 1), (0, 1, 2, 1), (1, 2, 2, 1)])
def test_batch_equals_single_instances(cfg):
    cases = small_batch()
    refs = [alone(c, 7) for c in cases]
    with dimhip.Batch(cases) as b:
        b.configure(*cfg)
        times, steps = b.advance(7, 0.8)
        for i, c in enumerate(cases):
            out = b.download(i)
            assert steps[i] == 7
            assert times[i] == refs[i][1]
            assert bits_equal(out, refs[i][0]), (i, mismatch_report(out, refs[i][0]))


def test_batch_graph_and_autotune_keep_results():
    cases = small_batch()
    refs = [alone(c, 9) for c in cases]
    with dimhip.Batch(cases) as b:
        report = b.autotune(0.8, trial_steps=2)
        # 2 x tiles x 3 y tiles x (2 persistent schedules x 3 block counts + one-tile blocks)
        assert "best" in report and len(report["trials"]) == 42
        times, steps = b.advance(9, 0.8, use_graph=True)
        for i in range(len(cases)):
            assert bits_equal(b.download(i), refs[i][0]), i
            assert times[i] == refs[i][1]


def test_batch_instances_differ():
    cases = W.batch_instances(count=4)
    shapes = {c.N for c in cases}
    assert len(shapes) == 4
    for c in cases:
        assert 256 <= c.N[0] <= 1024 and 256 <= c.N[1] <= 512
        assert np.all(np.isfinite(c.U))
