"""dimhip_output_rows_noalloc: the CSV row values of dimhip_output_rows with
its staging taken from the library's fixed per-GPU scratch array instead of
device allocations.  Called through the C ABI on two contexts in a row, the
second smaller than the first, each time against dimhip_output_rows bit for
bit; a grid above the scratch's 65536 cells fails with a message."""
import ctypes as C

import numpy as np
import pytest

from diffuse_vs_gfm_mms_amd import dimhip, workloads as W

pytestmark = pytest.mark.gpu


def rows_noalloc(s):
    out = np.full((s.ncells, 3 + s.dim + 3 * s.nmat), np.nan)
    rc = s.lib.dimhip_output_rows_noalloc(s.h, out.ctypes.data_as(C.POINTER(C.c_double)), s.ncells)
    return rc, out


@pytest.mark.parametrize("grid, precision, steps", [((256, 128), "fp64", 3), ((128, 64), "fp32_mixed", 5)])
def test_noalloc_rows_equal_allocating_rows(grid, precision, steps):
    c = W.helium_shock_bubble_2d(*grid)
    with dimhip.Solver.from_case(c, precision=precision) as s:
        s.advance(steps, c.cfl, 1.0, 0.0, use_graph=False)
        ref = s.output_rows()
        rc, out = rows_noalloc(s)
        assert rc == 0, s.lib.dimhip_last_error(s.h).decode()
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
        # a second call on the same state gives the same rows
        rc, again = rows_noalloc(s)
        assert rc == 0 and np.array_equal(again.view(np.uint64), ref.view(np.uint64))


def test_noalloc_rows_in_a_row_second_smaller():
    """Two contexts one after the other, the second with fewer cells: the
    scratch keeps the first call's contents, and the second call's rows are
    still exactly dimhip_output_rows's."""
    big = W.helium_shock_bubble_2d(256, 200)   # 51200 cells
    small = W.water_air_tube_1d(300)           # 300 cells, another layout
    for c, prec in ((big, "fp64"), (small, "fp32_mixed")):
        with dimhip.Solver.from_case(c, precision=prec) as s:
            s.advance(2, c.cfl, 1.0, 0.0, use_graph=False)
            ref = s.output_rows()
            rc, out = rows_noalloc(s)
            assert rc == 0, s.lib.dimhip_last_error(s.h).decode()
            assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), c.name


def test_noalloc_rows_reject_grids_above_the_scratch():
    c = W.helium_shock_bubble_2d(512, 256)  # 131072 cells
    with dimhip.Solver.from_case(c) as s:
        rc, _ = rows_noalloc(s)
        assert rc != 0
        msg = s.lib.dimhip_last_error(s.h).decode()
        assert "131072 cells exceed the 65536" in msg, msg
        # the context still works
        assert np.isfinite(s.output_rows()).all()
