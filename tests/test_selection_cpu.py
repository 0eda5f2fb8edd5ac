"""The GPU trace's choice of which crossing points need their angle, checked
on the CPU against the reference operations (no GPU).

The screened cooperative trace (es_kernels.hip, es_trace.h
d_first_cross_time_sel) orders the two crossing points of a solved disk by
their travel pseudo-angles from the start and compares them with the cell's
range (cell_range_pseudo), on margins of 1e-9 (1e-6 against the FP32 range
values), and evaluates the angle (asin) only of the points that can give the
disk's time.  The cell step (d_next_cell_lazy) picks the exit among the
cell-edge crossings by the same pseudo-angles from the entry point, and
decides whether the next cell is available (the arc from its entry to the
cell's exit holds the start angle) from pseudo-angles too; the exit angle is
evaluated only when a margin is not met.  oracle/liboracle_esm_sel.so runs
the same decisions beside the reference operations (ORACLE_CHECK_SELECTION):
over fields from sparse (long traces, many cells, near-whole turns) to the
production density, both rotation senses, both passes, low and high B, no
solve may get a different time, no point may be decided into (or out of) the
range that AngleInRange (details.h:119-136) puts out of (into) it, no decided
exit may differ from GetNextCell's (cell_grid.h:132-181) and no decided
availability from the reference's.  The GPU side is the bit-exact lifetime
tests (test_gpu_parity.py, test_gpu_configs.py, test_gpu_random_fields.py).

The check computes a pseudo-angle's quotient by IEEE division; the GPU takes
it from v_rcp_f64 and one Newton step (relative error below 1e-14).  Every
decision compares pseudo-angles on margins of at least 1e-9, so the two can
differ only for values within ~1e-14 of a margin, where both choices are
still on the safe side of the reference's rounding errors (~1e-13).
"""
import ctypes as C
import math
import os

import pytest

import oracle_lib as O


@pytest.fixture(scope="module")
def sel():
    lib = O.load("esm_sel")
    if lib is None:
        pytest.fail("oracle/liboracle_esm_sel.so is missing: run `make -C oracle`")
    lib.lib.orc_selection_stats16.argtypes = [C.POINTER(C.c_int64)]
    return lib


def _run(lib, og, ppr, ppq, rs, ext, tau, alpha, cw, speed, coherent, B):
    lib.lib.orc_selection_reset()
    prm = O.sim_params(ppr, ppq, rs, ext, og.cell_size, tau, alpha, cw, speed, nthreads=4)
    O.compute_sigmas(lib, og, prm, coherent, B, [1.0])
    a = (C.c_int64 * 16)()
    lib.lib.orc_selection_stats16(a)
    return dict(zip(("solves", "evaluated", "time_differs", "wrongly_in", "wrongly_out", "solve_undecided",
                     "exit_undecided", "exit_wrong", "avail_undecided", "avail_wrong", "range_undecided", "cells"),
                    list(a)))


def _golden(lib):
    return O.read_points(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_impurities.dat"))


CASES = [
    # name, grid factory, ppr, ppq, rs, ext, tau, alpha, speed, fields
    ("file", lambda L: O.OracleGrid.from_points(L, *_golden(L), 1e-7, 1e-7, 2e-9, 10), 31, 7, 1e-7, 1e-7, 1.0, 0.3,
     1.67834e5, [0.01, 12.008, 60.0]),
    ("production", lambda L: O.OracleGrid.generate(L, 101, 1e-6, 5e-7, 0.0, 1.1e-8, 15, 128, n=60000), 20, 9, 1e-6,
     5e-7, 1e-12, math.pi / 4, 7e5, [1.0, 40.0]),
    ("sparse", lambda L: O.OracleGrid.generate(L, 5, 1e-6, 5e-7, 0.0, 2e-8, 3, 0, n=300), 20, 9, 1e-6, 5e-7, 1e-12,
     0.3, 7e5, [0.1, 5.0, 40.0]),
    ("mid", lambda L: O.OracleGrid.generate(L, 11, 5e-7, 2.5e-7, 0.0, 1.1e-8, 15, 32, n=2000), 24, 9, 5e-7, 2.5e-7,
     1e-12, 0.5, 7e5, [0.5, 60.0]),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_selection_never_changes_a_solve(sel, case):
    name, make, ppr, ppq, rs, ext, tau, alpha, speed, fields = case
    og = make(sel)
    tot = {}
    for cw in (0, 1):
        for coherent in (0, 1):
            for B in fields:
                st = _run(sel, og, ppr, ppq, rs, ext, tau, alpha, cw, speed, coherent, B)
                wrong = ("time_differs", "wrongly_in", "wrongly_out", "exit_wrong", "avail_wrong")
                assert all(st[k] == 0 for k in wrong), (cw, coherent, B, st)
                for k, v in st.items():
                    tot[k] = tot.get(k, 0) + v
    solves, cells = tot["solves"], tot["cells"]
    assert solves > 1000 and cells > 1000
    # the point of the selection: about one angle per solve instead of two,
    # and no angle of a cell's exit point in all but a few cells
    print(f"{name}: {solves} solves, {tot['evaluated'] / solves:.3f} angles evaluated per solve (the reference: 2); "
          f"{cells} cells, undecided: exit {tot['exit_undecided'] / cells:.2e}, availability "
          f"{tot['avail_undecided'] / cells:.2e}, range {tot['range_undecided'] / cells:.2e}, "
          f"solves {tot['solve_undecided'] / solves:.2e}")
    assert tot["evaluated"] < 1.2 * solves
    assert tot["exit_undecided"] + tot["avail_undecided"] + tot["range_undecided"] < 0.05 * cells
