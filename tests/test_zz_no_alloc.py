"""es_grid_bin_noalloc against es_grid_set_impurities + es_grid_read (needs an MI355X).

The no-allocation entry point bins device-resident centres into caller
buffers with a fixed scratch array in the library.  Its CSR must equal the
allocating path's bit for bit: two calls in a row (the second smaller, so the
scratch holds the first call's contents when it starts), a cell grid from the
reference rule and one with disks wider than a cell; inputs beyond the
scratch size are refused.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import announce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def es():
    import elasticscattering_amd as es_mod

    es_mod.load_library()
    return es_mod


def _cpr(n, target, cpr):
    return cpr if cpr > 0 else max(int(round(math.sqrt(n / target))), 1)


def _bin_noalloc(es, x, y, rs, ext, r, target, cpr):
    from hip_mem import Hip

    hip = Hip()
    try:
        n = len(x)
        cells = _cpr(n, target, cpr) ** 2
        xd, yd = hip.upload(np.asarray(x, dtype=np.float64)), hip.upload(np.asarray(y, dtype=np.float64))
        ordered = hip.fill(max(9 * n, 1) * 16, np.nan, np.float64)
        index = hip.fill(cells * 4, -7, np.int32)
        total = hip.fill(8, -1, np.int64)
        d = es.GridDesc(rs, ext, r, 0.0, int(target), int(cpr))
        info = es.GridInfo()
        lib = es.load_library()
        rc = lib.es_grid_bin_noalloc(xd, yd, n, C.byref(d), ordered, index, total, C.byref(info), None)
        if rc != es.ES_OK:
            raise es.EsError(rc, lib.es_last_error(None).decode())
        hip.sync()
        t = int(hip.download(total, 1, np.int64)[0])
        assert info.cells_per_row * info.cells_per_row == cells and info.indexed_impurity_count == -1
        xy = hip.download(ordered, 2 * t, np.float64).reshape(-1, 2)
        return xy, hip.download(index, cells, np.int32), t, info
    finally:
        hip.close()


def _bin_allocating(es, x, y, rs, ext, r, target, cpr):
    with es.Simulation(31, 7) as sim:
        info = sim.set_impurities(x, y, rs, ext, r, target, cpr)
        xy, idx = sim.read_grid(info)
    return xy, idx, info


def _check(es, x, y, rs, ext, r, target, cpr):
    xy_n, idx_n, total_n, info_n = _bin_noalloc(es, x, y, rs, ext, r, target, cpr)
    xy_a, idx_a, info_a = _bin_allocating(es, x, y, rs, ext, r, target, cpr)
    assert info_n.cells_per_row == info_a.cells_per_row
    assert info_n.cell_size == info_a.cell_size
    assert total_n == info_a.indexed_impurity_count
    np.testing.assert_array_equal(idx_n, idx_a)
    assert np.array_equal(xy_n.view(np.uint64), xy_a.view(np.uint64))


def test_noalloc_binning_equals_allocating_path(es, oracle_esm):
    from elasticscattering_amd.workloads import PRODUCTION_FP64 as w, region

    rs, ext = region(w)
    n = w["impurity_count"]
    x = np.empty(n)
    y = np.empty(n)
    oracle_esm.generate_impurities(w["seeds"][0], n, -ext, rs + ext, O.dptr(x), O.dptr(y))
    announce("noalloc binning: production field, 60,000 disks, 128^2 cells")
    _check(es, x, y, rs, ext, w["impurity_radius"], 15, w["cells_per_row"])
    # second call, smaller: the reference's file field with the reference cell rule
    fx, fy = O.read_points(os.path.join(O.ROOT, "tests", "golden", "test_impurities.dat"))
    announce("noalloc binning: test_impurities.dat, target 10")
    _check(es, fx, fy, 1e-7, 1e-7, 2e-9, 10, 0)
    announce("noalloc binning: disks wider than a cell")
    _check(es, fx, fy, 1e-7, 1e-7, 4e-8, 0, 7)


def test_noalloc_refuses_inputs_beyond_its_scratch(es):
    from hip_mem import Hip

    lib = es.load_library()
    hip = Hip()
    try:
        big = es.ES_NOALLOC_MAX_IMPURITIES + 1
        xd = hip.alloc(big * 8)
        out, idx, tot = hip.alloc(16), hip.alloc(4), hip.alloc(8)
        d = es.GridDesc(1e-6, 5e-7, 1e-8, 0.0, 15, 4)
        rc = lib.es_grid_bin_noalloc(xd, xd, big, C.byref(d), out, idx, tot, None, None)
        assert rc == es.ES_ERR_INVALID_ARG and b"scratch" in lib.es_last_error(None)
        d = es.GridDesc(1e-6, 5e-7, 1e-8, 0.0, 15, 257)  # 257^2 cells > ES_NOALLOC_MAX_CELLS
        rc = lib.es_grid_bin_noalloc(xd, xd, 100, C.byref(d), out, idx, tot, None, None)
        assert rc == es.ES_ERR_INVALID_ARG and b"cells" in lib.es_last_error(None)
    finally:
        hip.close()
