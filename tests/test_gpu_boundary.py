"""The drop-in boundary for the reference's Grid (needs an MI355X).

A Simulation backend receives the Grid as the reference builds it
(Simulation::ComputeSigmas(B, T, const Grid&, ...), Simulation.h:128):
GetImpurities() — cell-ordered centres, a disk listed once per cell it
overlaps — and GetIndex() — inclusive end offsets (Grid.h:74-75,
Grid.cpp:92-110).  es_grid_set_binned takes that CSR as it is; fed the
oracle's restatement of Grid.cpp (duplicates included), every lifetime is
bit-exact, every metric (indexed count, impurities_tested, max counters)
equals the oracle's, and sigma agrees as in the other parity tests.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import announce

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
METRIC_KEYS = ("cells_passed", "impurities_tested", "particles_inside_impurity", "particles_escaped",
               "particles_at_bound", "max_impurities_tested", "max_cells_passed", "edge_hits", "cross_calls")


@pytest.fixture(scope="module")
def es():
    import elasticscattering_amd as es_mod

    es_mod.load_library()
    return es_mod


def _check_binned(es, lib, og, rs, ext, ppr, ppq, tau, alpha, cw, speed, fields, temps, itau=0.0):
    xy, idx = og.csr()
    assert og.total_indexed > og.unique_impurity_count  # the CSR carries duplicates
    prm = O.sim_params(ppr, ppq, rs, ext, og.cell_size, tau, alpha, cw, speed, itau, nthreads=8)
    gi = es.GridInfo(rs, ext, og.cell_size, og.total_indexed, og.unique_impurity_count,
                     og.cells_per_row ** 2, og.cells_per_row)
    with es.Simulation(ppr, ppq, grid_info=gi) as sim:
        info = sim.set_binned(xy, idx, og.cells_per_row, og.spawn_lo, og.spawn_hi - og.spawn_lo, og.radius)
        assert info.indexed_impurity_count == og.total_indexed
        assert info.cell_size == og.cell_size
        for coherent in (0, 1):
            sim.init_sample(tau, alpha, cw, speed, coherent, itau)
            for B in fields:
                announce(f"binned grid ppr={ppr} coherent={coherent} B={B}")
                sg, mg, lg = sim.compute_sigmas(B, temps, want_lifetimes=True)
                so, mo, lo = O.compute_sigmas(lib, og, prm, coherent, B, temps, want_lifetimes=True)
                bad = np.flatnonzero(lg.view(np.uint64) != lo.view(np.uint64))
                assert bad.size == 0, (coherent, B, bad[:5])
                for k in METRIC_KEYS:
                    assert mg[k] == getattr(mo.pm, k), (k, mg[k], getattr(mo.pm, k))
                for (gx, gy), s in zip(sg, so):
                    assert abs(gx - s.xx_comp) <= 1e-12 * abs(s.xx_comp) + 1e-300
                    assert abs(gy - s.xy_comp) <= 1e-12 * abs(s.xy_comp) + 1e-12 * abs(s.xx_comp)


def test_reference_grid_csr_file_field(es, oracle_esm):
    """tests/data/test_impurities.dat with the reference cell rule (target 10)."""
    x, y = O.read_points(os.path.join(GOLDEN, "test_impurities.dat"))
    og = O.OracleGrid.from_points(oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10)
    og.radius = 2e-9
    _check_binned(es, oracle_esm, og, 1e-7, 1e-7, 31, 7, 1.0, 0.3, 1, 1.67834e5, [0.01, 12.008, 60.0], [1.0, 10.0])


def test_reference_grid_csr_production_field(es, oracle_esm):
    """The production field (60,000 disks, 128^2 cells) at 32^2 positions, both senses."""
    from elasticscattering_amd.workloads import PRODUCTION_FP64 as w, region

    rs, ext = region(w)
    og = O.OracleGrid.generate(oracle_esm, w["seeds"][0], rs, ext, 0.0, w["impurity_radius"], 15,
                               w["cells_per_row"], n=w["impurity_count"])
    og.radius = w["impurity_radius"]
    for cw in (1, 0):
        _check_binned(es, oracle_esm, og, rs, ext, 32, 17, 1e-12, w["alpha"], cw, 7e5, [1.0, 40.0], [2.0, 100.0])


def test_binned_grid_equals_binning_on_the_gpu(es, oracle_esm):
    """The same field given binned or as centres: the same sigma bits and metrics."""
    x, y = O.read_points(os.path.join(GOLDEN, "test_impurities.dat"))
    og = O.OracleGrid.from_points(oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10)
    xy, idx = og.csr()
    gi = es.GridInfo(1e-7, 1e-7, og.cell_size, og.total_indexed, og.unique_impurity_count,
                     og.cells_per_row ** 2, og.cells_per_row)
    out = []
    for binned in (True, False):
        with es.Simulation(31, 7, grid_info=gi) as sim:
            if binned:
                sim.set_binned(xy, idx, og.cells_per_row, og.spawn_lo, og.spawn_hi - og.spawn_lo, 2e-9)
            else:
                sim.set_impurities(x, y, 1e-7, 1e-7, 2e-9, 10)
            sim.init_sample(1.0, 0.3, 1, 1.67834e5, False)
            out.append(sim.compute_sigmas(24.006, [1.0, 10.0], want_lifetimes=True))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][2].view(np.uint64), out[1][2].view(np.uint64))
    assert {k: out[0][1][k] for k in METRIC_KEYS} == {k: out[1][1][k] for k in METRIC_KEYS}


def test_binned_grid_rejects_a_bad_index(es):
    gi = es.GridInfo(1e-7, 1e-7, 1e-8, 0, 0, 4, 2)
    with es.Simulation(5, 3, grid_info=gi) as sim:
        with pytest.raises(es.EsError) as ei:
            sim.set_binned(np.zeros((3, 2)), np.array([2, 1, 3, 3], dtype=np.int32), 2, -1e-7, 3e-7, 1e-9)
        assert ei.value.code == es.ES_ERR_INVALID_ARG
    with es.Simulation(5, 3) as sim:  # no GridInformation: the start positions are unknown
        with pytest.raises(es.EsError):
            sim.set_binned(np.zeros((3, 2)), np.array([1, 2, 3, 3], dtype=np.int32), 2, -1e-7, 3e-7, 1e-9)
