"""HIP path vs the CPU oracle, through the C ABI (needs an MI355X).

Bars: binning and lifetimes bit-exact (the oracle built on es_math.h, the math
the kernels call); metrics equal; sigma within 1e-12 relative of the oracle's
compensated sum and 1e-8 relative of its reference-order serial sum, plus an
absolute floor of 1e-12 * |sigma_xx| for sigma_xy (it can pass through zero);
the reference's own golden DataRows within the reference tolerance
(1e-5 / B relative, tests/TestMacros.h:17).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import announce

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SIGMA_REL_COMP = 1e-12
SIGMA_REL_NAIVE = 1e-8


@pytest.fixture(scope="module")
def es():
    import elasticscattering_amd as es_mod

    es_mod.load_library()
    return es_mod


def _file_points(name):
    return O.read_points(os.path.join(GOLDEN, name))


def _small_case():
    from elasticscattering_amd.workloads import SMALL_FP64 as w, region

    rs, ext = region(w)
    return w, rs, ext


def _gpu_grid_matches(sim, info, og):
    xy_g, idx_g = sim.read_grid(info)
    xy_o, idx_o = og.csr()
    assert info.cells_per_row == og.cells_per_row
    assert info.indexed_impurity_count == og.total_indexed
    np.testing.assert_array_equal(idx_g, idx_o)
    assert np.array_equal(xy_g.view(np.uint64), xy_o.view(np.uint64))


@pytest.mark.parametrize("name,rs,ext,r,target,cpr", [
    ("test_impurities.dat", 1e-7, 1e-7, 2e-9, 10, 0),
    ("real_case_impurities.dat", 1e-7, 1e-7, 1.11e-9, 10, 0),
    ("test_impurities_lt.dat", 1e-6, 1e-6, 1e-8, 10, 0),
    ("test_impurities.dat", 1e-7, 1e-7, 4e-8, 0, 7),   # radius larger than a cell: wide probes
])
def test_binning_file_fields_bit_exact(es, oracle_esm, name, rs, ext, r, target, cpr):
    announce(f"binning {name} r={r} target={target} cpr={cpr}")
    x, y = _file_points(name)
    og = O.OracleGrid.from_points(oracle_esm, x, y, rs, ext, r, target, cpr)
    with es.Simulation(31, 7) as sim:
        info = sim.set_impurities(x, y, rs, ext, r, target, cpr)
        _gpu_grid_matches(sim, info, og)


# cell grids: 1, 32^2, 128^2 (one-block scan), 181^2 = 32,761 (its largest grid, 128 KB of
# LDS), 200^2 (two-level scan) and the reference rule (cpr 0)
@pytest.mark.parametrize("seed,n,r,cpr", [(101, 2000, 1.1e-8, 32), (7, 60000, 1.1e-8, 128), (3, 400, 2e-8, 0),
                                          (11, 5000, 5e-9, 1), (13, 120000, 6e-9, 181), (17, 90000, 5e-9, 200)])
def test_generated_fields_bit_exact(es, oracle_esm, seed, n, r, cpr):
    announce(f"generate+bin seed={seed} n={n} r={r} cpr={cpr}")
    rs, ext = 1e-6, 5e-7
    og = O.OracleGrid.generate(oracle_esm, seed, rs, ext, 0.0, r, 15, cpr, n=n)
    with es.Simulation(31, 7) as sim:
        info = sim.generate_impurities(seed, rs, ext, r, count=n, target_cell_population=15, cells_per_row=cpr)
        _gpu_grid_matches(sim, info, og)


def _run_pair(es, oracle_esm, x, y, rs, ext, r, target, cpr, ppr, ppq, tau, alpha, cw, speed, coherent, fields,
              temps, incoherent_tau=0.0):
    og = O.OracleGrid.from_points(oracle_esm, x, y, rs, ext, r, target, cpr)
    prm = O.sim_params(ppr, ppq, rs, ext, og.cell_size, tau, alpha, cw, speed, incoherent_tau, nthreads=8)
    with es.Simulation(ppr, ppq) as sim:
        sim.set_impurities(x, y, rs, ext, r, target, cpr)
        sim.init_sample(tau, alpha, cw, speed, coherent, incoherent_tau)
        for B in fields:
            announce(f"lifetimes ppr={ppr} ppq={ppq} cw={cw} coherent={coherent} B={B} nT={len(temps)}")
            sig_g, m_g, lt_g = sim.compute_sigmas(B, temps, want_lifetimes=True)
            sig_o, m_o, lt_o = O.compute_sigmas(oracle_esm, og, prm, coherent, B, temps, want_lifetimes=True)
            # lifetimes: bit-exact
            bad = np.flatnonzero(lt_g.view(np.uint64) != lt_o.view(np.uint64))
            assert bad.size == 0, f"{bad.size} lifetimes differ, first {bad[:5]}: {lt_g[bad[:5]]} vs {lt_o[bad[:5]]}"
            # metrics
            for k in ("cells_passed", "impurities_tested", "particles_inside_impurity", "particles_escaped",
                      "particles_at_bound", "max_impurities_tested", "max_cells_passed", "trace_cap_hits",
                      "edge_hits", "cross_calls"):
                assert m_g[k] == getattr(m_o.pm, k), k
            assert m_g["trace_cap_hits"] == 0
            assert m_g["real_particles"] == m_o.real_particles
            assert m_g["traced_particles"] == m_o.real_particles
            taus = [tau] * len(temps) if coherent else [incoherent_tau if incoherent_tau > 0 else
                                                        1.05457148e-34 / (1.380649e-23 * T) for T in temps]
            assert m_g["sigma_evals"] == sum(int(np.count_nonzero(np.minimum(lt_o, 15.0 * tt))) for tt in taus)
            assert abs(m_g["avg_particle_lifetime"] - m_o.avg_particle_lifetime) <= 1e-12 * abs(m_o.avg_particle_lifetime)
            # sigma
            for (gxx, gxy), so in zip(sig_g, sig_o):
                atol = 1e-12 * abs(so.xx_comp)
                assert abs(gxx - so.xx_comp) <= SIGMA_REL_COMP * abs(so.xx_comp) + 1e-300
                assert abs(gxy - so.xy_comp) <= SIGMA_REL_COMP * abs(so.xy_comp) + atol
                assert abs(gxx - so.xx_naive) <= SIGMA_REL_NAIVE * abs(so.xx_naive) + 1e-300
                assert abs(gxy - so.xy_naive) <= SIGMA_REL_NAIVE * abs(so.xy_naive) + atol


@pytest.mark.parametrize("cw", [1, 0])
@pytest.mark.parametrize("coherent", [0, 1])
def test_file1_lifetimes_bit_exact(es, oracle_esm, cw, coherent):
    x, y = _file_points("test_impurities.dat")
    _run_pair(es, oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10, 0, 31, 7, 1.0, 0.3, cw, 1.67834e5, coherent,
              [0.01, 12.008, 60.0], [0.01, 0.1, 1, 10, 100])


@pytest.mark.parametrize("cw", [1, 0])
def test_zero_width_coherent_window_bit_exact(es, oracle_esm, cw):
    """alpha = pi/4: the coherent window has zero width (Simulation.h:115-124), so the
    ppq phi points of a quadrant are one particle; the GPU traces it once and every
    lifetime, metric and the (zero) sigma must still equal the oracle's."""
    x, y = _file_points("test_impurities.dat")
    _run_pair(es, oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10, 0, 31, 7, 1.0, math.pi / 4, cw, 1.67834e5, 1,
              [0.01, 12.008, 60.0], [0.01, 1, 100])
    w, rs, ext = _small_case()
    xs = np.empty(w["impurity_count"])
    ys = np.empty(w["impurity_count"])
    oracle_esm.generate_impurities(w["seeds"][0], w["impurity_count"], -ext, rs + ext, O.dptr(xs), O.dptr(ys))
    _run_pair(es, oracle_esm, xs, ys, rs, ext, w["impurity_radius"], 15, w["cells_per_row"], 24, 9, w["tau"],
              math.pi / 4, cw, w["particle_speed"], 1, [1.0, 20.0], [2.0, 40.0])


def test_small_fp64_workload_bit_exact(es, oracle_esm):
    w, rs, ext = _small_case()
    x = np.empty(w["impurity_count"])
    y = np.empty(w["impurity_count"])
    oracle_esm.generate_impurities(w["seeds"][0], w["impurity_count"], -ext, rs + ext, O.dptr(x), O.dptr(y))
    _run_pair(es, oracle_esm, x, y, rs, ext, w["impurity_radius"], 15, w["cells_per_row"], w["positions_per_row"],
              w["particles_per_quadrant"], w["tau"], w["alpha"], w["clockwise"], w["particle_speed"], 0,
              w["fields"], w["temperatures"], w["incoherent_tau"])


def test_sparse_long_traces_bit_exact(es, oracle_esm):
    """Sparse field, low B: traces cross many cells and often leave the grid."""
    rs, ext = 1e-6, 5e-7
    x = np.empty(300)
    y = np.empty(300)
    oracle_esm.generate_impurities(5, 300, -ext, rs + ext, O.dptr(x), O.dptr(y))
    _run_pair(es, oracle_esm, x, y, rs, ext, 2e-8, 3, 0, 24, 9, 1e-12, 0.3, 1, 7e5, 0, [1.0, 5.0, 40.0],
              [2.0, 40.0, 100.0])


@pytest.mark.parametrize("cw", [1, 0])
def test_dense_field_screened_trace_bit_exact(es, oracle_esm, cw):
    """Dense field (about 20 indexed disks per cell): the solver runs the screened
    cooperative trace, which skips crossing disks that cannot lower the minimum;
    lifetimes stay bit-exact and fewer exact solves than crossing disks run."""
    rs, ext = 1e-6, 5e-7
    n = 40000
    x = np.empty(n)
    y = np.empty(n)
    oracle_esm.generate_impurities(23, n, -ext, rs + ext, O.dptr(x), O.dptr(y))
    _run_pair(es, oracle_esm, x, y, rs, ext, 8e-9, 15, 64, 24, 9, 1e-12, 0.3, cw, 7e5, 0, [0.5, 6.0, 40.0],
              [2.0, 40.0])
    with es.Simulation(24, 9) as sim:
        sim.set_impurities(x, y, rs, ext, 8e-9, 15, 64)
        sim.init_sample(1e-12, 0.3, cw, 7e5, 0)
        _, m, _ = sim.compute_sigmas(6.0, [2.0])
    assert 0 < m["cross_exact"] < m["cross_calls"], m


def test_sweep_equals_single_calls(es):
    w, rs, ext = _small_case()
    fields = [1.0, 7.5, 20.0, 40.0]
    temps = [2.0, 40.0, 100.0]
    announce("sweep vs single calls")
    with es.Simulation(w["positions_per_row"], w["particles_per_quadrant"]) as sim:
        sim.generate_impurities(101, rs, ext, w["impurity_radius"], count=w["impurity_count"],
                                cells_per_row=w["cells_per_row"])
        sim.init_sample(w["tau"], w["alpha"], 1, w["particle_speed"], False)
        sig, ms = sim.compute_sweep(fields, temps)
        for b, B in enumerate(fields):
            single, m, _ = sim.compute_sigmas(B, temps)
            for t in range(len(temps)):
                # nblk differs between one-field and many-field launches only if
                # the work item count differs; it does not, so results are identical.
                assert sig[b, t, 0] == single[t][0] and sig[b, t, 1] == single[t][1]
            assert ms[b]["cells_passed"] == m["cells_passed"]


def test_reference_golden_datarows(es):
    """tests/sim/SigmaIntegrationTest.h:80-141 on the GPU, at the reference tolerance."""
    fx = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))
    for key, rs, ext, r, ppr in (("datarows_file1", 1e-7, 1e-7, 2e-9, 31), ("datarows_file2", 1e-7, 1e-7, 1.11e-9, 127)):
        block = fx[key]
        x, y = _file_points(block["impurities"])
        temps = sorted({row["T"] for row in block["rows"]})
        fields = sorted({row["B"] for row in block["rows"]})
        announce(f"golden {key} ppr={ppr}")
        res = {}
        with es.Simulation(ppr, 7) as sim:
            sim.set_impurities(x, y, rs, ext, r, 10)
            for coherent in (1, 0):
                sim.init_sample(1.0, 0.3, 1, 1.67834e5, coherent)
                for B in fields:
                    sig, m, _ = sim.compute_sigmas(B, temps)
                    assert m["particles_escaped"] == 0
                    for t, T in enumerate(temps):
                        res[(coherent, B, T)] = sig[t]
        for row in block["rows"]:
            B, T = row["B"], row["T"]
            tol = 1e-5 / B
            for coherent, kx, ky in ((1, "coh_xx", "coh_xy"), (0, "inc_xx", "inc_xy")):
                gx, gy = res[(coherent, B, T)]
                assert abs((gx - row[kx]) / row[kx]) < tol, (key, B, T, kx, gx, row[kx])
                assert abs((gy - row[ky]) / row[ky]) < tol, (key, B, T, ky, gy, row[ky])


def test_start_outside_grid_is_an_error(es):
    announce("out-of-grid start position")
    x, y = _file_points("test_impurities.dat")
    with es.Simulation(31, 7) as sim:
        sim.set_impurities(x, y, 1e-7, 0.0, 2e-9, 10)  # no extends: the last row starts beyond the grid
        sim.init_sample(1.0, 0.3, 1, 1.67834e5, False)
        with pytest.raises(es.EsError) as ei:
            sim.compute_sigmas(1.0, [1.0])
        assert ei.value.code == es.ES_ERR_OUT_OF_GRID


def test_device_math_bit_exact_through_lifetimes(es, oracle_esm):
    """Near-axis phi grids: phi = 0 exactly (alpha = pi/4, odd ppq) and pi/2 multiples."""
    x, y = _file_points("test_impurities_lt.dat")
    _run_pair(es, oracle_esm, x, y, 1e-6, 1e-6, 1e-8, 10, 0, 17, 17, 1e-12, math.pi / 4, 1, 1.68e5, 0,
              [8.0, 10.0, 50.0], [1.0, 10.0])
