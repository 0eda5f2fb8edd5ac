"""Edge cases of the FP64 path against the oracle (needs an MI355X): an empty
field, a single impurity, more temperatures than one temperature-stage launch
holds (8), the smallest grids, and alpha = pi/4 (empty coherent window,
Simulation.h:115-117).  Same bars as test_gpu_parity.py."""
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


def _check(es, oracle, x, y, rs, ext, r, target, cpr, ppr, ppq, tau, alpha, cw, speed, coherent, fields, temps,
           incoherent_tau=0.0):
    og = O.OracleGrid.from_points(oracle, x, y, rs, ext, r, target, cpr)
    prm = O.sim_params(ppr, ppq, rs, ext, og.cell_size, tau, alpha, cw, speed, incoherent_tau, nthreads=8)
    out = []
    with es.Simulation(ppr, ppq) as sim:
        sim.set_impurities(x, y, rs, ext, r, target, cpr)
        sim.init_sample(tau, alpha, cw, speed, coherent, incoherent_tau)
        for B in fields:
            sig_g, m_g, lt_g = sim.compute_sigmas(B, temps, want_lifetimes=True)
            sig_o, m_o, lt_o = O.compute_sigmas(oracle, og, prm, coherent, B, temps, want_lifetimes=True)
            assert np.array_equal(lt_g.view(np.uint64), lt_o.view(np.uint64))
            assert m_g["cells_passed"] == m_o.pm.cells_passed
            assert m_g["particles_escaped"] == m_o.pm.particles_escaped
            for (gxx, gxy), so in zip(sig_g, sig_o):
                atol = 1e-12 * abs(so.xx_comp) + 1e-300
                assert abs(gxx - so.xx_comp) <= 1e-12 * abs(so.xx_comp) + 1e-300
                assert abs(gxy - so.xy_comp) <= 1e-12 * abs(so.xy_comp) + atol
            out.append((sig_g, m_g, lt_g))
    return out


def test_empty_field(es, oracle_esm):
    announce("empty field")
    res = _check(es, oracle_esm, np.empty(0), np.empty(0), 1e-6, 5e-7, 1e-8, 0, 8, 9, 5, 1e-12, 0.3, 1, 7e5, 0,
                 [1.0, 20.0], [10.0])
    for sig, m, lt in res:
        assert not lt.any() and m["particles_escaped"] == m["real_particles"] == 9 * 9 * 20
        assert all(s == (0.0, 0.0) for s in sig)


def test_single_impurity(es, oracle_esm):
    announce("single impurity")
    _check(es, oracle_esm, np.array([4.1e-7]), np.array([5.3e-7]), 1e-6, 5e-7, 6e-8, 0, 8, 17, 7, 1e-12, 0.3, 1,
           7e5, 0, [0.5, 3.0, 30.0], [2.0, 50.0])


def test_more_than_eight_temperatures(es, oracle_esm):
    announce("11 temperatures (two temperature-stage chunks)")
    x, y = O.read_points(os.path.join(os.path.dirname(__file__), "golden", "test_impurities.dat"))
    temps = [0.01, 0.1, 0.5, 1, 2, 5, 10, 20, 50, 100, 200]
    _check(es, oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10, 0, 15, 5, 1.0, 0.3, 1, 1.67834e5, 0, [5.0], temps)


def test_smallest_grids(es, oracle_esm):
    announce("2x2 positions, 2 phi points per quadrant")
    x, y = O.read_points(os.path.join(os.path.dirname(__file__), "golden", "test_impurities.dat"))
    _check(es, oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10, 0, 2, 2, 1.0, 0.3, 0, 1.67834e5, 1, [1.0, 12.0], [1.0])


def test_alpha_quarter_pi_coherent_is_zero(es, oracle_esm):
    announce("alpha = pi/4: empty coherent window")
    x, y = O.read_points(os.path.join(os.path.dirname(__file__), "golden", "test_impurities.dat"))
    res = _check(es, oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10, 0, 11, 5, 1.0, math.pi / 4, 1, 1.67834e5, 1, [3.0],
                 [1.0, 10.0])
    for sig, _, _ in res:
        assert all(abs(a) == 0.0 and abs(b) == 0.0 for a, b in sig)
