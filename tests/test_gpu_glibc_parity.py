"""HIP FP64 path vs the oracle built on glibc's libm (needs an MI355X).

The kernels call es_math.h's correctly rounded sin, cos, asin and exp; the
reference's CPU path calls the C runtime's libm (details.h:100, lifetime.h:43,
util.h:34-46).  On the fields and settings of oracle_lib.agreement_case, over
B in {0.01, 0.1, 1, 20, 40} T and both passes:
* GPU lifetimes equal the correctly rounded oracle build bit for bit;
* against the glibc build, sigma agrees within 1e-8 (relative) and every
  lifetime within 1e-9 but for at most one per (B, pass); the CPU tests show
  those come from glibc's asin, which misrounds ~0.15 % of arguments;
* the device's elementary functions equal the host's (the same source) bit for
  bit, including arguments whose fast phase fails the rounding test.
"""
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


def _gpu_setup(es, case):
    from elasticscattering_amd import workloads as W

    if case == "file1":
        import os

        x, y = O.read_points(os.path.join(O.ROOT, "tests", "golden", "test_impurities.dat"))
        sim = es.Simulation(31, 7)
        sim.set_impurities(x, y, 1e-7, 1e-7, 2e-9, 10)
        return sim, dict(tau=1.0, alpha=0.3, is_clockwise=1, particle_speed=1.67834e5), 0.0
    w = W.SMALL_FP64 if case == "small" else W.PRODUCTION_FP64
    rs, ext = W.region(w)
    sim = es.Simulation(64, 17 if case == "small" else 65)
    sim.generate_impurities(w["seeds"][0], rs, ext, w["impurity_radius"], count=w["impurity_count"],
                            cells_per_row=w["cells_per_row"])
    itau = 1e-12 if case == "small" else 0.0
    return sim, dict(tau=1e-12, alpha=w["alpha"], is_clockwise=1, particle_speed=7e5), itau


@pytest.mark.parametrize("case", ["file1", "small", "production"])
def test_fp64_path_agrees_with_glibc_oracle(es, case):
    L, E = O.load("libm"), O.load("esm")
    gl, pl, temps = O.agreement_case(L, case)
    ge, pe, _ = O.agreement_case(E, case)
    sim, st, itau = _gpu_setup(es, case)
    flips = 0
    with sim:
        for coherent in (1, 0):
            sim.init_sample(st["tau"], st["alpha"], st["is_clockwise"], st["particle_speed"], coherent, itau)
            for B in O.AGREEMENT_FIELDS:
                announce(f"glibc parity {case} coherent={coherent} B={B}")
                sig_g, m_g, lt_g = sim.compute_sigmas(B, temps, want_lifetimes=True)
                sig_l, m_l, lt_l = O.compute_sigmas(L, gl, pl, coherent, B, temps, want_lifetimes=True)
                sig_e, m_e, lt_e = O.compute_sigmas(E, ge, pe, coherent, B, temps, want_lifetimes=True)
                bad = np.flatnonzero(lt_g.view(np.uint64) != lt_e.view(np.uint64))
                assert bad.size == 0, (case, coherent, B, bad[:5])
                assert m_g["cells_passed"] == m_e.pm.cells_passed
                assert m_g["impurities_tested"] == m_e.pm.impurities_tested
                rel = np.abs(lt_g - lt_l) / np.maximum(np.abs(lt_l), 1e-300)
                n_far = int(np.count_nonzero(rel > 1e-9))
                assert n_far <= 1, (case, coherent, B, n_far, float(rel.max()))
                flips += n_far
                for (gx, gy), so in zip(sig_g, sig_l):
                    for g, o in ((gx, so.xx_comp), (gy, so.xy_comp)):
                        if o != 0.0 or g != 0.0:
                            assert abs(g - o) <= 1e-8 * abs(o), (case, coherent, B, g, o)
    print(f"{case}: lifetimes off by more than 1e-9 from the glibc oracle: {flips}")
    assert flips <= 1


def _math_args():
    rng = np.random.default_rng(5)
    n = 1 << 18
    near_one = 1.0 - np.ldexp(rng.random(n), -rng.integers(0, 52, n))
    tiny = np.ldexp(1.0 + rng.random(n), -rng.integers(0, 40, n)) * rng.choice([-1.0, 1.0], n)
    return {
        "asin": np.concatenate([rng.uniform(-1, 1, n), near_one, -near_one, tiny,
                                [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 1.0000000000000002, np.nan]]),
        "sin": np.concatenate([rng.uniform(-16, 16, n), tiny, rng.uniform(-1e6, 1e6, n // 4),
                               np.arange(-64, 65) * (np.pi / 4), [0.0, -0.0, np.inf, np.nan]]),
        "cos": np.concatenate([rng.uniform(-16, 16, n), tiny, np.arange(-64, 65) * (np.pi / 4), [0.0, np.nan]]),
        "exp": np.concatenate([rng.uniform(-20, 0, n), rng.uniform(-745.2, 709.79, n), tiny,
                               rng.uniform(-745.2, -707.0, n // 4), [0.0, -0.0, -746.0, 710.0, np.nan]]),
    }


@pytest.mark.parametrize("fn", ["sin", "cos", "asin", "exp"])
def test_device_math_equals_host_math(es, fn):
    announce(f"es_math_eval {fn}")
    E = O.load("esm")
    f = getattr(E, "m_" + fn)
    x = _math_args()[fn]
    y = es.math_eval(fn, x)
    want = np.array([f(float(v)) for v in x])
    same = (y.view(np.uint64) == want.view(np.uint64)) | (np.isnan(y) & np.isnan(want))
    assert same.all(), (fn, x[~same][:5], y[~same][:5], want[~same][:5])
