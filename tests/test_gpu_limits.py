"""Host-side failure checks of the library (needs an MI355X for the device).

* ES_MAX_DEVICE_BYTES caps what one solver or batch allocates on the device;
  an allocation beyond it is ES_ERR_OOM with a message, and the solver stays
  usable for work that fits (the reference prints and exit(0)s on errors,
  ErrorMacros.h:22-38).
* The mixed-precision trace hands the items it cannot decide to the FP64
  trace as int32 item numbers of their launch segment: a segment of 2^31 or
  more items is refused (ES_ERR_INVALID_ARG) before anything is allocated.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import announce

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def es():
    import elasticscattering_amd as es_mod

    es_mod.load_library()
    return es_mod


def test_device_byte_cap_is_oom(es, monkeypatch):
    announce("ES_MAX_DEVICE_BYTES = 2 MB")
    monkeypatch.setenv("ES_MAX_DEVICE_BYTES", str(2 << 20))
    x, y = O.read_points(os.path.join(GOLDEN, "test_impurities.dat"))
    rng = np.random.default_rng(4)
    bx, by = rng.uniform(-5e-7, 1.5e-6, 60000), rng.uniform(-5e-7, 1.5e-6, 60000)
    with es.Simulation(31, 7) as sim:
        with pytest.raises(es.EsError) as ei:
            sim.set_impurities(bx, by, 1e-6, 5e-7, 1.1e-8, 15, 128)  # 9 N slots of bins: > 2 MB
        assert ei.value.code == es.ES_ERR_OOM
        assert "cap" in str(ei.value)
    with es.Simulation(31, 7) as sim:  # a small field fits under the same cap and runs
        sim.set_impurities(x, y, 1e-7, 1e-7, 2e-9, 10)
        sim.init_sample(1.0, 0.3, 1, 1.67834e5, False)
        sig, _, _ = sim.compute_sigmas(24.006, [1.0])
        assert sig[0][0] > 0.0


def test_mixed_refuses_a_segment_beyond_int32(es, monkeypatch):
    announce("mixed precision, one segment of 2^31+ items")
    monkeypatch.setenv("ES_SEGMENT_ITEMS", str(10 ** 10))
    with es.Batch() as b:
        b.set_precision("mixed")
        # 400 disks over a 2e-6 m side: nearly all of 512^2 positions are free;
        # 512^2 x 4 x 65 x 40 fields = 2.7e9 items in one segment
        b.add(seed=5, region_size=1e-6, region_extends=5e-7, impurity_radius=5e-9, positions_per_row=512,
              particles_per_quadrant=65, fields=list(np.logspace(0, 1.6, 40)), temperatures=[10.0], tau=1e-12,
              alpha=np.pi / 4, is_clockwise=1, particle_speed=7e5, coherent=False, count=400, cells_per_row=8)
        for call in (b.counts, b.run):  # both plan the batch; neither allocates the per-item buffers
            with pytest.raises(es.EsError) as ei:
                call()
            assert ei.value.code == es.ES_ERR_INVALID_ARG and "2^31" in str(ei.value)
