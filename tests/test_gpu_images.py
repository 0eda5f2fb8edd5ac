"""Images mode on the GPU (ComputeSigmasWithImages, SimulationCPU.cpp:159-171,
224-243; LogImages, Logger.cpp:278-334), through the C ABI: the per-position
images bit-exact against the oracle's restatement on the same lifetimes, the
sigmas equal to ComputeSigmas', and the runner's PNG files (output_type all)
decoding to the pixels of the oracle's images."""
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import announce
from test_images_cpu import expected_pixels, read_png_rgb

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def es():
    import elasticscattering_amd as es_mod

    es_mod.load_library()
    return es_mod


def _same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("cw,coherent", [(1, 0), (0, 1), (1, 1), (0, 0)])
def test_images_bit_exact_file_field(es, oracle_esm, cw, coherent):
    announce(f"images test_impurities.dat cw={cw} coherent={coherent}")
    x, y = O.read_points(os.path.join(GOLDEN, "test_impurities.dat"))
    ppr, ppq, tau, alpha, speed = 31, 7, 1.0, 0.3, 1.67834e5
    temps = [0.01, 1.0, 100.0]
    og = O.OracleGrid.from_points(oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10)
    prm = O.sim_params(ppr, ppq, 1e-7, 1e-7, og.cell_size, tau, alpha, cw, speed, 0.0, nthreads=8)
    with es.Simulation(ppr, ppq) as sim:
        sim.set_impurities(x, y, 1e-7, 1e-7, 2e-9, 10)
        sim.init_sample(tau, alpha, cw, speed, coherent)
        for B in (0.5, 12.008):
            sig, m, img = sim.compute_sigmas_with_images(B, temps)
            sig2, _, lts = sim.compute_sigmas(B, temps, want_lifetimes=True)
            assert sig == sig2
            for t, T in enumerate(temps):
                ref = O.derive_images(oracle_esm, og, prm, coherent, B, lts, T, ppr)
                assert _same_bits(img[t], ref), (B, T, np.abs(img[t] - ref).max())


def test_images_bit_exact_generated_field(es, oracle_esm):
    from elasticscattering_amd.workloads import SMALL_FP64 as w, region

    rs, ext = region(w)
    announce("images small workload")
    ppr, ppq = w["positions_per_row"], w["particles_per_quadrant"]
    og = O.OracleGrid.generate(oracle_esm, 101, rs, ext, 0.0, w["impurity_radius"], 15, w["cells_per_row"],
                               n=w["impurity_count"])
    prm = O.sim_params(ppr, ppq, rs, ext, og.cell_size, w["tau"], w["alpha"], 1, w["particle_speed"], 2e-13,
                       nthreads=8)
    with es.Simulation(ppr, ppq) as sim:
        sim.generate_impurities(101, rs, ext, w["impurity_radius"], count=w["impurity_count"],
                                cells_per_row=w["cells_per_row"])
        sim.init_sample(w["tau"], w["alpha"], 1, w["particle_speed"], False, 2e-13)
        sig, m, img = sim.compute_sigmas_with_images(20.0, [0.0])
        _, _, lts = sim.compute_sigmas(20.0, [0.0], want_lifetimes=True)
    ref = O.derive_images(oracle_esm, og, prm, 0, 20.0, lts, 0.0, ppr)
    assert _same_bits(img[0], ref)


CONFIG = """output_directory {out}
output_type all
num_samples 1
dimension 17
phi_steps 5
temperatures 2 40
magnetic_field_min 5
magnetic_field_max 40
magnetic_field_n 2
particle_speed 7e5
tau 1e-12
alpha 0.3
clockwise 1
region_size 1e-6
region_extends 5e-7
impurity_density 1e15
impurity_radius 1.1e-8
start_seed 13
"""


def test_es_run_writes_images(tmp_path, es, oracle_esm):
    cfg_path = tmp_path / "run.config"
    cfg_path.write_text(CONFIG.format(out=str(tmp_path / "out")))
    cfg = es.parse_config(str(cfg_path))
    announce("es_run output_type all")
    out_dir = es.run(str(cfg_path))
    seed = cfg["sample_seeds"][0]
    rs, ext = cfg["region_size"], cfg["region_extends"]
    ppr, ppq = cfg["positions_per_row"], cfg["particles_per_quadrant"]
    og = O.OracleGrid.generate(oracle_esm, seed, rs, ext, cfg["impurity_density"], cfg["impurity_radius"],
                               cfg["target_cell_population"])
    prm = O.sim_params(ppr, ppq, rs, ext, og.cell_size, cfg["tau"], cfg["alpha"], cfg["is_clockwise"],
                       cfg["particle_speed"], nthreads=8)
    n = 0
    for coherent, sub in ((1, "Coherent"), (0, "Incoherent")):
        for b, B in enumerate(cfg["magnetic_fields"]):
            _, _, lts = O.compute_sigmas(oracle_esm, og, prm, coherent, B, cfg["temperatures"], want_lifetimes=True)
            for t, T in enumerate(cfg["temperatures"]):
                path = os.path.join(out_dir, "Sample 0", sub, f"T{t} MF{b}.png")
                assert os.path.exists(path), path
                img = O.derive_images(oracle_esm, og, prm, coherent, B, lts, T, ppr)
                assert np.array_equal(read_png_rgb(path), expected_pixels(*img)), path
                n += 1
    assert n == 2 * len(cfg["magnetic_fields"]) * len(cfg["temperatures"])
