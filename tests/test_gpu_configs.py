"""The benchmark workloads themselves against the CPU oracle (needs an MI355X).

* Production (60,000 disks, 128^2 cells, 65 phi points per window, 40 fields
  log-spaced over 1-40 T, T in {2, 40, 100} K): every bench seed's field at
  64^2 positions, both passes, every field, through the production launch
  shape (the screened cooperative trace, chosen for fields with >= 10
  indexed disks per cell): lifetimes bit-exact, every metric equal, sigma
  within 1e-12 of the oracle; the 8-sample batch (one launch for all samples,
  passes and fields, as the bench runs it) gives the solver's sigma bits.
* Batched uneven (64 samples, 400-60,000 disks, radii 5-20 nm, 32^2-512^2
  positions): a stratified subset — both density ends, both radius extremes,
  every position class, at their full position grids, fields 1, 6.3, 40 T —
  FP64 bit-exact against the oracle; the mixed path: every lifetime t within
  1e-5 t + 5e-7 / omega, sigma within 1e-6 (measured <= 1.2e-8); and the batch
  of them (one launch) equal to the solver bit for bit.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from conftest import announce

pytestmark = pytest.mark.gpu

METRIC_KEYS = ("cells_passed", "impurities_tested", "particles_inside_impurity", "particles_escaped",
               "particles_at_bound", "max_impurities_tested", "max_cells_passed", "edge_hits", "cross_calls")
E_CHARGE, M_EFF = 1.60217646e-19, 5.0 * 9.10938188e-31
# The mixed path's travel-angle error on this workload: up to 3.9e-7 rad at 40 T
# (FP32 in the start frame of orbits of ~0.5 um).
MIXED_ANGLE_FLOOR = 5e-7
THREADS = 16


@pytest.fixture(scope="module")
def es():
    import elasticscattering_amd as es_mod

    es_mod.load_library()
    return es_mod


def _check_fp64(sg, mg, lg, so, mo, lo, tag):
    bad = np.flatnonzero(lg.view(np.uint64) != lo.view(np.uint64))
    assert bad.size == 0, (tag, bad[:5], lg[bad[:5]], lo[bad[:5]])
    for k in METRIC_KEYS:
        assert mg[k] == getattr(mo.pm, k), (tag, k, mg[k], getattr(mo.pm, k))
    for (gx, gy), s in zip(sg, so):
        assert abs(gx - s.xx_comp) <= 1e-12 * abs(s.xx_comp) + 1e-300, (tag, gx, s.xx_comp)
        assert abs(gy - s.xy_comp) <= 1e-12 * abs(s.xy_comp) + 1e-12 * abs(s.xx_comp), (tag, gy, s.xy_comp)


@pytest.mark.parametrize("seed_index", range(8))
def test_production_fields_against_oracle(es, oracle_esm, seed_index):
    from elasticscattering_amd.workloads import PRODUCTION_FP64 as w, region

    rs, ext = region(w)
    seed = w["seeds"][seed_index]
    ppr, ppq = 64, w["particles_per_quadrant"]
    og = O.OracleGrid.generate(oracle_esm, seed, rs, ext, 0.0, w["impurity_radius"], 15, w["cells_per_row"],
                               n=w["impurity_count"])
    prm = O.sim_params(ppr, ppq, rs, ext, og.cell_size, w["tau"], w["alpha"], w["clockwise"], w["particle_speed"],
                       nthreads=THREADS)
    sweeps = {}
    with es.Simulation(ppr, ppq) as sim:
        sim.generate_impurities(seed, rs, ext, w["impurity_radius"], count=w["impurity_count"],
                                cells_per_row=w["cells_per_row"])
        screened = 0
        for coherent in (0, 1):
            sim.init_sample(w["tau"], w["alpha"], w["clockwise"], w["particle_speed"], coherent)
            for B in w["fields"]:
                announce(f"production seed={seed} coherent={coherent} B={B:.4g}")
                sg, mg, lg = sim.compute_sigmas(B, w["temperatures"], want_lifetimes=True)
                so, mo, lo = O.compute_sigmas(oracle_esm, og, prm, coherent, B, w["temperatures"],
                                              want_lifetimes=True)
                _check_fp64(sg, mg, lg, so, mo, lo, (seed, coherent, B))
                screened += mg["cross_calls"] - mg["cross_exact"]
            sweeps[coherent] = sim.compute_sweep(w["fields"], w["temperatures"])[0]
    assert screened > 0  # the screened trace ran and skipped disks
    # the bench's launch: both passes of the sample in one batch
    with es.Batch() as b:
        kw = dict(seed=seed, region_size=rs, region_extends=ext, impurity_radius=w["impurity_radius"],
                  positions_per_row=ppr, particles_per_quadrant=ppq, fields=w["fields"],
                  temperatures=w["temperatures"], tau=w["tau"], alpha=w["alpha"], is_clockwise=w["clockwise"],
                  particle_speed=w["particle_speed"], count=w["impurity_count"], cells_per_row=w["cells_per_row"])
        i = b.add(coherent=False, **kw)
        b.add(coherent=True, share_field_of=i, **kw)
        sig, _, _ = b.run()
    for coherent in (0, 1):
        assert np.array_equal(sig[coherent].view(np.uint64), sweeps[coherent].view(np.uint64))


UNEVEN_SUBSET = [4, 1, 42, 27, 14, 6, 45]  # 410 / 58,752 disks, r 5.1 / 19.5 nm, ppr 32, 64, 128, 256, 512
UNEVEN_FIELDS = [1.0, 6.3, 40.0]


@pytest.mark.parametrize("index", UNEVEN_SUBSET)
def test_batched_uneven_sample_against_oracle(es, oracle_esm, index):
    from elasticscattering_amd.workloads import batched_samples

    s = batched_samples()[index]
    ppr, ppq, T = s["positions_per_row"], s["particles_per_quadrant"], s["temperatures"]
    og = O.OracleGrid.generate(oracle_esm, s["seed"], s["region_size"], s["region_extends"], s["density"],
                               s["impurity_radius"], s["target_cell_population"])
    prm = O.sim_params(ppr, ppq, s["region_size"], s["region_extends"], og.cell_size, s["tau"], s["alpha"],
                       s["is_clockwise"], s["particle_speed"], nthreads=THREADS)
    res = {}
    for prec in ("fp64", "mixed"):
        with es.Simulation(ppr, ppq, precision=prec) as sim:
            sim.generate_impurities(s["seed"], s["region_size"], s["region_extends"], s["impurity_radius"],
                                    density=s["density"], target_cell_population=s["target_cell_population"])
            sim.init_sample(s["tau"], s["alpha"], s["is_clockwise"], s["particle_speed"], False)
            res[prec] = [sim.compute_sigmas(B, T, want_lifetimes=True) for B in UNEVEN_FIELDS]
            if prec == "fp64":
                sweep = sim.compute_sweep(UNEVEN_FIELDS, T)[0]
    for B, (sg, mg, lg), (sm, mm, lm) in zip(UNEVEN_FIELDS, res["fp64"], res["mixed"]):
        announce(f"uneven sample {index} (ppr {ppr}, {og.unique_impurity_count} disks) B={B}")
        so, mo, lo = O.compute_sigmas(oracle_esm, og, prm, 0, B, T, want_lifetimes=True)
        _check_fp64(sg, mg, lg, so, mo, lo, (index, B))
        omega = E_CHARGE * B / M_EFF
        d = np.abs(lm - lo)
        bad = np.flatnonzero(d > 1e-5 * np.abs(lo) + MIXED_ANGLE_FLOOR / omega)
        print(f"sample {index} B={B}: mixed lifetimes outside 1e-5 t + 5e-7/omega: {bad.size} of {lo.size}")
        assert bad.size == 0, (index, B, bad[:5], lo[bad[:5]], lm[bad[:5]])
        for (x, y), o in zip(sm, so):
            norm = math.hypot(o.xx_comp, o.xy_comp)
            if norm > 0:
                assert max(abs(x - o.xx_comp), abs(y - o.xy_comp)) <= 1e-6 * norm, (index, B, x, y)
    with es.Batch() as b:
        b.add(**dict(s, fields=UNEVEN_FIELDS))
        sig, _, _ = b.run()
    assert np.array_equal(sig[0].view(np.uint64), sweep.view(np.uint64))


def test_batched_uneven_subset_in_one_launch(es):
    """The whole stratified subset as one batch (one launch per stage) gives each
    sample's own solver bits."""
    from elasticscattering_amd.workloads import batched_samples

    samples = [dict(batched_samples()[i], fields=UNEVEN_FIELDS) for i in UNEVEN_SUBSET]
    announce("uneven subset: one batch")
    with es.Batch() as b:
        for s in samples:
            b.add(**s)
        sig, _, _ = b.run()
    for k, s in enumerate(samples):
        with es.Simulation(s["positions_per_row"], s["particles_per_quadrant"]) as sim:
            sim.generate_impurities(s["seed"], s["region_size"], s["region_extends"], s["impurity_radius"],
                                    density=s["density"], target_cell_population=s["target_cell_population"])
            sim.init_sample(s["tau"], s["alpha"], s["is_clockwise"], s["particle_speed"], False)
            ref, _ = sim.compute_sweep(UNEVEN_FIELDS, s["temperatures"])
        assert np.array_equal(sig[k].view(np.uint64), ref.view(np.uint64)), k
