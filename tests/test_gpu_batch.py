"""Batched path: many uneven samples in one launch per stage (needs an MI355X).

Bars: every entry of a batch gives bit-identical sigmas and equal metrics to
the same sample run alone through es_solver (each pair keeps its own work
items and its own fixed reduction order, whatever else shares the launch);
entries are also checked against the CPU oracle at the bars of the one-sample
parity tests; graph replays, launch segments and every tuner shape give the
eager result bit for bit.
"""
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import announce

pytestmark = pytest.mark.gpu

METRIC_KEYS = ("cells_passed", "impurities_tested", "particles_inside_impurity", "particles_escaped",
               "particles_at_bound", "max_impurities_tested", "max_cells_passed", "trace_cap_hits",
               "real_particles", "avg_particle_lifetime", "traced_particles", "edge_hits", "cross_calls",
               "sigma_evals")
# cross_exact counts the crossing solves a kernel actually ran: the screened
# cooperative trace skips disks that cannot lower the minimum, so it depends on
# the launch shape and is only bounded here.

RS, EXT = 1e-6, 5e-7


def _entries():
    """Uneven entries: sizes, grids, fields, temperatures, passes, a shared field."""
    common = dict(region_size=RS, region_extends=EXT, is_clockwise=1, particle_speed=7e5)
    return [
        dict(seed=101, count=2000, impurity_radius=1.1e-8, cells_per_row=32, positions_per_row=64,
             particles_per_quadrant=17, fields=[20.0, 5.0, 1.3], temperatures=[0.0], tau=1e-12,
             alpha=math.pi / 4, coherent=False, incoherent_tau=1e-12, **common),
        dict(seed=7, count=600, impurity_radius=2e-8, target_cell_population=4, positions_per_row=24,
             particles_per_quadrant=9, fields=[1.0, 40.0], temperatures=[2.0, 40.0, 100.0], tau=1e-12,
             alpha=math.pi / 4, coherent=False, **common),
        dict(seed=0, share_field_of=1, impurity_radius=2e-8, target_cell_population=4, positions_per_row=24,
             particles_per_quadrant=9, fields=[3.0, 11.0], temperatures=[2.0, 150.0], tau=1e-12, alpha=0.3,
             coherent=True, **common),
        dict(seed=13, count=3000, impurity_radius=5e-9, target_cell_population=15, positions_per_row=40,
             particles_per_quadrant=5, fields=[3.0], temperatures=[], tau=1e-12, alpha=0.3, coherent=False,
             incoherent_tau=2e-13, **common),
        dict(seed=17, count=50, impurity_radius=1e-8, target_cell_population=2, positions_per_row=16,
             particles_per_quadrant=3, fields=[0.7, 2.0, 9.0, 30.0],
             temperatures=[1.0, 2.0, 5.0, 10.0, 20.0, 50.0, 100.0, 150.0], tau=1e-12, alpha=math.pi / 4,
             coherent=False, **common),
    ]


def _add(batch, e):
    kw = dict(e)
    return batch.add(**kw)


def _solo(es, e, entries):
    """The entry run alone through es_solver: (sigma [nB, nT, 2], metrics per field)."""
    src = entries[e["share_field_of"]] if e.get("share_field_of", -1) >= 0 else e
    with es.Simulation(e["positions_per_row"], e["particles_per_quadrant"]) as sim:
        sim.generate_impurities(src["seed"], RS, EXT, src["impurity_radius"], count=src.get("count", 0),
                                target_cell_population=src.get("target_cell_population", 15),
                                cells_per_row=src.get("cells_per_row", 0))
        sim.init_sample(e["tau"], e["alpha"], e["is_clockwise"], e["particle_speed"], e["coherent"],
                        e.get("incoherent_tau", 0.0))
        return sim.compute_sweep(e["fields"], e["temperatures"])


def _same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _same_results(r0, r1):
    sig0, m0, _ = r0
    sig1, m1, _ = r1
    assert len(sig0) == len(sig1)
    for e, (a, b) in enumerate(zip(sig0, sig1)):
        assert _same_bits(a, b), f"entry {e}: {a} vs {b}"
    if m0 is not None and m1 is not None:
        for e, (ma, mb) in enumerate(zip(m0, m1)):
            for fa, fb in zip(ma, mb):
                for k in METRIC_KEYS:
                    assert fa[k] == fb[k], (e, k, fa[k], fb[k])
                for f in (fa, fb):
                    assert 0 <= f["cross_exact"] <= f["cross_calls"], f


@pytest.fixture(scope="module")
def es():
    import elasticscattering_amd as es_mod

    es_mod.load_library()
    return es_mod


@pytest.fixture(scope="module")
def batch_eager(es):
    entries = _entries()
    announce("batch of 5 uneven entries: eager run")
    b = es.Batch()
    for e in entries:
        _add(b, e)
    res = b.run()
    yield entries, b, res
    b.close()


def test_batch_counts(batch_eager):
    entries, b, res = batch_eager
    c = b.counts()
    assert c["samples"] == len(entries)
    assert c["pairs"] == sum(len(e["fields"]) for e in entries)
    assert c["outputs"] == sum(len(e["fields"]) * len(e["temperatures"]) for e in entries)
    traced = sum(m["traced_particles"] for per in res[1] for m in per)
    assert c["traced_items"] == traced


def test_batch_entries_equal_solo_runs(es, batch_eager):
    entries, b, (sig, mets, secs) = batch_eager
    assert secs > 0
    for i, e in enumerate(entries):
        announce(f"batch entry {i} vs es_solver alone")
        sig_s, ms_s = _solo(es, e, entries)
        assert _same_bits(sig[i], sig_s), f"entry {i}: {sig[i]} vs {sig_s}"
        for bi, (mb, msolo) in enumerate(zip(mets[i], ms_s)):
            for k in METRIC_KEYS:
                assert mb[k] == msolo[k], (i, bi, k, mb[k], msolo[k])


@pytest.mark.parametrize("entry", [1, 2, 4])
def test_batch_entries_against_oracle(es, oracle_esm, batch_eager, entry):
    entries, b, (sig, mets, _) = batch_eager
    e = entries[entry]
    src = entries[e["share_field_of"]] if e.get("share_field_of", -1) >= 0 else e
    announce(f"batch entry {entry} vs oracle")
    og = O.OracleGrid.generate(oracle_esm, src["seed"], RS, EXT, 0.0, src["impurity_radius"],
                               src.get("target_cell_population", 15), src.get("cells_per_row", 0),
                               n=src.get("count", 0))
    prm = O.sim_params(e["positions_per_row"], e["particles_per_quadrant"], RS, EXT, og.cell_size, e["tau"],
                       e["alpha"], e["is_clockwise"], e["particle_speed"], e.get("incoherent_tau", 0.0), nthreads=8)
    for bi, B in enumerate(e["fields"]):
        so, mo, _ = O.compute_sigmas(oracle_esm, og, prm, int(e["coherent"]), B, e["temperatures"])
        m = mets[entry][bi]
        assert m["cells_passed"] == mo.pm.cells_passed
        assert m["impurities_tested"] == mo.pm.impurities_tested
        assert m["particles_escaped"] == mo.pm.particles_escaped
        assert m["real_particles"] == mo.real_particles
        for t, o in enumerate(so):
            gxx, gxy = sig[entry][bi, t]
            atol = 1e-12 * abs(o.xx_comp)
            assert abs(gxx - o.xx_comp) <= 1e-12 * abs(o.xx_comp) + 1e-300
            assert abs(gxy - o.xy_comp) <= 1e-12 * abs(o.xy_comp) + atol
            assert abs(gxx - o.xx_naive) <= 1e-8 * abs(o.xx_naive) + 1e-300
            assert abs(gxy - o.xy_naive) <= 1e-8 * abs(o.xy_naive) + atol


def test_graph_replays_equal_eager(es, batch_eager):
    entries, b, eager = batch_eager
    announce("hipGraph capture + 3 replays with eager work between")
    b.set_capture(True)
    try:
        other = es.Simulation(24, 5)
        other.generate_impurities(99, RS, EXT, 1e-8, count=500, target_cell_population=4)
        other.init_sample(1e-12, 0.3, 1, 7e5, False)
        for rep in range(3):
            r = b.run()
            _same_results(eager, r)
            other.compute_sigmas(2.0 + rep, [10.0])  # eager work on another stream between replays
            np.sort(np.random.default_rng(rep).random(100000))  # and host work
        other.close()
    finally:
        b.set_capture(False)
    _same_results(eager, b.run())


def test_segments_equal_one_launch(es, batch_eager):
    entries, b, eager = batch_eager
    announce("same batch cut into launch segments of <= 20000 items")
    old = os.environ.get("ES_SEGMENT_ITEMS")
    os.environ["ES_SEGMENT_ITEMS"] = "20000"
    try:
        b2 = es.Batch()
    finally:
        if old is None:
            del os.environ["ES_SEGMENT_ITEMS"]
        else:
            os.environ["ES_SEGMENT_ITEMS"] = old
    with b2:
        for e in entries:
            _add(b2, e)
        _same_results(eager, b2.run())
        b2.set_capture(True)
        _same_results(eager, b2.run())
        _same_results(eager, b2.run())


def test_tuner_shapes_give_same_results(es, batch_eager):
    entries, b, eager = batch_eager
    announce("tuner over launch shapes")
    best, t = b.tune()
    assert best["waves_per_eu"] in (2, 3) and best["chunk"] in (64, 128, 256, 512)
    assert best["grid_mult"] in (1, 2) and best["cooperative"] in (1, 2) and t > 0
    assert b.shape == best
    _same_results(eager, b.run())
    # one lane per particle and the wave-cooperative crossing solves: same bits
    for shape in ((2, 32, 2, 1), (3, 512, 1, 1), (2, 64, 1, 2), (3, 32, 2, 2), (2, 32, 2, 0), (3, 512, 1, 0),
                  (3, 64, 1, 0)):
        b.set_shape(*shape)
        _same_results(eager, b.run())
    b.set_shape(3, 256, 1, -1)  # back to the default shape


def test_clear_reuses_fields_with_same_bits(es, batch_eager):
    """After es_batch_clear, entries with a known recipe take the kept field
    (its positions recomputed when they differ); results equal a fresh batch."""
    entries, _, _ = batch_eager
    announce("batch clear: kept fields against a fresh batch")
    later = [dict(entries[0], fields=[4.0, 20.0]), dict(entries[3], positions_per_row=24, temperatures=[7.0]),
             dict(entries[0], temperatures=[3.0]), dict(entries[4], seed=18)]
    with es.Batch() as b:
        for e in entries:
            _add(b, e)
        b.run()
        b.clear()
        for e in later:
            _add(b, e)
        got = b.run()
        b.clear()
        with pytest.raises(es.EsError):
            b.run()  # empty again
    with es.Batch() as fresh:
        for e in later:
            _add(fresh, e)
        _same_results(fresh.run(), got)


def test_reset_redraws_into_kept_memory_with_same_bits(es, batch_eager):
    """After es_batch_reset every entry draws and bins its field again (a changed
    seed, count or radius lands in the memory of a forgotten field); results
    equal a fresh batch, and a second reset + refill gives the same bits again."""
    entries, _, _ = batch_eager
    announce("batch reset: redrawn fields against a fresh batch")
    later = [dict(entries[1], seed=77), dict(entries[0], temperatures=[3.0]),
             dict(entries[3], impurity_radius=entries[3]["impurity_radius"] * 1.5)]
    with es.Batch() as b:
        for e in entries:
            _add(b, e)
        b.run()
        b.reset()
        for e in later:
            _add(b, e)
        got = b.run()
        b.reset()
        for e in later:
            _add(b, e)
        again = b.run()
    with es.Batch() as fresh:
        for e in later:
            _add(fresh, e)
        want = fresh.run()
    _same_results(want, got)
    _same_results(want, again)


def test_launch_collect_pipelined_with_same_bits(es, batch_eager):
    """es_batch_launch / es_batch_collect on two batches whose runs overlap on
    the GPU give the bits of es_batch_run; a launched batch refuses other
    calls until collected."""
    entries, _, eager = batch_eager
    announce("batch launch / collect: two pipelined batches")
    with es.Batch() as b1, es.Batch() as b2:
        for e in entries:
            _add(b1, e)
        b1.launch()
        with pytest.raises(es.EsError):
            _add(b1, entries[0])  # not collected yet
        with pytest.raises(es.EsError):
            b1.launch()
        for e in entries:
            _add(b2, e)
        b2.launch()
        r1 = b1.collect()
        r2 = b2.collect()
        with pytest.raises(es.EsError):
            b1.collect()  # nothing launched
        _same_results(eager, r1)
        _same_results(eager, r2)
        b1.reset()
        for e in entries:
            _add(b1, e)
        _same_results(eager, b1.run())


def test_prepare_ahead_of_another_trace_with_same_bits(es, batch_eager):
    """es_batch_prepare enqueues a batch's binning before another batch's trace
    is enqueued, three batches in the rotation of a pipelined loop: every run
    gives the bits of es_batch_run; prepare with nothing pending is a no-op, and
    a launched batch refuses it."""
    entries, _, eager = batch_eager
    announce("batch prepare: three batches, binning one step ahead")
    with es.Batch() as b0, es.Batch() as b1, es.Batch() as b2:
        bufs = (b0, b1, b2)
        results = []
        for e in entries:
            _add(b0, e)
        b0.prepare()
        b0.prepare()  # nothing pending
        n = 4
        for k in range(n):
            if k + 1 < n:
                nxt = bufs[(k + 1) % 3]
                nxt.reset()
                for e in entries:
                    _add(nxt, e)
                nxt.prepare()
            bufs[k % 3].launch()
            with pytest.raises(es.EsError):
                bufs[k % 3].prepare()  # launched, not collected
            if k > 0:
                results.append(bufs[(k - 1) % 3].collect())
        results.append(bufs[(n - 1) % 3].collect())
        for r in results:
            _same_results(eager, r)


def test_batch_errors(es):
    announce("batch argument errors")
    with es.Batch() as b:
        with pytest.raises(es.EsError):
            b.run()  # empty batch
        e = _entries()[1]
        _add(b, e)
        bad = dict(e, is_clockwise=0)
        _add(b, bad)
        with pytest.raises(es.EsError) as ei:
            b.run()  # two rotation senses
        assert "rotation" in str(ei.value)
    with es.Batch() as b:
        with pytest.raises(es.EsError):
            _add(b, dict(_entries()[1], temperatures=list(range(1, 4098))))  # more than 4096 temperatures
    with es.Batch() as b:
        _add(b, dict(_entries()[1], temperatures=list(range(1, 10))))
        with pytest.raises(es.EsError) as ei:
            b.launch()  # launch / collect take up to 8 temperatures per entry
        assert "es_batch_run" in str(ei.value)
        with pytest.raises(es.EsError):
            _add(b, dict(_entries()[1], share_field_of=3))  # no such entry
        with pytest.raises(es.EsError):
            b.set_shape(4, 128, 1)


def test_batched_uneven_workload_subset(es):
    """Six samples of the 64-sample uneven workload (position grids capped at
    64 per row) against their solo runs."""
    from elasticscattering_amd.workloads import batched_samples

    samples = batched_samples(n_samples=6, max_ppr=64)
    announce("uneven workload subset: 6 samples, ppr <= 64")
    with es.Batch() as b:
        for s in samples:
            b.add(**s)
        sig, mets, _ = b.run()
    for i, s in enumerate(samples):
        with es.Simulation(s["positions_per_row"], s["particles_per_quadrant"]) as sim:
            sim.generate_impurities(s["seed"], s["region_size"], s["region_extends"], s["impurity_radius"],
                                    density=s["density"], target_cell_population=s["target_cell_population"])
            sim.init_sample(s["tau"], s["alpha"], s["is_clockwise"], s["particle_speed"], s["coherent"])
            ref, _ = sim.compute_sweep(s["fields"], s["temperatures"])
        assert _same_bits(sig[i], ref), i


MANY_T = [0.5 + 7.3 * k for k in range(19)]  # 19 temperatures: chunks of 8, 8 and 3


@pytest.mark.parametrize("segment_items", [None, "20000"])
def test_more_than_eight_temperatures(es, segment_items):
    """Entries with 19, 3 and 0 temperatures in one batch (the reference takes
    any number, SimulationConfiguration.h:186-196): every sigma bit-equal to the
    solver's (which chunks its temperatures the same way) and to the batch's
    own 8-temperature runs; metrics unchanged; with one launch segment (the
    later chunks reuse the traced lifetimes) and with several (traced again);
    with graph capture on."""
    announce(f"batch with 19 temperatures, segments {segment_items or 'one'}")
    base = _entries()
    entries = [dict(base[1], temperatures=MANY_T), dict(base[2], temperatures=MANY_T[:3], share_field_of=0),
               dict(base[3], temperatures=[])]
    old = os.environ.get("ES_SEGMENT_ITEMS")
    if segment_items:
        os.environ["ES_SEGMENT_ITEMS"] = segment_items
    try:
        b = es.Batch()
    finally:
        if old is None:
            os.environ.pop("ES_SEGMENT_ITEMS", None)
        else:
            os.environ["ES_SEGMENT_ITEMS"] = old
    with b:
        for e in entries:
            _add(b, e)
        if segment_items:
            assert b.launches()["segments"] > 1
        sig, mets, secs = b.run()
        assert secs > 0
        assert sig[0].shape == (len(base[1]["fields"]), 19, 2) and sig[2].size == 0
        sig_s, ms_s = _solo(es, entries[0], entries)
        assert _same_bits(sig[0], sig_s)
        sig_s, _ = _solo(es, entries[1], entries)
        assert _same_bits(sig[1], sig_s)
        for bi, (mb, msolo) in enumerate(zip(mets[0], ms_s)):
            for k in METRIC_KEYS:
                assert mb[k] == msolo[k], (bi, k)
        b.set_capture(True)
        for _ in range(2):
            sig2, mets2, _ = b.run()
            for x, y in zip(sig, sig2):
                assert _same_bits(x, y)
    with es.Batch() as b8:  # the first 8 and the next 8 as batches of their own
        for t0 in (0, 8, 16):
            _add(b8, dict(entries[0], temperatures=MANY_T[t0:t0 + 8]))
        parts, _, _ = b8.run()
    assert _same_bits(sig[0], np.concatenate(parts, axis=1))
