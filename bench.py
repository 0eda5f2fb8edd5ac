#!/usr/bin/env python3
"""Throughput of the lifetime / sigma path on one MI355X.

A step is the whole production sweep: 8 impurity samples, each field drawn
and binned on the GPU, and for each sample the incoherent and the coherent
pass over the 40 fields with the 3 temperatures (SimulationRunner.cpp:88-98),
run as one batch: one launch per stage for all samples, passes and fields.
Steps are pipelined over two batches: the next step's samples are added (and
their fields binned) while the current one runs.  The roofline's kernel times
come from two more steps run alone.

Prints one JSON line: lifetimes/s of the FP64 path (incoherent lifetimes, the
coherent pass is run but not counted; `value`), the sweep wall time, the
roofline of the trace kernel (algorithmic FP64 work per launch from the
kernel's own work counters over its HIP-event duration, against the
78.6 TFLOP/s FP64 vector peak), the CPU oracle timed on a bounded sample of
the same workload on this host, and beside them:
  * "mixed": the same steps with the mixed-precision trace, and its sigma error
    against the FP64 steps;
  * "batched_uneven": the 64-sample uneven batch, FP64 and mixed;
  * "sweep": 400-field sigma(B) curves of seeded materials at 128 x 128
    positions — the full FP64 sweep against the adaptive sweep (mixed
    precision, batched, surrogate-guided; the surrogate trained first on
    solver runs), with the adaptive curves' error against the full sweep.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_VECTOR_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

# Algorithmic FP64 operations per event (add, sub, mul, div, sqrt, floor,
# rint = 1; fma = 2), counted from the operation sequence of the reference
# functions as restated in es_trace.h / es_math.h.  The trace kernel
# (k_trace) does CreateParticle + TraceOrbit; k_sigma the temperature stage.
TRACE_WORK = {
    "traced_particles": 85,   # CreateParticle (sincos 58, orbit centre, bound time), start cell
    "cells_passed": 95,       # GetNextCell: corners, 4 edge projections, snap, stop test
    "edge_hits": 52,          # per accepted edge intersection: sqrt, points, GetAngle (asin), GetCrossAngle
    "impurities_tested": 5,   # CirclesCross
    "cross_exact": 121,       # GetCrossPoints + 2 GetAngle + 2 AngleInRange + times (per crossing disk solved)
}
# The reference solves every crossing disk (cross_calls); the screened trace
# skips those that cannot lower the minimum (cross_exact of them are solved).
REFERENCE_TRACE_WORK = dict(TRACE_WORK, cross_exact=0, cross_calls=121)
SIGMA_WORK = {
    "traced_particles": 70,   # phi, sincos, Simpson weight, lifetime sum
    "sigma_evals": 78,        # truncation, exp, sincos, GetSigma, two weighted terms, compensated adds
}


def _json(obj):
    """One JSON line; numpy scalars (metrics records) as plain numbers."""
    return json.dumps(obj, default=lambda o: o.item() if hasattr(o, "item") else str(o))


def progress(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_flops(m, model):
    """Work of one metrics record, or of a record array (summed)."""
    return float(sum(model[k] * m[k].sum() if hasattr(m[k], "sum") else model[k] * m[k] for k in model))


def cpu_baseline(w, seconds_target=20.0):
    """The oracle (libm build) on the first sample's incoherent pass at 4 fields spread over the sweep:
    1 thread (the reference CPU path is single-threaded, SimulationCPU.cpp:112-135), then the same
    sample over the host's cores (rows split across threads, up to 16)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from elasticscattering_amd.workloads import region

    lib = O.load("libm")
    if lib is None:
        return None
    rs, ext = region(w)
    og = O.OracleGrid.generate(lib, w["seeds"][0], rs, ext, 0.0, w["impurity_radius"], 15, w["cells_per_row"],
                               n=w["impurity_count"])
    fs = w["fields"]
    fields = [fs[0], fs[len(fs) // 3], fs[2 * len(fs) // 3], fs[-1]]
    P = w["positions_per_row"] ** 2 * 4 * w["particles_per_quadrant"]

    def timed_run(nthreads):
        prm = O.sim_params(w["positions_per_row"], w["particles_per_quadrant"], rs, ext, og.cell_size, w["tau"],
                           w["alpha"], w["clockwise"], w["particle_speed"], w["incoherent_tau"], nthreads=nthreads)
        t0 = time.perf_counter()
        done = 0
        for B in fields:
            O.compute_sigmas(lib, og, prm, 0, B, w["temperatures"])
            done += 1
            if time.perf_counter() - t0 > seconds_target:
                break
        return done, time.perf_counter() - t0

    done, dt = timed_run(1)
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    done_t, dt_t = timed_run(threads)
    sample = (f"CPU oracle (restated reference CPU path, glibc libm, -O2) on sample seed {w['seeds'][0]}, "
              f"incoherent pass, B in {[round(float(f), 6) for f in fields[:done]]} T, "
              f"{len(w['temperatures'])} temperatures each: {done * P} lifetimes + sigma stage in {dt:.1f} s")
    return {
        "value": done * P / dt,
        "unit": "lifetimes/s",
        "cores": 1,
        "kind": "port",
        "sample": sample,
        "seconds": dt,
        "threads": {"value": done_t * P / dt_t, "cores": threads, "seconds": dt_t},
    }


def calibrate_copy(mib):
    """Plain device copy of known size: FETCH_SIZE / WRITE_SIZE per dispatch of this
    kernel scale the counters of the trace kernel to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_per_dispatch": n * 4}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="production")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the FP64 production line (profiling runs)")
    ap.add_argument("--part", default="all", choices=("all", "production", "batched", "sweep"),
                    help="batched / sweep: print only that part's JSON (profiling runs)")
    ap.add_argument("--sweep-materials", type=int, default=200)
    ap.add_argument("--train-runs", type=int, default=20000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 4 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by sample and field, "
              "see DESIGN.md", file=sys.stderr)
        return 2

    import numpy as np

    import elasticscattering_amd as es
    from elasticscattering_amd.workloads import PRODUCTION_FP64, SMALL_FP64, region, lifetimes_per_pass

    if args.part == "batched":
        print(_json({"batched_uneven": batched_uneven(es)}), flush=True)
        return 0
    if args.part == "sweep":
        print(_json({"sweep": sweep_comparison(args.sweep_materials, args.train_runs)}), flush=True)
        return 0
    if args.part == "production":
        args.no_extras = True

    w = PRODUCTION_FP64 if args.workload == "production" else SMALL_FP64
    rs, ext = region(w)
    ppr, ppq = w["positions_per_row"], w["particles_per_quadrant"]
    fields, temps = w["fields"], w["temperatures"]
    P = lifetimes_per_pass(w)

    def production_entries(b):
        """The 8 samples as batch entries: per sample its incoherent pass (drawn and binned
        on the GPU) and its coherent pass on the same field (SimulationRunner.cpp:88-98)."""
        b.reset()  # every step draws and bins its fields again
        for seed in w["seeds"]:
            kw = dict(seed=seed, region_size=rs, region_extends=ext, impurity_radius=w["impurity_radius"],
                      positions_per_row=ppr, particles_per_quadrant=ppq, fields=fields, temperatures=temps,
                      tau=w["tau"], alpha=w["alpha"], is_clockwise=w["clockwise"],
                      particle_speed=w["particle_speed"], incoherent_tau=w["incoherent_tau"],
                      count=w["impurity_count"], cells_per_row=w["cells_per_row"])
            i = b.add(coherent=False, **kw)
            b.add(coherent=True, share_field_of=i, **kw)

    COUNTERS = ("traced_particles", "cells_passed", "impurities_tested", "edge_hits", "cross_calls", "cross_exact",
                "sigma_evals")

    def reduce_runs(runs):
        st = {"launch_s": [], "flops": [], "ref_flops": [], "sigma_flops": [], "sigma_s": [], "sigma": [],
              "counters": []}
        tot = dict(traced=0, cross=0, exact=0, recomputed=0)
        for sig, mets, _ in runs:
            # Every (entry, field) pair of the one launch, with the share of its counters
            # that was executed: the coherent window at alpha = pi/4 has zero width, so its
            # ppq phi points are one particle, traced once, and its counters (which count
            # all ppq lifetimes, as the reference's metrics do) are ppq times the work done.
            ms = [(m, 1.0 if e % 2 == 0 else 1.0 / ppq) for e, per in enumerate(mets) for m in per]
            st["launch_s"].append(sum(m["time_elapsed_lifetimes"] for m, _ in ms))
            st["sigma_s"].append(sum(m["time_elapsed_temperatures"] for m, _ in ms))
            st["flops"].append(sum(x * algorithmic_flops(m, TRACE_WORK) for m, x in ms))
            st["ref_flops"].append(sum(x * algorithmic_flops(m, REFERENCE_TRACE_WORK) for m, x in ms))
            st["sigma_flops"].append(sum(x * algorithmic_flops(m, SIGMA_WORK) for m, x in ms))
            st["sigma"].append(np.stack(sig[0::2]))  # incoherent passes [sample, field, T, 2]
            # work the launch executed: the coherent window's particles counted once
            st["counters"].append({k: int(round(sum(m[k] * x for m, x in ms))) for k in COUNTERS})
            for key, col in (("traced", "traced_particles"), ("cross", "cross_calls"), ("exact", "cross_exact"),
                             ("recomputed", "fp64_recomputed")):
                tot[key] += sum(round(m[col] * x) for m, x in ms)
        st.update(tot)
        return st

    def timed(precision):
        """K pipelined steps over three batches: step k's fields are drawn and binned
        (prepare) before step k-1's trace is enqueued, so that binning runs ahead and
        step k's launch finds its read-backs done and enqueues its trace at once; then
        step k-2's results are collected.  Then two steps run alone, for the kernel
        times of the roofline."""
        with es.Batch() as b0, es.Batch() as b1, es.Batch() as b2:
            bufs = (b0, b1, b2)
            for b in bufs:
                b.set_precision(precision)

            def pipelined(n):
                production_entries(bufs[0])
                bufs[0].prepare()
                out = []
                for k in range(n):
                    if k + 1 < n:
                        nxt = bufs[(k + 1) % 3]
                        production_entries(nxt)
                        nxt.prepare()
                    bufs[k % 3].launch()
                    if k > 0:
                        out.append(bufs[(k - 1) % 3].collect())
                out.append(bufs[(n - 1) % 3].collect())
                return out

            pipelined(max(1, args.warmup))
            t0 = time.perf_counter()
            runs = pipelined(args.steps)
            dt_ = time.perf_counter() - t0
            alone = []
            for _ in range(2):
                production_entries(b0)
                alone.append(b0.run())
        st = reduce_runs(runs)
        iso = reduce_runs(alone)
        st["launch_s"], st["flops"], st["ref_flops"] = iso["launch_s"], iso["flops"], iso["ref_flops"]
        st["sigma_flops"], st["sigma_s"] = iso["sigma_flops"], iso["sigma_s"]
        st["counters"] = iso["counters"]
        return st, dt_

    progress("production FP64")
    stats, dt = timed("fp64")
    mixed = None
    if not args.no_extras:
        progress("production mixed")
        mstats, mdt = timed("mixed")
        err = 0.0
        for a, b in zip(mstats["sigma"], stats["sigma"]):
            n = np.hypot(b[..., 0], b[..., 1])
            err = max(err, float((np.abs(a - b).max(axis=-1) / n).max()))
        mixed = {
            "value": args.steps * len(w["seeds"]) * len(fields) * P / mdt, "unit": "lifetimes/s",
            "ms_per_step": mdt / args.steps * 1e3, "sweep_wall_s": mdt / args.steps,
            "trace_ms_per_launch": sum(mstats["launch_s"]) / len(mstats["launch_s"]) * 1e3,
            "max_sigma_error_vs_fp64": err, "fp64_recomputed_fraction": mstats["recomputed"] / max(mstats["traced"], 1),
            "dtype": "f32 trace + f64 recompute, f64 start data / Simpson / sigma sums",
        }

    lifetimes = args.steps * len(w["seeds"]) * len(fields) * P
    avg_launch = sum(stats["launch_s"]) / len(stats["launch_s"])
    avg_flops = sum(stats["flops"]) / len(stats["flops"])
    achieved = avg_flops / avg_launch / 1e12
    sigma_tf = (sum(stats["sigma_flops"]) / sum(stats["sigma_s"]) / 1e12) if sum(stats["sigma_s"]) > 0 else None
    traffic = None
    tf_path = os.path.join(ROOT, "profiles", "k_trace_production_fp64_traffic.json")
    if w is PRODUCTION_FP64 and os.path.exists(tf_path):
        traffic = json.load(open(tf_path)).get("hbm_bytes_per_launch")
    line = {
        "metric": "orbit lifetimes/sec + sigma(B) sweep wall-time at <=1e-3 rel err vs FP64",
        "value": lifetimes / dt,
        "unit": "lifetimes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": w["name"],
            "impurities": w["impurity_count"], "side_m": w["side"], "impurity_radius_m": w["impurity_radius"],
            "cells": f"{w['cells_per_row']}x{w['cells_per_row']}", "positions": f"{ppr}x{ppr}",
            "phi_points": 4 * ppq, "fields": len(fields), "temperatures": temps, "samples_per_step": len(w["seeds"]),
            "passes": "incoherent (counted) + coherent (run, alpha=pi/4 gives zero weight)",
        },
        "sweep_wall_s": dt / args.steps,
        "traced_lifetimes_per_s": stats["traced"] / dt,
        "exact_cross_solves_fraction": stats["exact"] / max(stats["cross"], 1),
        "roofline": {
            "bound": "mfma",
            "bound_detail": "FP64 vector ALU (VALU) work: the kernel issues no MFMA; 'mfma' is the contract's compute "
                            "category and MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s.  HBM traffic is "
                            "a few % of peak (profiles/k_trace_production_fp64.md)",
            "kernel": "k_trace (CreateParticle + TraceOrbit; one launch per step: 8 samples, both passes, 40 fields)",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": traffic,
            "avg_launch_ms": avg_launch * 1e3,
            "algorithmic_flops_per_launch": avg_flops,
            # the same launch priced at the reference's work (every crossing disk solved)
            "reference_work_tflops": sum(stats["ref_flops"]) / len(stats["ref_flops"]) / avg_launch / 1e12,
            "sigma_stage_tflops": sigma_tf,
            "sigma_stage_ms": sum(stats["sigma_s"]) / len(stats["sigma_s"]) * 1e3,
            "work_per_launch": stats["counters"][-1],
        },
    }
    if mixed is not None:
        line["mixed"] = mixed
        progress("batched uneven (config 3)")
        line["batched_uneven"] = batched_uneven(es)
        progress("sweep (config 5)")
        line["sweep"] = sweep_comparison(args.sweep_materials, args.train_runs)
    if not args.no_cpu_baseline:
        progress("cpu baseline")
        cb = cpu_baseline(w)
        if cb is not None:
            line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
            # the same sample over the host's cores (a second baseline, not the contract's)
            line["cpu_baseline_threads"] = dict(cb["threads"], unit="lifetimes/s", kind="port")
    print(_json(line), flush=True)
    return 0


def batched_uneven(es):
    """The 64-sample uneven batch (config 3), FP64 and mixed.

    The launch shape comes from the project's tuner (es_batch_tune) run before
    the timed region on a stratified copy of the batch: all 64 samples at 4 of
    the 40 fields (1, 3.4, 11.7, 40 T), a tenth of the work, so that trying
    every shape takes seconds.  The full batch then runs with that shape and
    hipGraph capture on: the first (untimed) run records every stage launch of
    every segment in one graph, the timed run replays it."""
    import numpy as np
    from elasticscattering_amd.workloads import BATCHED_UNEVEN, batched_samples

    samples = batched_samples()
    fields = BATCHED_UNEVEN["fields"]
    tune_fields = [fields[0], fields[13], fields[26], fields[39]]
    out = {"workload": BATCHED_UNEVEN["name"]}
    sig = {}
    for prec in ("fp64", "mixed"):
        t0 = time.perf_counter()
        with es.Batch() as tb:
            tb.set_precision(prec)
            for smp in samples:
                tb.add(**dict(smp, fields=tune_fields))
            progress(f"tuning {prec}")
            shape, tuned_s = tb.tune()
        tune_wall = time.perf_counter() - t0
        with es.Batch() as b:
            b.set_precision(prec)
            b.set_shape(**shape)
            b.set_capture(True)
            for smp in samples:
                b.add(**smp)
            c = b.counts()
            b.run(metrics=False)  # records the graph
            t0 = time.perf_counter()
            sig[prec], mets, secs = b.run()
            wall = time.perf_counter() - t0
            launches = b.launches()
        traced = sum(m["traced_particles"] for per in mets for m in per)
        out[prec] = {
            "traced_lifetimes_per_s": traced / secs, "device_s": secs, "wall_s": wall,
            "fp64_recomputed": sum(m["fp64_recomputed"] for per in mets for m in per),
            "tuned_shape": shape, "tune_trace_s_on_stratified_copy": tuned_s, "tune_wall_s": tune_wall,
            "graph": True, "kernel_launches_per_run": launches["kernels"], "segments": launches["segments"],
            "graph_launches_per_run": launches["graph_launches"],
        }
        out.update(samples=c["samples"], pairs=c["pairs"], traced_items=c["traced_items"])
    err = 0.0
    for a, b in zip(sig["mixed"], sig["fp64"]):
        n = np.maximum(np.hypot(b[..., 0], b[..., 1]), 1e-300)
        err = max(err, float((np.abs(a - b).max(axis=-1) / n).max()))
    out["mixed"]["max_sigma_error_vs_fp64"] = err
    out["mixed_speedup"] = out["fp64"]["device_s"] / out["mixed"]["device_s"]
    return out


def sweep_comparison(n_materials, train_runs):
    """sigma(B) curves on 400 log-spaced fields (0.1-60 T) for seeded materials:
    the full FP64 sweep (every field, batched) against the adaptive sweep
    (mixed precision, batched rounds, surrogate-guided)."""
    import math

    import numpy as np
    from elasticscattering_amd import surrogate as S

    out = {"materials": n_materials, "fields": 400, "positions": "128x128", "phi_points": 260,
           "tolerance": 1e-3, "error": "max(|d xx|, |d xy|) / |sigma| against the full FP64 sweep"}
    t0 = time.perf_counter()
    n_mat = max(1, train_runs // 40)
    progress("surrogate training data")
    X, Y, gen_s, _ = S.training_set(n_mat, 20, 2, seed=1)
    sur = S.Surrogate(members=4, seed=3)
    sur.fit(X, Y, epochs=600, batch=256, lr=2e-3, lr_final=2e-5)
    train_s, train_steps = sur.mlp.last_train_time()
    out["surrogate"] = {"solver_runs": len(X), "solver_device_s": gen_s, "train_device_s": train_s,
                        "train_steps": train_steps, "wall_s": time.perf_counter() - t0}
    B = np.logspace(-1, math.log10(60), 400)
    mats = S.draw_materials(n_materials, seed=77)
    t0 = time.perf_counter()
    progress("full FP64 sweep")
    full, full_s, full_items = S.solver_runs(mats, [B] * n_materials)
    full_wall = time.perf_counter() - t0
    full = np.stack([f[:, 0, :] for f in full])
    # The solver's own jitter per curve: how far each point lies from the mean
    # of its two neighbours (400 log-spaced fields resolve the curvature).
    norm = np.maximum(np.hypot(full[..., 0], full[..., 1]), 1e-300)
    jitter = np.abs(full[:, 1:-1] - 0.5 * (full[:, :-2] + full[:, 2:])).max(axis=-1) / norm[:, 1:-1]
    jitter = jitter.max(axis=1)
    out["full_fp64"] = {"device_s": full_s, "wall_s": full_wall, "traced_items": full_items,
                        "curves_with_jitter_above_half_tolerance": int((jitter > 5e-4).sum())}
    for name, sg, prec in (("adaptive_mixed_surrogate", sur, "mixed"), ("adaptive_mixed", None, "mixed"),
                           ("adaptive_fp64", None, "fp64")):
        # one untimed run first: the first use of a precision loads its kernels
        # and sizes the batch buffers, which would land on whichever variant ran first
        progress(f"adaptive sweep {name}")
        S.adaptive_sweep(mats, B, surrogate=sg, precision=prec)
        t0 = time.perf_counter()
        est, rep = S.adaptive_sweep(mats, B, surrogate=sg, precision=prec)
        wall = time.perf_counter() - t0
        err = S.tensor_error(est, full)
        cmax = err.max(axis=1)
        out[name] = {"device_s": rep["solver_seconds"], "wall_s": wall, "solver_points": rep["solver_points"],
                     "points_per_curve": rep["points_per_curve"], "rounds": rep["rounds"],
                     "traced_items": rep["traced_items"], "work_vs_full": full_items / max(rep["traced_items"], 1),
                     "rough_points": rep["rough_points"], "gaps_closed_by_surrogate": rep["gaps_skipped"],
                     "curves_surrogate_trusted": rep["surrogate_trusted"], "bisect_checks_passed": rep["bisect_passed"],
                     "points_by_reason": rep["points_by_reason"], "work_share_by_reason": rep["work_share_by_reason"],
                     "accepted": rep["accepted"], "true_max_error": float(cmax.max()),
                     "curves_within_tolerance": int((cmax <= 1e-3).sum()),
                     "over_tolerance": [{"material": int(i), "error": float(cmax[i]), "solver_jitter": float(jitter[i])}
                                        for i in np.flatnonzero(cmax > 1e-3)],
                     "speedup_vs_full_fp64_wall": full_wall / wall}
        if sg is not None:
            # the network's cost: the solver runs that make its training data and its training
            build = out["surrogate"]["wall_s"]
            out[name]["wall_s_with_surrogate_build"] = wall + build
            out[name]["speedup_vs_full_fp64_wall_with_surrogate_build"] = full_wall / (wall + build)
    sur.close()
    g, p = out["adaptive_mixed_surrogate"], out["adaptive_mixed"]
    # what the network buys: solver work and wall time against the same method without it,
    # the second with the solver runs and training that build the network counted
    out["surrogate_gain"] = {"solver_work": p["traced_items"] / g["traced_items"], "wall": p["wall_s"] / g["wall_s"],
                             "wall_with_build": p["wall_s"] / g["wall_s_with_surrogate_build"]}
    return out


if __name__ == "__main__":
    sys.exit(main())
