#!/usr/bin/env python3
"""Throughput of the FP64 lifetime / sigma path on one MI355X.

A step is one impurity sample of the production sweep: the field is drawn
and binned on the GPU, then the incoherent and the coherent pass each run the
40-field sweep with the 3 temperatures (SimulationRunner.cpp:88-98 order).

Prints one JSON line: lifetimes/s (incoherent lifetimes, the coherent pass is
run but not counted), the sweep wall time, the roofline of the trace kernel
(algorithmic FP64 work per launch from the kernel's own work counters over its
HIP-event duration, against the 78.6 TFLOP/s FP64 vector peak), and the CPU
oracle timed on a bounded sample of the same workload on this host.
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
    "cross_exact": 121,       # GetCrossPoints + 2 GetAngle + 2 AngleInRange + times (every crossing disk)
}
SIGMA_WORK = {
    "traced_particles": 70,   # phi, sincos, Simpson weight, lifetime sum
    "sigma_evals": 78,        # truncation, exp, sincos, GetSigma, two weighted terms, compensated adds
}


def algorithmic_flops(m, model):
    return sum(model[k] * m[k] for k in model)


def cpu_baseline(w, seconds_target=15.0):
    """The oracle (libm build, 1 thread) on the first sample's incoherent pass at the lowest and highest field."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from elasticscattering_amd.workloads import region

    lib = O.load("libm")
    if lib is None:
        return None
    rs, ext = region(w)
    og = O.OracleGrid.generate(lib, w["seeds"][0], rs, ext, 0.0, w["impurity_radius"], 15, w["cells_per_row"],
                               n=w["impurity_count"])
    prm = O.sim_params(w["positions_per_row"], w["particles_per_quadrant"], rs, ext, og.cell_size, w["tau"],
                       w["alpha"], w["clockwise"], w["particle_speed"], w["incoherent_tau"], nthreads=1)
    fields = [w["fields"][0], w["fields"][-1]]
    t0 = time.perf_counter()
    done = 0
    for B in fields:
        O.compute_sigmas(lib, og, prm, 0, B, w["temperatures"])
        done += 1
        if time.perf_counter() - t0 > seconds_target:
            break
    dt = time.perf_counter() - t0
    P = w["positions_per_row"] ** 2 * 4 * w["particles_per_quadrant"]
    return {
        "value": done * P / dt,
        "unit": "lifetimes/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"CPU oracle (restated reference CPU path, glibc libm, -O2) on sample seed {w['seeds'][0]}, "
                   f"incoherent pass, B in {fields[:done]} T, {len(w['temperatures'])} temperatures each: "
                   f"{done * P} lifetimes + sigma stage in {dt:.1f} s"),
        "seconds": dt,
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
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="production")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 4 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by sample and field, "
              "see DESIGN.md", file=sys.stderr)
        return 2

    import elasticscattering_amd as es
    from elasticscattering_amd.workloads import PRODUCTION_FP64, SMALL_FP64, region, lifetimes_per_pass

    w = PRODUCTION_FP64 if args.workload == "production" else SMALL_FP64
    rs, ext = region(w)
    ppr, ppq = w["positions_per_row"], w["particles_per_quadrant"]
    fields, temps = w["fields"], w["temperatures"]
    P = lifetimes_per_pass(w)

    sim = es.Simulation(ppr, ppq)

    def step(s, stats):
        seed = w["seeds"][s % len(w["seeds"])]
        sim.generate_impurities(seed, rs, ext, w["impurity_radius"], count=w["impurity_count"],
                                cells_per_row=w["cells_per_row"])
        for coherent in (0, 1):
            sim.init_sample(w["tau"], w["alpha"], w["clockwise"], w["particle_speed"], coherent, w["incoherent_tau"])
            sig, ms = sim.compute_sweep(fields, temps)
            if stats is not None and coherent == 0:
                stats["launch_s"].append(sum(m["time_elapsed_lifetimes"] for m in ms))
                stats["flops"].append(sum(algorithmic_flops(m, TRACE_WORK) for m in ms))
                stats["sigma_flops"].append(sum(algorithmic_flops(m, SIGMA_WORK) for m in ms))
                stats["sigma_s"].append(sum(m["time_elapsed_temperatures"] for m in ms))
                stats["traced"] += sum(m["traced_particles"] for m in ms)
                stats["cross"] += sum(m["cross_calls"] for m in ms)
                stats["exact"] += sum(m["cross_exact"] for m in ms)
                stats["sigma"].append(sig)

    for s in range(args.warmup):
        step(s, None)
    stats = {"launch_s": [], "flops": [], "sigma_flops": [], "sigma_s": [], "traced": 0, "sigma": [], "cross": 0,
             "exact": 0}
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(s, stats)
    dt = time.perf_counter() - t0

    lifetimes = args.steps * len(fields) * P
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
            "phi_points": 4 * ppq, "fields": len(fields), "temperatures": temps, "samples_timed": args.steps,
            "passes": "incoherent (counted) + coherent (run, alpha=pi/4 gives zero weight)",
        },
        "sweep_wall_s": dt * len(w["seeds"]) / args.steps,
        "traced_lifetimes_per_s": stats["traced"] / dt,
        "exact_cross_solves_fraction": stats["exact"] / max(stats["cross"], 1),
        "roofline": {
            "bound": "fp64-valu",
            "kernel": "k_trace (CreateParticle + TraceOrbit, one launch per sample and pass, all fields)",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": traffic,
            "avg_launch_ms": avg_launch * 1e3,
            "algorithmic_flops_per_launch": avg_flops,
            "sigma_stage_tflops": sigma_tf,
        },
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(w)
        if cb is not None:
            line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
    print(json.dumps(line), flush=True)
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
