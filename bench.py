"""Benchmark of the mHM + mRM hot path on one MI355X.

Workload: one synthetic basin of 1024 x 1024 L1 cells (L0 = L1 = L11,
1 km, D8 network on an fBm DEM, ~1400 routing levels), 3 soil horizons,
hourly time step for 90 days (2160 steps), Hargreaves PET on daily forcing,
Muskingum routing.  One bench step = the full 90-day time loop (cells and
routing) over that basin with all inputs resident in HBM.

Prints one JSON line: cell-steps/s of the time loop (plus plain-path and
parameter-ensemble figures), the roofline of the dominant kernel measured
with HIP events inside the timed region, and the oracle's OpenMP CPU rate
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_unit(kernel, steps_per_launch, steps_per_day):
    """Algorithmic HBM bytes per work unit (cell-step or node-step)."""
    if kernel == "cell_step":
        # forcing 32 + PET params 24 + mHM params 232 + states 64 r/64 w + 23 fluxes 184 w
        return 600.0
    if kernel == "cell_fused":
        # forcing once per forcing day, total runoff out every step, params + states
        # in/out once per launch: 32/F + 8 + (128 + 256)/S
        return 32.0 / steps_per_day + 8.0 + 384.0 / steps_per_launch
    # routing node-step: qOUT source, C1, C2, own qTIN/qTR, upstream qTR, node
    # metadata read; qTIN, qTR written
    return 80.0


def roofline(prof, steps_per_launch, steps_per_day, traffic):
    best = max(prof.items(), key=lambda kv: kv[1][0])
    name, (ms, launches, units) = best
    if launches == 0 or ms <= 0:
        return None, {}
    bpu = bytes_per_unit(name, steps_per_launch, steps_per_day)
    achieved = units * bpu / (ms * 1e-3) / 1e9
    detail = {}
    for k, (kms, kn, ku) in prof.items():
        if kn:
            b = bytes_per_unit(k, steps_per_launch, steps_per_day)
            detail[k] = {"ms": round(kms, 3), "launches": kn, "units": ku,
                         "avg_us": round(kms * 1e3 / kn, 3),
                         "bytes_per_unit": round(b, 3),
                         "GBs": round(ku * b / (kms * 1e-3) / 1e9, 1)}
    t = traffic.get(name)
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": t, "avg_launch_us": round(ms * 1e3 / launches, 3),
            "algorithmic_bytes_per_launch": round(units * bpu / launches, 1)}, detail


def cpu_baseline(dom, gam, n_days, sample_steps, threads):
    """The oracle (C restatement, OpenMP over cells, serial routing) on the
    first `sample_steps` hourly steps of the same basin."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_sim import OracleSim
    from mhm_amd import calendar, parameters as PM, synth, workloads
    sim = OracleSim([dom], n_sets=1, timestep=1, n_forcing_days=n_days, nthreads=threads)
    sim.set_params(synth.cell_params(dom, gam[0]), 0)
    sim.set_routing_params(PM.routing_gamma(gam[0]))
    sim.default_states()
    sim.set_domain_forcing()
    recs = calendar.step_records(workloads.START, sample_steps, 1)
    sim.step(recs[0])  # first touch
    t0 = time.perf_counter()
    for r in recs[1:]:
        sim.step(r)
    dt = time.perf_counter() - t0
    n = dom.n_cells * (sample_steps - 1)
    return {"value": n / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle OpenMP ({threads} threads): steps 2..{sample_steps} of the same "
                      f"{dom.nrows}x{dom.ncols} basin, meteo + mHM + serial mRM routing, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--days", type=int, default=90)
    ap.add_argument("--steps-per-launch", type=int, default=24)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=25)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (every basin, "
              "parameter set and request here fits one MI355X; see the sharding design)",
              file=sys.stderr)
        sys.exit(2)

    from mhm_amd import _native as N, workloads

    n_steps = args.days * 24
    t0 = time.perf_counter()
    model, steps, dom, gam = workloads.plain_domain(args.grid, args.grid, n_steps, seed=1)
    log(f"setup {time.perf_counter() - t0:.1f}s: cells={dom.n_cells} dmax={dom.dmax} "
        f"steps={n_steps} device_bytes={model.device_bytes / 2**30:.2f} GiB")
    S = args.steps_per_launch
    for _ in range(args.warmup):
        model.run(steps, mode=N.RUN_FUSED, steps_per_launch=S, want_gauges=True)
    model.set_profiling(True)
    model.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        model.run(steps, mode=N.RUN_FUSED, steps_per_launch=S, want_gauges=True)
        log(f"bench step {k + 1}/{args.steps} done")
    model.sync()
    elapsed = time.perf_counter() - t0
    prof = model.profile()
    model.set_profiling(False)
    cell_steps = float(dom.n_cells) * n_steps * args.steps
    value = cell_steps / elapsed
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))
    roof, kernels = roofline(prof, S, 24, traffic)
    log(f"fused: {value:.4g} cell-steps/s  kernels={json.dumps(kernels)}")

    extra = {}
    if not args.no_plain:
        model.set_profiling(True)
        t0 = time.perf_counter()
        model.run(steps, mode=N.RUN_PLAIN, want_gauges=True)
        model.sync()
        tp = time.perf_counter() - t0
        pprof = model.profile()
        model.set_profiling(False)
        proof, pk = roofline({"cell_step": pprof["cell_step"]}, 1, 24, traffic)
        extra["plain"] = {"value": dom.n_cells * n_steps / tp, "unit": "cell-steps/s",
                          "cell_step_roofline": proof, "kernels": pk}
        log(f"plain: {extra['plain']['value']:.4g} cell-steps/s  {json.dumps(pk)}")
    model.close()

    if not args.no_ensemble:
        em, esteps, edom, _ = workloads.plain_domain(64, 64, 3650, timestep=24, seed=3, n_sets=256)
        em.run(esteps[:48], mode=N.RUN_FUSED, steps_per_launch=48)
        em.set_profiling(True)
        t0 = time.perf_counter()
        em.run(esteps, mode=N.RUN_FUSED, steps_per_launch=365)
        te = time.perf_counter() - t0
        eprof = em.profile()
        em.close()
        _, ek = roofline(eprof, 365, 1, traffic)
        extra["ensemble"] = {"value": 256 * edom.n_cells * 3650 / te, "unit": "cell-steps/s",
                             "workload": "256 parameter sets x 4096 cells, daily x 3650",
                             "kernels": ek}
        log(f"ensemble: {extra['ensemble']['value']:.4g} (cell,set)-steps/s {json.dumps(ek)}")

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "16"))
        cpu = cpu_baseline(dom, gam, dom.n_days, args.cpu_steps, threads)
        log(f"cpu: {cpu['value']:.4g} cell-steps/s")

    out = {
        "metric": "cell-steps/s FP64 plain vs batched; p50/p99 request latency at fixed QPS",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single basin {args.grid}x{args.grid} L1 cells, 3 horizons, "
                               f"hourly x {n_steps} steps, D8 routing (dmax {dom.dmax}), "
                               f"fused S={S}",
                   "cells": dom.n_cells, "time_steps": n_steps, "routing_levels": dom.dmax + 1,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
