"""Benchmark of the mHM + mRM hot path on one MI355X.

Headline workload: one synthetic basin of 1024 x 1024 L1 cells (L0 = L1 =
L11, 1 km, D8 network on an fBm DEM, ~1400 routing levels), 3 soil
horizons, hourly time step for 90 days (2160 steps), Hargreaves PET on daily
forcing, Muskingum routing.  One bench step = the full 90-day time loop
(cells and routing) over that basin with all inputs resident in HBM.

Besides the headline (fused time loop) the same JSON line carries
  plain       the one-step-per-launch path on the same basin (every flux
              written every step), with its HBM roofline
  ensemble    256 parameter sets x 4096 cells, daily x 3650, settings
              chosen by the autotuner
  multi_basin 64 basins of 500..200k cells packed in one context, hourly x 720
  serving     Poisson request stream (1k..100k cells, 30..365 daily steps,
              1..64 sets), continuous batching, p50/p99 latency
  cpu_baseline  the oracle (C restatement, OpenMP) on a bounded sample of
              the headline workload, on this host's cores

`--summarize-pmc FETCH_DIR WRITE_DIR` turns two rocprofv3 --pmc runs
(FETCH_SIZE, then WRITE_SIZE, each over `bench.py --pmc-workload`) into
profiles/traffic.json: HBM bytes per launch of each kernel, scaled by what
the same counters read for a plain copy of known size.
"""
import argparse
import csv
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
PROF_STRIDE = 16           # one launch in 16 of each kernel is bracketed by events
COPY_CALIB_BYTES = 1 << 30  # 1 GiB read + 1 GiB written by the calibration copy


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_unit(kernel, steps_per_launch, steps_per_forcing, window=16, sets_sharing=1):
    """Algorithmic HBM bytes per work unit (cell-step or node-step).
    steps_per_forcing: model steps per forcing column (iMeteoTS changes every
    24 steps for daily forcing, hourly and daily model steps alike);
    sets_sharing: parameter sets reading one forcing column."""
    if kernel == "cell_step":
        # forcing 32 + PET params 24 + mHM params 232 + states 64 r/64 w + 23 fluxes 184 w
        return 600.0
    if kernel == "cell_fused":
        # forcing once per forcing column, shared by the sets (32/(F n_sets)), qOUT
        # to the ring every step (8); once per launch: params 256 + states 64 +
        # ring lag 1 read, states 64 + 23 fluxes 184 written (569/S)
        return 32.0 / (steps_per_forcing * sets_sharing) + 8.0 + 569.0 / steps_per_launch
    if kernel == "route_tile":
        # node-step: qOUT from the ring 8; once per window of B steps: slot metadata
        # 16, C1 + C2 16, qTIN read + write 16, qTR read + write 16
        return 8.0 + 64.0 / window
    # routing node-step: qOUT from the ring 8, C1 + C2 16, own qTIN read + write 16,
    # own qTR(IT1) read 8 + qTR(IT2) write 8, upstream qTR read 8 (one per edge),
    # node metadata (child index + packed depth/flags) 8
    return 72.0


def load_json(path):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def kernel_table(prof, steps_per_launch, steps_per_day, window=8, sets_sharing=1):
    detail = {}
    for k, (kms, kn, ku) in prof.items():
        if kn:
            b = bytes_per_unit(k, steps_per_launch, steps_per_day, window, sets_sharing)
            # the timed launches only (one in PROF_STRIDE of each kernel)
            detail[k] = {"ms": round(kms, 3), "timed_launches": kn, "units": ku,
                         "avg_us": round(kms * 1e3 / kn, 3),
                         "bytes_per_unit": round(b, 3),
                         "GBs": round(ku * b / (kms * 1e-3) / 1e9, 1)}
    return detail


def hbm_roofline(name, stats, steps_per_launch, steps_per_day, traffic):
    ms, launches, units = stats
    if launches == 0 or ms <= 0:
        return None
    bpu = bytes_per_unit(name, steps_per_launch, steps_per_day)
    achieved = units * bpu / (ms * 1e-3) / 1e9
    t = traffic.get(name)
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if t is None else t.get("bytes_per_launch"),
            "avg_launch_us": round(ms * 1e3 / launches, 3),
            "algorithmic_bytes_per_launch": round(units * bpu / launches, 1)}


N_CU, CLOCK_GHZ = 256, 2.4


def fp64_roofline(name, stats, flops_per_unit, valu_per_cell_step=None):
    """FP64 VALU roofline: flops per cell-step (PMC count of the FP64 VALU
    instructions of this kernel on this workload, 64 lanes each, FMA = 2)
    x cell-steps per launch / measured launch time.  With the PMC count of
    all VALU wave instructions per cell-step, also the share of the CUs'
    VALU issue rate (one wave instruction per CU and cycle) the kernel uses."""
    ms, launches, units = stats
    if launches == 0 or ms <= 0 or not flops_per_unit:
        return None
    achieved = units * flops_per_unit / (ms * 1e-3) / 1e12
    out = {"bound": "fp64-valu", "kernel": name, "achieved": round(achieved, 2),
           "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
           "flops_per_cell_step": round(flops_per_unit, 1),
           "avg_launch_us": round(ms * 1e3 / launches, 3)}
    if valu_per_cell_step:
        rate = units * valu_per_cell_step / (ms * 1e-3)  # wave instructions per second
        out["valu_issue"] = {"wave_insts_per_cell_step": round(valu_per_cell_step, 4),
                             "frac": round(rate / (N_CU * CLOCK_GHZ * 1e9), 4)}
    return out


def host_cores():
    """The host cores this process may use, measured: logical CPUs in the
    affinity mask, the physical cores among them (lscpu), the cgroup CPU
    quota, and OMP_NUM_THREADS when the environment sets one; the oracle runs
    one thread per usable physical core (the smallest of these limits)."""
    import math
    import subprocess
    aff = sorted(os.sched_getaffinity(0))
    info = {"nproc": len(aff)}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                info["lscpu_model"] = line.split(":", 1)[1].strip()
            if line.startswith("Thread(s) per core:"):
                info["threads_per_core"] = int(line.split(":", 1)[1])
        par = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        cores = set()
        for line in par.splitlines():
            if line.startswith("#") or not line.strip():
                continue
            cpu, core, sock = (int(v) if v else -1 for v in line.split(",")[:3])
            if cpu in aff:
                cores.add((sock, core))
        info["physical_cores"] = len(cores) if cores else len(aff)
    except (OSError, ValueError, subprocess.SubprocessError):
        info["physical_cores"] = len(aff)
    limits = {"physical_cores": info["physical_cores"]}
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            info["cgroup_cpus"] = int(quota) / int(period)
            limits["cgroup_cpus"] = max(1, math.floor(info["cgroup_cpus"]))
    except (OSError, ValueError):
        pass
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        info["OMP_NUM_THREADS"] = int(os.environ["OMP_NUM_THREADS"])
        limits["OMP_NUM_THREADS"] = info["OMP_NUM_THREADS"]
    info["threads"] = min(limits.values())
    info["bound_by"] = min(limits, key=limits.get)
    info["OMP_PROC_BIND"] = os.environ.get("OMP_PROC_BIND")
    info["OMP_PLACES"] = os.environ.get("OMP_PLACES")
    return info


def cpu_baseline(dom, gam, n_days, sample_steps, threads, hosts=None, label="same"):
    """The oracle (C restatement, OpenMP over cells, serial routing) on the
    first `sample_steps` hourly steps of a basin (step 1 untimed)."""
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
    out = {"value": n / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
           "sample": f"oracle, OpenMP {threads} threads: hourly steps 2..{sample_steps} of the "
                     f"{label} {dom.nrows}x{dom.ncols} basin, meteo + mHM + serial mRM routing, "
                     f"{dt:.1f} s"}
    if hosts:
        out["host"] = hosts
    return out


def timed_runs(model, steps, k, **kw):
    model.sync()
    t0 = time.perf_counter()
    for _ in range(k):
        model.run(steps, **kw)
    model.sync()
    return time.perf_counter() - t0


def bench_ensemble(N, workloads, autotune, traffic):
    t0 = time.perf_counter()
    em, esteps, edom, _ = workloads.plain_domain(64, 64, 3650, timestep=24, seed=3, n_sets=256,
                                                 device_params=True)
    setup = time.perf_counter() - t0
    best, timings = autotune.autotune(em, esteps, trial_steps=365,
                                      cands=list(autotune.candidates(em, S_values=(73, 365, 730))),
                                      repeats=2)
    snap = em.snapshot()
    # a few launches per run: time them all
    em.set_profiling(True, 1)
    em.run(esteps[:730], want_gauges=False, **best)  # graph capture, if chosen
    em.restore(snap)
    em.set_profiling(True, 1)
    te = timed_runs(em, esteps, 1, **best)
    eprof = em.profile()
    em.close()
    units = 256 * edom.n_cells * 3650
    S = best["steps_per_launch"]
    # daily steps on daily forcing: iMeteoTS = ceiling(tt / 24) moves every 24 steps
    return {"value": units / te, "unit": "cell-steps/s",
            "workload": "256 parameter sets x 4096 cells (64x64), daily x 3650, shared forcing, "
                        "each set's parameter rows built on the device from its gamma vector",
            "settings": best, "autotune_trials": len(timings), "setup_s": round(setup, 1),
            "kernels": kernel_table(eprof, S, 24, best.get("route_window") or 8, 256)}


def bench_per_step(model, steps, dom, n, tile_height):
    """The drop-in as the reference's time loop calls it: per step one
    mhmhip_mhm_step (meteo + mHM(), every flux written) and one
    mhmhip_mrm_step (mRM_routing) handing the gauge discharge back to the
    host, on the headline basin; n steps timed after a warm-up step."""
    model.mhm_step(steps[0])
    model.mrm_step(steps[0])
    model.sync()
    t0 = time.perf_counter()
    for k in range(1, n + 1):
        model.mhm_step(steps[k])
        model.mrm_step(steps[k])
    dt = time.perf_counter() - t0
    model.set_profiling(True, 1)
    model.mhm_step(steps[n + 1])
    model.mrm_step(steps[n + 1])
    prof = model.profile()
    model.set_profiling(False)
    plan_launches = prof["route_tile"][1]
    return {"value": dom.n_cells * n / dt, "unit": "cell-steps/s", "ms_per_step": dt / n * 1e3,
            "workload": f"same basin, {n} hourly steps, each one mhm_step + one mrm_step "
                        "(gauges to the host every step), as mhm_interface_run_do_time_step "
                        f"would call them; routing tiles of {tile_height} levels",
            "one_step": {"cell_step_us": round(prof["cell_step"][0] * 1e3, 1),
                         "route_tile_launches": plan_launches,
                         "route_tile_us": round(prof["route_tile"][0] * 1e3, 1)}}


def bench_year_basin(N, workloads, autotune):
    """64 x 64 cells (4096), 3 horizons, hourly x 8760 (one year): the
    fused time loop with autotuned settings, the plain path, and the oracle
    over the whole year on the host (all usable cores, then one thread)."""
    n_steps = 8760
    t0 = time.perf_counter()
    model, steps, dom, gam = workloads.plain_domain(64, 64, n_steps, seed=2)
    setup = time.perf_counter() - t0
    cands = [dict(c) for c in autotune.candidates(model, S_values=(144, 730, 2190),
                                                  blocks=(64, 128), graphs=(True,),
                                                  windows=(64,))]
    best, timings = autotune.autotune(model, steps, trial_steps=2190, cands=cands, repeats=2)
    snap = model.snapshot()
    model.run(steps, want_gauges=True, **best)  # warm-up (graph capture)
    model.restore(snap)
    tf = timed_runs(model, steps, 3, want_gauges=True, **best) / 3
    model.restore(snap)
    model.set_profiling(True, 1)
    model.run(steps, want_gauges=True, **best)
    prof = model.profile()
    model.set_profiling(False)
    model.restore(snap)
    pkw = dict(mode=N.RUN_PLAIN, want_gauges=True, route_mode=best["route_mode"],
               route_window=best.get("route_window", 0))
    model.run(steps[:48], **pkw)
    model.restore(snap)
    tp = timed_runs(model, steps, 1, **pkw)
    model.close()
    units = dom.n_cells * n_steps
    out = {"value": units / tf, "ms_per_run": round(tf * 1e3, 3), "unit": "cell-steps/s",
           "workload": "one basin 64x64 L1 cells (4096), 3 horizons, hourly x 8760 (one year), "
                       f"D8 routing (max depth {dom.dmax}), fused time loop",
           "settings": best, "autotune_trials": len(timings), "setup_s": round(setup, 1),
           "kernels": kernel_table(prof, best["steps_per_launch"], 24,
                                   best.get("route_window") or 8),
           "plain": {"value": units / tp, "ms_per_run": round(tp * 1e3, 3),
                     "unit": "cell-steps/s",
                     "workload": "same basin and steps, one launch per step, every flux written "
                                 "every step"}}
    return out, dom, gam


def bench_multi_basin(N, workloads, S, window):
    t0 = time.perf_counter()
    mm, msteps, mdoms, _ = workloads.multi_basin(64, 500, 200000, 720, seed=7)
    setup = time.perf_counter() - t0
    cells = sum(d.n_cells for d in mdoms)
    kw = dict(mode=N.RUN_FUSED, steps_per_launch=S, want_gauges=True, use_graph=True,
              route_window=window)
    mm.run(msteps, **kw)
    mm.set_profiling(True, PROF_STRIDE)
    mm.run(msteps, **kw)
    mm.set_profiling(True, PROF_STRIDE)
    tm = timed_runs(mm, msteps, 1, **kw)
    mprof = mm.profile()
    mm.close()
    return {"value": cells * len(msteps) / tm, "unit": "cell-steps/s",
            "workload": f"64 basins, log-uniform 500..200k cells ({cells} in all, largest "
                        f"{max(d.n_cells for d in mdoms)}, max depth {max(d.dmax for d in mdoms)}),"
                        f" own DEM/network/forcing/parameters each, hourly x 720, fused S={S},"
                        f" tiled routing window {window}",
            "setup_s": round(setup, 1), "kernels": kernel_table(mprof, S, 24, window)}


def bench_serving(serve, n_req, qps, pool, seed=11):
    t0 = time.perf_counter()
    reqs = serve.poisson_requests(n_req, qps=qps, seed=seed, pool=pool)
    gen_s = time.perf_counter() - t0
    # warm the code paths (module loads, first context) outside the stream
    serve.run_alone(serve.poisson_requests(1, qps=1.0, seed=99, cells=(1000, 1000),
                                           horizon=(30, 30), sets=(1, 1))[0])
    srv = serve.Server(hbm_budget_bytes=64 << 30, max_slots=8, chunk_steps=73)
    res = srv.serve(reqs)
    s = serve.summarize(res, reqs, srv.wall, srv.timings)
    s.update({"qps": qps, "peak_slots": srv.peak_slots,
              "peak_admitted_GiB": round(srv.peak_bytes / 2**30, 2),
              "basin_pool": pool, "generate_s": round(gen_s, 1),
              "workload": f"{n_req} Poisson arrivals at {qps} req/s: log-uniform 1k..100k cells "
                          f"(basins drawn from {pool} pre-built ones), 30..365 daily steps, "
                          "1..64 parameter sets; budget 64 GiB, 8 slots"})
    return s


# -- PMC traffic ------------------------------------------------------------------

def pmc_workload(args):
    """Short run for counter collection: the headline kernels on the same
    basin (2 plain steps, then 6 fused launches routed in the headline's
    windows), then a plain device copy of COPY_CALIB_BYTES."""
    import torch
    from mhm_amd import _native as N, workloads
    n_steps = 2 + 6 * args.steps_per_launch  # 2 plain steps, then whole fused launches
    model, steps, dom, _ = workloads.plain_domain(args.grid, args.grid, n_steps, seed=1)
    model.run(steps[:2], mode=N.RUN_PLAIN)
    model.run(steps[2:], mode=N.RUN_FUSED, steps_per_launch=args.steps_per_launch,
              route_window=args.route_window)
    model.sync()
    model.close()
    plan, _, _ = N.plan_network(dom)
    n_fused = n_steps - 2
    meta_dir = os.path.dirname(os.path.abspath(args.pmc_meta))
    os.makedirs(meta_dir, exist_ok=True)
    with open(args.pmc_meta, "w") as f:
        json.dump({"cells": dom.n_cells, "steps_per_launch": args.steps_per_launch,
                   "fused_launches": -(-n_fused // args.steps_per_launch), "plain_launches": 2,
                   "route_window": args.route_window,
                   "route_tile_fused_launches": -(-n_fused // args.route_window) + plan.tau_max,
                   "route_node_steps": dom.n_cells * n_fused}, f)
    a = torch.empty(COPY_CALIB_BYTES // 8, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"pmc workload done: cells={dom.n_cells}")


def _pmc_rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    rows = []
    for f in files:
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _per_kernel(rows, counter):
    """{kernel short name: [value per dispatch, in dispatch order]} for one counter."""
    out = {}
    for r in sorted(rows, key=lambda r: int(r.get("Dispatch_Id", 0))):
        if r.get("Counter_Name") != counter:
            continue
        name = r["Kernel_Name"]
        short = ("copy" if ("copy" in name.lower() or "elementwise" in name.lower())
                 and "mhm" not in name.lower() else None)
        for k in ("cell_fused", "cell_step", "route_diag", "route_wg", "route_tile"):
            if k in name:
                short = k
        if short:
            out.setdefault(short, []).append(float(r["Counter_Value"]))
    return out


def summarize_pmc(fetch_dir, write_dir, out_path, meta_path):
    fetch = _per_kernel(_pmc_rows(fetch_dir), "FETCH_SIZE")
    write = _per_kernel(_pmc_rows(write_dir), "WRITE_SIZE")
    if "copy" not in fetch or "copy" not in write:
        raise SystemExit("calibration copy not found in the counter runs")
    kib = COPY_CALIB_BYTES / 1024.0
    f_scale = kib / max(fetch["copy"])
    w_scale = kib / max(write["copy"])
    res = {"_calibration": {"copy_bytes": COPY_CALIB_BYTES, "fetch_kib": max(fetch["copy"]),
                            "write_kib": max(write["copy"]), "fetch_scale": f_scale,
                            "write_scale": w_scale}}
    for k in fetch:
        if k == "copy" or k not in write:
            continue
        fv = sum(fetch[k]) / len(fetch[k]) * f_scale * 1024.0
        wv = sum(write[k]) / len(write[k]) * w_scale * 1024.0
        res[k] = {"bytes_per_launch": round(fv + wv), "fetch_bytes": round(fv),
                  "write_bytes": round(wv), "dispatches": len(fetch[k])}
    if os.path.exists(meta_path) and "route_tile" in fetch and "route_tile" in write:
        # the fused run's tile launches are the last ones: bytes per node-step
        meta = load_json(meta_path)
        n = meta["route_tile_fused_launches"]
        fv = sum(fetch["route_tile"][-n:]) * f_scale * 1024.0
        wv = sum(write["route_tile"][-n:]) * w_scale * 1024.0
        res["route_tile_fused"] = {"bytes_per_node_step": (fv + wv) / meta["route_node_steps"],
                                   "fetch_bytes": round(fv), "write_bytes": round(wv),
                                   "dispatches": n, "node_steps": meta["route_node_steps"],
                                   "window": meta["route_window"]}
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


FP64_COUNTERS = {"SQ_INSTS_VALU_ADD_F64": 64, "SQ_INSTS_VALU_MUL_F64": 64,
                 "SQ_INSTS_VALU_FMA_F64": 128, "SQ_INSTS_VALU_TRANS_F64": 64}


def summarize_flops(pmc_dir, meta_path, out_path):
    """FP64 lane-ops per cell-step of the cell kernels from one --pmc run of
    the FP64 VALU instruction counters (per wave instruction: 64 lanes, FMA 2)."""
    rows = _pmc_rows(pmc_dir)
    with open(meta_path) as f:
        meta = json.load(f)
    units = {"cell_fused": meta["cells"] * meta["steps_per_launch"], "cell_step": meta["cells"]}
    res = {}
    for k, u in units.items():
        tot, parts = 0.0, {}
        for c, w in FP64_COUNTERS.items():
            v = _per_kernel(rows, c).get(k)
            if not v:
                continue
            parts[c] = sum(v) / len(v)
            tot += w * parts[c]
        if parts:
            res[k] = {"flops_per_unit": tot / u, "unit": "cell-step",
                      "wave_insts_per_dispatch": parts}
            valu = _per_kernel(rows, "SQ_INSTS_VALU").get(k)
            if valu:
                res[k]["valu_wave_insts_per_unit"] = sum(valu) / len(valu) / u
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--days", type=int, default=90)
    ap.add_argument("--steps-per-launch", type=int, default=144)
    ap.add_argument("--route-window", type=int, default=64,
                    help="routing window of the headline loop (ring memory grows with it)")
    ap.add_argument("--no-graph", action="store_true", help="headline loop without hipGraph")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-block", type=int, default=0, choices=(0, 64, 128, 256),
                    help="threads per workgroup of the plain path's cell kernel (0: 256)")
    ap.add_argument("--no-year", action="store_true")
    ap.add_argument("--no-per-step", action="store_true")
    ap.add_argument("--tile-height", type=int, default=0,
                    help="levels per routing tile of the headline context (0: 16)")
    ap.add_argument("--per-step-steps", type=int, default=48)
    ap.add_argument("--per-step-tile-height", type=int, default=32,
                    help="levels per routing tile of the per-step context (16: the headline's)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-multi", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=241)
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--serve-pool", type=int, default=32,
                    help="pre-built basins the serving requests draw from")
    ap.add_argument("--serve-qps", type=float, default=8.0)
    ap.add_argument("--pmc-workload", action="store_true")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--summarize-flops", metavar="PMC_DIR")
    ap.add_argument("--pmc-meta", default="pmc_meta.json",
                    help="workload description written by --pmc-workload and read by the "
                         "--summarize-* steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (every basin, "
              "parameter set and request here fits one MI355X; see the sharding design)",
              file=sys.stderr)
        sys.exit(2)
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, os.path.join(ROOT, "profiles", "traffic.json"),
                      args.pmc_meta)
        return
    if args.summarize_flops:
        summarize_flops(args.summarize_flops, args.pmc_meta,
                        os.path.join(ROOT, "profiles", "fp64_flops.json"))
        return
    if args.pmc_workload:
        pmc_workload(args)
        return

    # the oracle's OpenMP threads: one per core, packed (read when its
    # runtime starts, so set before anything loads it)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    hosts = host_cores()
    from mhm_amd import _native as N, autotune, serve, workloads

    traffic = load_json(os.path.join(ROOT, "profiles", "traffic.json"))
    flops = load_json(os.path.join(ROOT, "profiles", "fp64_flops.json"))
    n_steps = args.days * 24
    t0 = time.perf_counter()
    model, steps, dom, gam = workloads.plain_domain(args.grid, args.grid, n_steps, seed=1,
                                                    tile_height=args.tile_height)
    log(f"setup {time.perf_counter() - t0:.1f}s: cells={dom.n_cells} dmax={dom.dmax} "
        f"steps={n_steps} device_bytes={model.device_bytes / 2**30:.2f} GiB")
    S = args.steps_per_launch
    kw = dict(mode=N.RUN_FUSED, steps_per_launch=S, want_gauges=True, use_graph=not args.no_graph,
              route_window=args.route_window)
    # profiling on before the warm-up so a captured graph holds its event nodes
    model.set_profiling(True, PROF_STRIDE)
    for _ in range(args.warmup):
        model.run(steps, **kw)
    model.set_profiling(True, PROF_STRIDE)  # counters from zero for the timed region
    model.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        model.run(steps, **kw)
        log(f"bench step {k + 1}/{args.steps} done")
    model.sync()
    elapsed = time.perf_counter() - t0
    device_gib = model.device_bytes / 2**30
    prof = model.profile()
    model.set_profiling(False)
    value = float(dom.n_cells) * n_steps * args.steps / elapsed
    kernels = kernel_table(prof, S, 24, args.route_window)
    dominant = max(prof.items(), key=lambda kv: kv[1][0])[0]
    hbm = hbm_roofline(dominant, prof[dominant], S, 24, traffic)
    fpu = flops.get(dominant, {}).get("flops_per_unit")
    fp64 = fp64_roofline(dominant, prof[dominant], fpu,
                         flops.get(dominant, {}).get("valu_wave_insts_per_unit"))
    roof = fp64 if (fp64 and dominant == "cell_fused") else hbm
    log(f"fused: {value:.4g} cell-steps/s  kernels={json.dumps(kernels)}")

    extra = {"roofline_hbm": hbm}
    if not args.no_plain:
        pkw = dict(mode=N.RUN_PLAIN, want_gauges=True, route_window=args.route_window)
        if args.plain_block:
            pkw["block"] = args.plain_block
        model.run(steps[:48], **pkw)
        model.set_profiling(True, PROF_STRIDE)
        tp = timed_runs(model, steps, 1, **pkw)
        pprof = model.profile()
        model.set_profiling(False)
        extra["plain"] = {"value": dom.n_cells * n_steps / tp, "unit": "cell-steps/s",
                          "workload": "same basin and steps, one launch per step, every flux "
                                      f"written every step, tiled routing window {args.route_window}",
                          "cell_step_roofline": hbm_roofline("cell_step", pprof["cell_step"], 1,
                                                             24, traffic),
                          "kernels": kernel_table(pprof, 1, 24, args.route_window)}
        log(f"plain: {extra['plain']['value']:.4g} cell-steps/s")
    model.close()
    if not args.no_per_step:
        # the drop-in's own context: tiles of per_step_tile_height levels
        pm, psteps = workloads.model_on(dom, args.per_step_steps + 2, gam,
                                        tile_height=args.per_step_tile_height)
        extra["per_step"] = bench_per_step(pm, psteps, dom, args.per_step_steps,
                                           args.per_step_tile_height)
        pm.close()
        log(f"per step: {extra['per_step']['value']:.4g} cell-steps/s, "
            f"{extra['per_step']['ms_per_step']:.3f} ms per step")

    year = None
    if not args.no_year:
        extra["year_basin"], ydom, ygam = bench_year_basin(N, workloads, autotune)
        year = (ydom, ygam)
        log(f"year basin: {extra['year_basin']['value']:.4g} "
            f"plain {extra['year_basin']['plain']['value']:.4g} {extra['year_basin']['settings']}")
    if not args.no_ensemble:
        extra["ensemble"] = bench_ensemble(N, workloads, autotune, traffic)
        log(f"ensemble: {extra['ensemble']['value']:.4g} {extra['ensemble']['settings']}")
    if not args.no_multi:
        extra["multi_basin"] = bench_multi_basin(N, workloads, S, args.route_window)
        log(f"multi_basin: {extra['multi_basin']['value']:.4g}")
    if not args.no_serve:
        extra["serving"] = bench_serving(serve, args.serve_requests, args.serve_qps,
                                         args.serve_pool)
        log(f"serving: {json.dumps(extra['serving'])}")

    # The oracle last: its OpenMP runtime, started with OMP_PROC_BIND, pins
    # the calling thread (and every thread started from it later) to one core.
    cpu = None
    if not args.no_cpu:
        aff = os.sched_getaffinity(0)
        cpu = cpu_baseline(dom, gam, dom.n_days, args.cpu_steps, hosts["threads"], hosts)
        log(f"cpu: {cpu['value']:.4g} cell-steps/s ({json.dumps(hosts)})")
        one = cpu_baseline(dom, gam, dom.n_days, 25, 1)  # one simulated day after step 1
        extra["cpu_baseline_1thread"] = one
        log(f"cpu 1 thread: {one['value']:.4g} cell-steps/s")
        if year is not None:
            ydom, ygam = year
            yb = extra["year_basin"]
            yb["cpu_baseline"] = cpu_baseline(ydom, ygam, ydom.n_days, 8760, hosts["threads"],
                                              hosts, label="year")
            yb["cpu_baseline_1thread"] = cpu_baseline(ydom, ygam, ydom.n_days, 8760, 1,
                                                      label="year")
            log(f"cpu year basin: {yb['cpu_baseline']['value']:.4g}, 1 thread "
                f"{yb['cpu_baseline_1thread']['value']:.4g} cell-steps/s")
        os.sched_setaffinity(0, aff)

    out = {
        "metric": "cell-steps/s FP64 plain vs batched; p50/p99 request latency at fixed QPS",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single basin {args.grid}x{args.grid} L1 cells, 3 horizons, "
                               f"hourly x {n_steps} steps, D8 routing (max depth {dom.dmax}), "
                               f"fused S={S}, tiled routing window {args.route_window}, "
                               f"{'eager' if args.no_graph else 'hipGraph'} time loop",
                   "device_GiB": round(device_gib, 2),
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
