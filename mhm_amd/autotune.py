"""Autotuner for the time loop: picks steps per launch (S), the cell-kernel
block size, the routing schedule (tiles, grid diagonals or workgroup-resident
trees) and
hipGraph capture by timing short trials on the model itself.  The device
states are snapshotted before and restored after every trial, so tuning
leaves the simulation where it was.
"""
import itertools
import time

from . import _native as N


def candidates(model, S_values=(8, 24, 72, 168), blocks=(64, 128, 256), graphs=(False, True),
               windows=(8, 64)):
    """Fused-path settings worth trying for this model (the routing window
    only matters for the tiled routing)."""
    route_modes = [3, 1, 2] if _small_trees(model) else [3, 1]
    for S, b, rm, g in itertools.product(S_values, blocks, route_modes, graphs):
        for w in (windows if rm == 3 else (0,)):
            yield {"mode": N.RUN_FUSED, "steps_per_launch": S, "block": b, "route_mode": rm,
                   "use_graph": g, "route_window": w}


def _small_trees(model):
    starts = model.cell_start
    return max(int(starts[i + 1] - starts[i]) for i in range(len(starts) - 1)) <= 65536


def autotune(model, steps, trial_steps=None, cands=None, repeats=2, log=None):
    """Returns (best settings, {settings tuple: seconds}) for runs of `steps`."""
    n = min(len(steps), trial_steps or len(steps))
    trial = steps[:n]
    snap = model.snapshot()
    timings = {}
    best, best_t = None, float("inf")
    for c in (list(cands) if cands is not None else list(candidates(model))):
        model.restore(snap)
        model.run(trial, want_gauges=False, **c)  # warm-up (and graph capture)
        ts = []
        for _ in range(repeats):
            model.restore(snap)
            model.sync()
            t0 = time.perf_counter()
            model.run(trial, want_gauges=False, **c)
            ts.append(time.perf_counter() - t0)
        t = min(ts)
        key = tuple(sorted(c.items()))
        timings[key] = t
        if log:
            log(f"autotune {c}: {t * 1e3:.2f} ms")
        if t < best_t:
            best, best_t = c, t
    model.restore(snap)
    return best, timings
