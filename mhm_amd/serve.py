"""Serving loop: simulation requests of different basin sizes, horizons and
parameter-set counts arrive over time and are run with continuous batching.

* Admission by HBM footprint: a request is admitted when a slot is free and
  its estimated device footprint fits the budget next to the requests on the
  device; otherwise it waits (first-fit over the queue, so a small request
  is not held behind a large one).
* Each admitted request owns a device context (its states stay resident for
  its whole horizon) and a worker thread; the C-ABI releases the GIL and each
  context runs on its own HIP stream, so the work of requests in flight
  overlaps on the GPU.  Requests advance in chunks of `chunk_steps`, and a
  request that arrives while others run is admitted at once when memory
  allows -- not after the batch drains.
* No state or reduction is shared between requests, so a result is
  bitwise the same alone and in the mix.

Latency = completion - arrival; p50/p99 are reported beside cell-steps/s.
"""
import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import calendar, engine, parameters as PM, synth


@dataclass
class Request:
    rid: int
    domain: synth.Domain            # with forcing (synth.make_forcing)
    gammas: np.ndarray              # (n_sets, n_gamma)
    n_steps: int
    timestep: int = 24
    start: object = None
    arrival: float = 0.0            # seconds after the stream starts

    @property
    def cell_steps(self):
        return self.domain.n_cells * self.gammas.shape[0] * self.n_steps


@dataclass
class Result:
    rid: int
    gauges: np.ndarray              # (n_steps, n_sets, n_gauges)
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


TILE_WINDOW_MAX = 64  # longest routing window; boundary series are allocated at this length
TILE_HEIGHT = 16  # levels per routing tile; the runoff ring has this many rows past its length


STEP_BYTES = 104  # one step of the device step table (StepDev)
# forcing rows a context holds per PET case (pre, tavg and what the case reads)
FORCING_ROWS = {-1: 3, 0: 3, 1: 4, 2: 3, 3: 5}


def estimate_bytes(n_cells, n_sets, n_horizons, n_forcing_days, plan, chunk_steps, n_steps,
                   n_gauges=1, n_outlets=1, process_pet=1):
    """Device footprint of one context run in chunks of `chunk_steps` fused
    steps, allocation by allocation as api.hip makes them; `plan` is the
    network's mhmhip_plan (N.plan_network)."""
    sstride = (n_cells + 63) // 64 * 64
    ld = sstride * n_sets
    rows = (19 + 5 * n_horizons + 60) + (5 + n_horizons) + (17 + 2 * n_horizons)
    # parameter, state and flux rows; forcing columns; MPR-lite's 8 input rows
    cells = (rows * ld * 8 + FORCING_ROWS[process_pet] * n_forcing_days * sstride * 8
             + 8 * sstride * 8)
    window = 8  # default routing window (mhmhip_run_opts)
    rows_run = min(chunk_steps, n_steps)  # steps of the longest call
    # ring rows of a call: (tile depth + 1) windows + S + 1, at most the
    # call's steps, rounded up to whole windows, and the lag rows past them
    ring_rows = min((plan.tau_max + 1) * window + chunk_steps + 1, rows_run)
    ring = (-(-ring_rows // window) * window + TILE_HEIGHT) * ld * 8
    routing = (21 * n_cells + 6 * ld * 8 + 20 + 4 * max(1, n_cells - n_outlets)
               + 20 * plan.n_slots + 16 * plan.n_bins
               + 2 * n_sets * max(1, plan.n_bnd) * TILE_WINDOW_MAX * 8
               + 4 * max(1, n_gauges) + 8)
    series = rows_run * STEP_BYTES + rows_run * n_sets * max(1, n_gauges) * 8
    return cells + ring + routing + series


def prepare_model(req, device=0, timings=None):
    """Context with network, parameters, default states and forcing of a
    request; timings (a dict, optional) receives the seconds of each part."""
    dom = req.domain
    n_days = calendar.n_forcing_days_needed(req.n_steps)
    t0 = time.perf_counter()
    m = engine.Model([dom], n_sets=req.gammas.shape[0], timestep=req.timestep, route=True,
                     n_forcing_days=n_days, device=device)
    t1 = time.perf_counter()
    m.bind_network()
    t2 = time.perf_counter()
    # the sets' parameter rows come from MPR-lite on the device: the cell
    # inputs cross once, then each set is its gamma vector and one launch
    m.set_mprlite_inputs()
    for s_, g in enumerate(req.gammas):
        m.mprlite_params(g, s_)
        m.set_routing_params(PM.routing_gamma(g), s_)
    t3 = time.perf_counter()
    m.default_states()
    m.set_domain_forcing(n_days=n_days)
    m.sync()
    if timings is not None:
        timings.update(create_s=t1 - t0, network_s=t2 - t1, params_s=t3 - t2,
                       states_forcing_s=time.perf_counter() - t3)
    return m


def run_chunks(m, req, chunk_steps=73):
    """The request's horizon on context m in fused runs of chunk_steps."""
    steps = calendar.step_table(req.start, req.n_steps, req.timestep)
    parts = []
    for a in range(0, req.n_steps, chunk_steps):
        parts.append(m.run(steps[a:a + chunk_steps], mode=N.RUN_FUSED,
                           steps_per_launch=chunk_steps))
    return np.concatenate(parts)


def run_alone(req, chunk_steps=73, device=0, timings=None):
    """The request on its own (reference result for the mix); timings (a
    dict, optional) receives the seconds spent preparing the context on the
    host (parameter fields, network plan, uploads) and running it."""
    t0 = time.perf_counter()
    m = prepare_model(req, device, timings)
    t1 = time.perf_counter()
    g = run_chunks(m, req, chunk_steps)
    t2 = time.perf_counter()
    m.close()
    if timings is not None:
        timings["prepare_s"] = t1 - t0
        timings["run_s"] = t2 - t1
    return g


class Server:
    def __init__(self, hbm_budget_bytes, max_slots=8, chunk_steps=73, device=0, runner=None,
                 footprint=None):
        """runner(req) -> gauges runs one request (default: its own device
        context, chunked fused runs); footprint(req) -> bytes (default:
        estimate_bytes of its context)."""
        self.timings = {}
        self.runner = runner or (lambda req: run_alone(req, chunk_steps, device,
                                                       self.timings.setdefault(req.rid, {})))
        self.footprint = footprint or self._footprint
        self.budget = int(hbm_budget_bytes)
        self.max_slots = int(max_slots)
        self.chunk = int(chunk_steps)
        self.device = device
        self._lock = threading.Condition()
        self._queue = []
        self._in_use = 0
        self._running = 0
        self._results = {}
        self._errors = []
        self._t0 = None
        self.peak_slots = 0
        self.peak_bytes = 0

    def _footprint(self, req):
        dom = req.domain
        plan, _, _ = N.plan_network(dom)
        return estimate_bytes(dom.n_cells, req.gammas.shape[0], dom.horizon_depth.shape[0],
                              calendar.n_forcing_days_needed(req.n_steps), plan,
                              self.chunk, req.n_steps, dom.gauge_nodes.shape[0], dom.n_outlets)

    def submit(self, req):
        fp = self.footprint(req)
        if fp > self.budget:
            raise ValueError(f"request {req.rid} needs {fp} B of device memory, over the "
                             f"budget of {self.budget} B")
        with self._lock:
            self._queue.append((req, fp, time.perf_counter()))
            self._lock.notify_all()

    def _worker(self, req, fp, t_arr, t_adm):
        try:
            g = self.runner(req)
            t_fin = time.perf_counter()
            with self._lock:
                self._results[req.rid] = Result(req.rid, g, t_arr - self._t0, t_adm - self._t0,
                                                t_fin - self._t0)
        except Exception as e:  # pragma: no cover - surfaced by serve()
            with self._lock:
                self._errors.append((req.rid, repr(e)))
        finally:
            with self._lock:
                self._in_use -= fp
                self._running -= 1
                self._lock.notify_all()

    def _admit_loop(self, n_total, stop):
        threads = []
        finished = False
        while not finished:
            with self._lock:
                while True:
                    if len(self._results) + len(self._errors) >= n_total or stop.is_set():
                        finished = True
                        break
                    pick = None
                    if self._running < self.max_slots:
                        for i, (req, fp, t_arr) in enumerate(self._queue):
                            if self._in_use + fp <= self.budget:
                                pick = i
                                break
                    if pick is not None:
                        break
                    self._lock.wait(timeout=0.05)
                if finished:
                    break
                req, fp, t_arr = self._queue.pop(pick)
                self._in_use += fp
                self._running += 1
                self.peak_slots = max(self.peak_slots, self._running)
                self.peak_bytes = max(self.peak_bytes, self._in_use)
            t = threading.Thread(target=self._worker, args=(req, fp, t_arr, time.perf_counter()),
                                 daemon=True)
            t.start()
            threads.append(t)
        for t in threads:
            t.join()

    def serve(self, requests):
        """Feed requests at their arrival times, run them, return results by id."""
        self._t0 = time.perf_counter()
        stop = threading.Event()
        adm = threading.Thread(target=self._admit_loop, args=(len(requests), stop), daemon=True)
        adm.start()
        for req in sorted(requests, key=lambda r: r.arrival):
            wait = self._t0 + req.arrival - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            self.submit(req)
        adm.join()
        stop.set()
        if self._errors:
            raise RuntimeError(f"requests failed: {self._errors}")
        self.wall = time.perf_counter() - self._t0
        return dict(self._results)


def poisson_requests(n, qps, seed=0, cells=(1000, 100000), horizon=(30, 365), sets=(1, 64),
                     timestep=24, start=None, pool=None):
    """Requests with Poisson arrivals: basin size log-uniform, horizon and
    parameter-set count uniform within the given bounds.  pool: the requests
    draw their basins (terrain, network, forcing over the longest horizon)
    from that many pre-built ones, shared read-only, instead of one each --
    long streams without a basin and a forcing year built per request."""
    import datetime as dt
    rng = np.random.default_rng(seed)
    start = start or dt.datetime(1990, 1, 1)

    def basin(i, days):
        nc = int(math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1]))))
        nr = max(2, int(round(math.sqrt(nc))))
        dom = synth.make_domain(nr, max(2, nc // nr), seed=seed * 10007 + i)
        synth.make_forcing(dom, days, seed * 10007 + 5000 + i)
        return dom

    basins = None
    if pool:
        days = calendar.n_forcing_days_needed(horizon[1])
        basins = [basin(i, days) for i in range(pool)]
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / qps)
        h = int(rng.integers(horizon[0], horizon[1] + 1))
        if basins is None:
            dom = basin(i, calendar.n_forcing_days_needed(h))
        else:
            dom = basins[int(rng.integers(0, len(basins)))]
        k = int(rng.integers(sets[0], sets[1] + 1))
        gam = PM.draw(rng, k)
        reqs.append(Request(i, dom, gam, h, timestep, start, t))
    return reqs


def summarize(results, requests, wall, timings=None):
    lat = np.array([results[r.rid].latency for r in requests])
    cs = sum(r.cell_steps for r in requests)
    out = {"requests": len(requests), "p50_s": float(np.percentile(lat, 50)),
           "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
           "cell_steps_per_s": cs / wall, "wall_s": wall}
    if timings:
        for k in ("prepare_s", "run_s", "create_s", "network_s", "params_s", "states_forcing_s"):
            v = np.array([timings[r.rid][k] for r in requests if k in timings.get(r.rid, {})])
            if v.size:
                out[k.replace("_s", "") + "_p50_s"] = float(np.percentile(v, 50))
                out[k.replace("_s", "") + "_p99_s"] = float(np.percentile(v, 99))
    return out
