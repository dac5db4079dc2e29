"""Serving loop scheduling on the CPU: a stand-in runner replaces the device
work, so admission by footprint, continuous batching and latency accounting
are checked without a GPU."""
import datetime as dt
import threading
import time

import numpy as np
import pytest

from mhm_amd import serve, synth


def _req(rid, arrival, n_steps=10, n_cells_side=4):
    dom = synth.make_domain(n_cells_side, n_cells_side, seed=rid + 1)
    return serve.Request(rid, dom, np.zeros((1, 5)), n_steps, 24, dt.datetime(1990, 1, 1), arrival)


class _FakeGPU:
    def __init__(self, dur):
        self.dur = dur
        self.lock = threading.Lock()
        self.active = 0
        self.peak = 0
        self.starts = {}

    def __call__(self, req):
        with self.lock:
            self.active += 1
            self.peak = max(self.peak, self.active)
            self.starts[req.rid] = time.perf_counter()
        time.sleep(self.dur[req.rid])
        with self.lock:
            self.active -= 1
        return np.full((req.n_steps, 1, 1), float(req.rid))


def test_requests_overlap_and_late_arrivals_join():
    reqs = [_req(0, 0.0), _req(1, 0.05), _req(2, 0.10), _req(3, 0.12)]
    fake = _FakeGPU({0: 0.4, 1: 0.4, 2: 0.1, 3: 0.1})
    srv = serve.Server(hbm_budget_bytes=10, max_slots=4, runner=fake, footprint=lambda r: 1)
    res = srv.serve(reqs)
    assert sorted(res) == [0, 1, 2, 3]
    assert fake.peak >= 3  # requests that arrived while others ran were admitted at once
    for r in reqs:
        assert res[r.rid].gauges[0, 0, 0] == r.rid
        assert res[r.rid].latency >= 0.0
        assert res[r.rid].admitted >= res[r.rid].arrival - 1e-6


def test_admission_respects_the_memory_budget_first_fit():
    reqs = [_req(0, 0.0), _req(1, 0.01), _req(2, 0.02)]
    sizes = {0: 6, 1: 6, 2: 3}
    fake = _FakeGPU({0: 0.3, 1: 0.1, 2: 0.1})
    srv = serve.Server(hbm_budget_bytes=10, max_slots=8, runner=fake,
                       footprint=lambda r: sizes[r.rid])
    res = srv.serve(reqs)
    assert srv.peak_bytes <= 10
    # request 1 does not fit next to 0, the smaller request 2 does and goes first
    assert fake.starts[2] < fake.starts[1]
    assert res[1].admitted >= res[0].finished - 1e-3


def test_oversized_request_is_rejected():
    srv = serve.Server(hbm_budget_bytes=10, runner=lambda r: None, footprint=lambda r: 11)
    with pytest.raises(ValueError, match="over the budget"):
        srv.submit(_req(0, 0.0))


def test_summary_percentiles():
    reqs = [_req(i, 0.0) for i in range(3)]
    res = {i: serve.Result(i, None, 0.0, 0.0, float(i + 1)) for i in range(3)}
    s = serve.summarize(res, reqs, wall=3.0)
    assert s["p50_s"] == 2.0 and s["requests"] == 3
    assert s["cell_steps_per_s"] == pytest.approx(3 * 16 * 10 / 3.0)


def test_footprint_estimate_grows_with_the_request():
    from mhm_amd import _native as N, synth
    plan, _, _ = N.plan_network(synth.make_domain(30, 30, seed=1))
    a = serve.estimate_bytes(1000, 1, 3, 10, plan, 73, 100)
    assert serve.estimate_bytes(2000, 1, 3, 10, plan, 73, 100) > a
    assert serve.estimate_bytes(1000, 8, 3, 10, plan, 73, 100) > 4 * a
    assert serve.estimate_bytes(1000, 1, 3, 40, plan, 73, 100) > a
