"""Serving on the GPU: requests of different sizes, horizons and set counts
arriving over time run with continuous batching; every result is bitwise
the same as the request run alone, and one request matches the oracle."""
import datetime as dt

import numpy as np
import pytest

from mhm_amd import calendar, parameters as PM, serve, synth
from oracle_sim import OracleSim, rel_close

pytestmark = pytest.mark.gpu


def test_mix_equals_alone_and_oracle():
    print("case: serving mix", flush=True)
    # arrivals well inside one request's preparation time, so that requests
    # are on the device together
    reqs = serve.poisson_requests(8, qps=2000.0, seed=5, cells=(300, 3000), horizon=(20, 90),
                                  sets=(1, 6))
    srv = serve.Server(hbm_budget_bytes=2 << 30, max_slots=4, chunk_steps=30)
    res = srv.serve(reqs)
    assert srv.peak_slots >= 2, "no two requests were on the device together"
    for r in reqs:
        alone = serve.run_alone(r, chunk_steps=30)
        assert np.array_equal(res[r.rid].gauges, alone), r.rid
    # one request against the oracle (daily steps, reference iMeteoTS rule)
    r = reqs[0]
    n_days = calendar.n_forcing_days_needed(r.n_steps)
    sim = OracleSim([r.domain], n_sets=r.gammas.shape[0], timestep=24, n_forcing_days=n_days)
    for s, g in enumerate(r.gammas):
        sim.set_params(synth.cell_params(r.domain, g), s)
        sim.set_routing_params(PM.routing_gamma(g), s)
    sim.default_states()
    sim.set_domain_forcing()
    recs = calendar.step_records(r.start, r.n_steps, 24)
    go = np.stack([sim.step(rec) for rec in recs])
    assert rel_close(res[r.rid].gauges, go, 1e-10, 1e-12)[0]
    s = serve.summarize(res, reqs, srv.wall)
    assert s["p99_s"] >= s["p50_s"] > 0.0


def test_footprint_estimate_matches_the_context():
    """Admission's estimate equals what the context allocates on the device."""
    print("case: serving footprint", flush=True)
    for r in serve.poisson_requests(3, qps=10.0, seed=8, cells=(500, 5000), horizon=(20, 60),
                                    sets=(1, 9)):
        m = serve.prepare_model(r)
        serve.run_chunks(m, r, 30)
        srv = serve.Server(hbm_budget_bytes=1 << 40, chunk_steps=30)
        est = srv.footprint(r)
        assert est == m.device_bytes, (r.rid, est, m.device_bytes)
        m.close()



def test_mix_equals_alone_at_the_stream_bounds():
    """Requests at the stream's upper bounds (50k..100k cells, 32..64
    parameter sets) arriving together: each result in the mix is bitwise the
    one it has alone."""
    print("case: serving mix at 50k..100k cells x 32..64 sets", flush=True)
    reqs = serve.poisson_requests(3, qps=1000.0, seed=17, cells=(50000, 100000),
                                  horizon=(30, 45), sets=(32, 64))
    srv = serve.Server(hbm_budget_bytes=64 << 30, max_slots=3, chunk_steps=73)
    res = srv.serve(reqs)
    assert srv.peak_slots >= 2, "no two requests were on the device together"
    for r in reqs:
        assert r.domain.n_cells >= 49000 and r.gammas.shape[0] >= 32
        assert np.array_equal(res[r.rid].gauges, serve.run_alone(r, chunk_steps=73)), r.rid
