"""The workloads at their full sizes.  The headline basin: 1024 x 1024
cells, D8 network of depth 1848 (1,048,575 links), 3 horizons, hourly steps;
the 256-set ensemble (64 x 64 cells, daily x 3650) and the 64-basin pack
(hourly x 720) likewise.

* Parity: 48 steps of the pipelined time loop (fused cells S = 24, tiled
  routing in windows of 64, hipGraph) against the oracle on the same
  inputs: every gauge discharge, cell state and routing state within the
  suite's 1e-10 relative bar.  The oracle finishes them in a few seconds.
* Schedule independence over the full 2160 steps (90 days), where the
  oracle would take minutes: from the same starting states, the fused loop
  (144 steps per launch, as the bench runs it) with tiled routing and the
  plain loop (one launch per step, every flux
  written) with diagonal routing give the same gauge series and the same
  final states bit for bit -- the two share no kernel, so a fault in either
  schedule's indexing at full size (τmax 123 tile launches, 4008
  diagonals) shows here.
"""
import sys

import numpy as np
import pytest

from mhm_amd import _native as N, calendar, synth
from oracle_sim import rel_close
from pairs import START, compare_cells, compare_routing, make_pair

pytestmark = pytest.mark.gpu

RTOL = 1e-10
ATOL = 1e-12
STATES = ("INTER", "SNOWPACK", "SEALSTW", "UNSATSTW", "SATSTW", "SOILMOIST")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def basin():
    _log("case: full size 1024x1024 setup")
    dom = synth.make_domain(1024, 1024, seed=1)
    model, sim, steps, recs = make_pair([dom], n_sets=1, n_steps=48, seed=5)
    yield dom, model, sim, steps, recs
    model.close()


def test_full_size_time_loop_matches_oracle(basin):
    dom, model, sim, steps, recs = basin
    _log(f"case: full size parity, {dom.n_cells} cells, dmax {dom.dmax}, 48 steps")
    g = model.run(steps, mode=N.RUN_FUSED, steps_per_launch=24, route_window=64, use_graph=True)
    go = np.stack([sim.step(r) for r in recs])
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []


def test_full_size_schedules_agree_bit_for_bit(basin):
    dom, model, _, _, _ = basin
    n_steps = 2160
    _log(f"case: full size schedules, {n_steps} steps")
    # a fresh context on the full 90 days of forcing
    synth.make_forcing(dom, calendar.n_forcing_days_needed(n_steps), 300,
                       doy0=START.timetuple().tm_yday)
    m2, _, steps, _ = make_pair([dom], n_sets=1, n_steps=n_steps, seed=5)
    snap = m2.snapshot()
    g_fused = m2.run(steps, mode=N.RUN_FUSED, steps_per_launch=144, route_window=64,
                     use_graph=True)
    s_fused = {k: m2.get(k) for k in STATES}
    r_fused = m2.routing_state(0)
    m2.restore(snap)
    g_plain = m2.run(steps, mode=N.RUN_PLAIN, route_mode=1)
    s_plain = {k: m2.get(k) for k in STATES}
    r_plain = m2.routing_state(0)
    m2.close()
    assert np.isfinite(g_fused).all() and np.abs(g_fused).max() > 0.0
    assert np.array_equal(g_fused, g_plain)
    for k in STATES:
        assert np.array_equal(s_fused[k], s_plain[k]), k
    for k in ("qTIN", "qTR", "Qmod"):
        assert np.array_equal(r_fused[k], r_plain[k]), k


def _parity_then_schedules(model, sim, steps, recs, n_check, S_check, S_full):
    """The first n_check steps against the oracle, then the whole horizon
    fused/tiled against plain/diagonal from the same states, bit for bit."""
    snap = model.snapshot()
    g = model.run(steps[:n_check], mode=N.RUN_FUSED, steps_per_launch=S_check, route_mode=3,
                  use_graph=True)
    go = np.stack([sim.step(r) for r in recs[:n_check]])
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []
    model.restore(snap)
    g_fused = model.run(steps, mode=N.RUN_FUSED, steps_per_launch=S_full, route_mode=3,
                        route_window=64, use_graph=True)
    s_fused = [{k: model.get(k, s) for k in STATES} for s in range(model.n_sets)]
    r_fused = [model.routing_state(s) for s in range(model.n_sets)]
    model.restore(snap)
    g_plain = model.run(steps, mode=N.RUN_PLAIN, route_mode=1)
    assert np.isfinite(g_fused).all()
    assert np.array_equal(g_fused, g_plain)
    for s in range(model.n_sets):
        for k in STATES:
            assert np.array_equal(s_fused[s][k], model.get(k, s)), (s, k)
        r = model.routing_state(s)
        for k in ("qTIN", "qTR", "Qmod"):
            assert np.array_equal(r_fused[s][k], r[k]), (s, k)


def test_ensemble_full_size():
    """256 parameter sets x 4096 cells (64 x 64), daily x 3650 steps."""
    _log("case: full size ensemble 256 sets x 64x64, daily x 3650")
    dom = synth.make_domain(64, 64, seed=3)
    model, sim, steps, recs = make_pair([dom], n_sets=256, timestep=24, n_steps=3650, seed=8)
    _parity_then_schedules(model, sim, steps, recs, n_check=30, S_check=10, S_full=365)
    model.close()


def test_multi_basin_full_size():
    """64 basins, log-uniform 500..200k cells, each its own DEM, network,
    forcing and parameters, packed in one context; hourly x 720 steps."""
    import math
    rng = np.random.default_rng(7)
    sizes = np.exp(rng.uniform(math.log(500), math.log(200000), 64)).astype(int)
    doms = []
    for b, n in enumerate(sizes):
        nr = max(2, int(round(math.sqrt(n) * rng.uniform(0.7, 1.3))))
        nc = max(2, int(round(n / nr)))
        doms.append(synth.make_domain(nr, nc, seed=7000 + b))
    _log(f"case: full size multi-basin, {sum(d.n_cells for d in doms)} cells in 64 basins, "
         f"hourly x 720")
    model, sim, steps, recs = make_pair(doms, n_sets=1, n_steps=720, seed=9)
    _parity_then_schedules(model, sim, steps, recs, n_check=24, S_check=12, S_full=24)
    model.close()


def test_year_basin_whole_horizon_matches_oracle():
    """The 64 x 64 basin (4096 cells, 3 horizons) hourly x 8760 (one year),
    as the bench runs it (fused S = 144, tiled routing in windows of 64,
    hipGraph; the ring wraps many times): every gauge value of the whole
    year, the final cell states and fluxes and the routing states against
    the oracle at the suite's 1e-10 bar -- the chain's own exp/log/pow and
    the device acos/sin against libm over 8760 steps -- and the plain loop
    with diagonal routing from the same states bit for bit."""
    n_steps = 8760
    _log(f"case: year basin 64x64, hourly x {n_steps}, whole horizon against the oracle")
    dom = synth.make_domain(64, 64, seed=2)
    model, sim, steps, recs = make_pair([dom], n_sets=1, n_steps=n_steps, seed=6)
    snap = model.snapshot()
    g = model.run(steps, mode=N.RUN_FUSED, steps_per_launch=144, route_window=64,
                  use_graph=True)
    go = np.stack([sim.step(r) for r in recs])
    assert np.isfinite(g).all() and np.abs(g).max() > 0.0
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []
    s_fused = {k: model.get(k) for k in STATES}
    r_fused = model.routing_state(0)
    model.restore(snap)
    g_plain = model.run(steps, mode=N.RUN_PLAIN, route_mode=1)
    assert np.array_equal(g, g_plain)
    for k in STATES:
        assert np.array_equal(s_fused[k], model.get(k)), k
    r = model.routing_state(0)
    for k in ("qTIN", "qTR", "Qmod"):
        assert np.array_equal(r_fused[k], r[k]), k
    model.close()
