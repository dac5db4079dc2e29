"""Routing on an L11 grid coarser than the L1 grid (resolution_Routing a
multiple of resolution_Hydrology, as in the reference's check case_04):
L11_runoff_acc's aggregating branch (map_flag, mo_mrm_pre_routing.f90:
110-126) on the device -- every node's qOUT is its L1 cells' runoff times
their area, summed in L1 order, * 1000 / TST -- followed by the routing on
the coarse network.  Gauges, every L1 state and flux and every routing state
against the oracle, for the time loop (fused, captured, plain; tiled,
diagonal and workgroup-resident routing), the per-step path, basins of
different refinement packed with an L1 = L11 basin, inflow gauges, routing
sub-steps and routing steps longer than the model step."""
import sys

import numpy as np
import pytest

from mhm_amd import synth
from oracle_sim import rel_close
from pairs import compare_cells, compare_routing, make_pair

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-10, 1e-12


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _domains():
    # refinement 2 and 3 and an L1 = L11 basin, packed in one context
    return [synth.refine_domain(synth.make_domain(7, 8, seed=71), 2, 72),
            synth.refine_domain(synth.make_domain(5, 6, seed=73), 3, 74),
            synth.make_domain(6, 7, seed=75)]


def _check(model, sim, g, recs):
    go = np.stack([sim.step(r) for r in recs])
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []


@pytest.mark.parametrize("how,route_mode", [("fused", 3), ("graph", 3), ("plain", 3),
                                            ("fused", 1), ("fused", 2), ("per_step", 3)])
def test_coarse_routing_grid_matches_oracle(how, route_mode):
    _log(f"case: coarse routing grid {how} route_mode {route_mode}")
    doms = _domains()
    model, sim, steps, recs = make_pair(doms, n_sets=2, n_steps=30, seed=76)
    assert model.n_nodes == 56 + 30 + 42 and model.n_cells == 224 + 270 + 42
    if how == "per_step":
        g = []
        for k in range(30):
            model.mhm_step(steps[k])
            g.append(model.mrm_step(steps[k]))
        g = np.stack(g)
    else:
        kw = {"fused": dict(mode=1, steps_per_launch=8), "plain": dict(mode=0),
              "graph": dict(mode=1, steps_per_launch=8, use_graph=True)}[how]
        kw["route_mode"] = route_mode
        g = np.concatenate([model.run(steps[:12], **kw), model.run(steps[12:], **kw)])
    _check(model, sim, g, recs)
    model.close()


def test_coarse_routing_grid_with_inflow_gauges():
    _log("case: coarse routing grid, inflow gauges")
    doms = _domains()[:2]
    rng = np.random.default_rng(77)
    inflow = {"domain": [0, 1], "node": [3, 8], "headwater": [1, 0],
              "q": rng.uniform(0.5, 4.0, (30, 2))}
    model, sim, steps, recs = make_pair(doms, n_sets=1, n_steps=30, seed=78, inflow=inflow)
    g = model.run(steps, mode=1, steps_per_launch=10)
    _check(model, sim, g, recs)
    model.close()


@pytest.mark.parametrize("celerity,path", [(2.0, "fused"), (2.0, "per_step"), (0.05, "per_step")])
def test_coarse_routing_grid_adaptive_steps(celerity, path):
    """Routing sub-steps (celerity 2 m/s: routing steps shorter than the
    hour) and routing steps longer than the model step (0.05 m/s: RunToRout
    summed per L1 cell over the steps and aggregated on the routing step)."""
    _log(f"case: coarse routing grid adaptive celerity {celerity} {path}")
    doms = _domains()[:2]
    n = 24
    model, sim, steps, recs = make_pair(doms, n_sets=1, n_steps=n, seed=79)
    model.set_routing_celerity(celerity, 0)
    sim.set_routing_celerity(celerity, 0)
    if path == "fused":
        g = model.run(steps, mode=1, steps_per_launch=8)
        _check(model, sim, g, recs)
    else:
        routed = 0
        for k in range(n):
            model.mhm_step(steps[k])
            g = model.mrm_step(steps[k])
            go = sim.step(recs[k])
            assert (g is None) == bool(np.isnan(go).all()), k
            if g is not None:
                routed += 1
                ok, worst = rel_close(g, go, RTOL, ATOL)
                assert ok, (k, worst)
        assert routed > 0
        assert compare_cells(model, sim, RTOL, ATOL) == []
        assert compare_routing(model, sim, RTOL, ATOL) == []
    model.close()


def test_coarse_routing_grid_limits():
    _log("case: coarse routing grid limits")
    doms = _domains()[:1]
    model, sim, steps, recs = make_pair(doms, n_sets=1, n_steps=8, seed=80)
    with pytest.raises(RuntimeError, match="coarser"):
        model.run(steps, mode=1, steps_per_launch=8, noalloc=True)
    model.close()


@pytest.mark.parametrize("coarse", [False, True])
def test_binding_after_data_keeps_every_cell_value(coarse):
    """Binding a network re-lays the cell data already on the device into the
    new cell order (routing order, or grouped by node on a coarser routing
    grid): a context that ran 6 steps and started output sums before its
    network was bound continues exactly like one bound from the start --
    every state, flux and output sum bit for bit."""
    from mhm_amd import calendar, engine, parameters as PM
    from pairs import START
    _log(f"case: binding after data coarse={coarse}")
    dom = (synth.refine_domain(synth.make_domain(6, 7, seed=81), 2, 82) if coarse
           else synth.make_domain(12, 11, seed=81))
    n = 16
    synth.make_forcing(dom, calendar.n_forcing_days_needed(n), 83)
    steps = calendar.step_table(START, n, 1)
    gam = PM.draw(np.random.default_rng(84), 2)
    flags = [1] * 21
    flags[17] = 0
    models = []
    for bind_first in (True, False):
        m = engine.Model([dom], n_sets=2, timestep=1, route=True,
                         n_forcing_days=calendar.n_forcing_days_needed(n))
        if bind_first:
            m.bind_network()
        for s in range(2):
            m.set_params(synth.cell_params(dom, gam[s]), s)
        m.default_states()
        m.set_domain_forcing()
        m.output_accumulate(flags)
        m.run(steps[:6], mode=1, steps_per_launch=3, want_gauges=False)
        if not bind_first:
            m.bind_network()
        for s in range(2):
            m.set_routing_params(PM.routing_gamma(gam[s]), s)
        m.run(steps[6:], mode=1, steps_per_launch=5)
        models.append(m)
    a, b = models
    for s in range(2):
        for name in ("INTER", "SNOWPACK", "SEALSTW", "UNSATSTW", "SATSTW", "SOILMOIST",
                     "TOTAL_RUNOFF", "AETSOIL", "PET_CALC"):
            assert np.array_equal(a.get(name, s), b.get(name, s)), (s, name)
        for k in (2, 3, 11, 19):
            sa, ca = a.output_sum(k, 1, s)
            sb, cb = b.output_sum(k, 1, s)
            assert ca == cb == n and np.array_equal(sa, sb), (s, k)
    a.close()
    b.close()
