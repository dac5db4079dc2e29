"""Inflow gauges (add_inflow, mo_mrm_pre_routing.f90:179-214, and the cut of
the links into a non-headwater inflow node, mo_mrm_routing.f90:452-457) on
every routing schedule, against the oracle at 1e-10."""
import sys

import numpy as np
import pytest

from mhm_amd import synth
from oracle_sim import rel_close
from pairs import compare_routing, make_pair

pytestmark = pytest.mark.gpu


def _inflow_nodes(dom):
    """Two nodes with upstream links (1-based): the one with most upstream
    links (no headwater: its links are cut) and another (headwater)."""
    indeg = np.bincount(dom.toN - 1, minlength=dom.n_cells)
    order = np.argsort(-indeg, kind="stable")
    return int(order[0]) + 1, int(order[3]) + 1


@pytest.mark.parametrize("route_mode", [1, 2, 3])
def test_inflow_gauges_match_oracle(route_mode):
    print(f"case: inflow route_mode={route_mode}", flush=True)
    sys.stdout.flush()
    dom = synth.make_domain(20, 24, seed=17)
    a, b = _inflow_nodes(dom)
    n_steps = 40
    rng = np.random.default_rng(3)
    q = 2.0 + rng.gamma(2.0, 1.5, size=(n_steps, 2))
    inflow = {"domain": [0, 0], "node": [a, b], "headwater": [0, 1], "q": q}
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=n_steps, seed=5, inflow=inflow)
    g = model.run(steps, mode=1, steps_per_launch=8, route_mode=route_mode)
    go = np.stack([sim.step(r) for r in recs])
    ok, worst = rel_close(g, go, 1e-10, 1e-12)
    assert ok, worst
    assert compare_routing(model, sim, 1e-10, 1e-12) == []
    model.close()


def test_rebind_with_more_gauges_and_inflow_points():
    """A network bound again with more gauges and inflow gauges: the gauge
    series and inflow buffers of the first network are dropped (the inflow
    discharge must be given again) and the new lists route as the oracle."""
    print("case: rebind 20x24, 1 -> 3 gauges, 0 -> 2 inflow gauges", flush=True)
    from mhm_amd import parameters as PM
    dom = synth.make_domain(20, 24, seed=17)
    n1, n2 = 10, 30
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=n1 + n2, seed=5)
    g = model.run(steps[:n1], mode=1, steps_per_launch=4, route_mode=3)
    go = np.stack([sim.step(r) for r in recs[:n1]])
    assert rel_close(g, go, 1e-10, 1e-12)[0]
    a, b = _inflow_nodes(dom)
    order = np.argsort(-dom.depth, kind="stable")
    dom.gauge_nodes = np.array([dom.gauge_nodes[0], a, int(order[5]) + 1], np.int32)
    rng = np.random.default_rng(4)
    q = 1.0 + rng.gamma(2.0, 1.5, size=(n2, 2))
    inflow = {"domain": [0, 0], "node": [a, b], "headwater": [0, 1], "q": q}
    model.bind_network(inflow=inflow)
    with pytest.raises(RuntimeError, match="inflow"):
        model.run(steps[n1:], mode=1, steps_per_launch=4, route_mode=3)
    model.set_inflow(q)
    gam = PM.draw(np.random.default_rng(5), 2).reshape(2, 1, -1)
    for s in range(2):
        model.set_routing_params(PM.routing_gamma(gam[s, 0]), s, 0)
        o = sim.ds[0][s]
        for arr in (o.qOUT, o.qTIN, o.qTR, o.Qmod):
            arr[:] = 0.0  # binding a network resets the routing states (mo_mrm_init.f90:453-494)
    dom._oracle_inflow = {"node": np.array([a, b]), "headwater": np.array([0, 1]), "q": q,
                          "tt0": n1}
    g = model.run(steps[n1:], mode=1, steps_per_launch=4, route_mode=3)
    go = np.stack([sim.step(r) for r in recs[n1:]])
    assert g.shape == (n2, 2, 3)
    ok, worst = rel_close(g, go, 1e-10, 1e-12)
    assert ok, worst
    assert compare_routing(model, sim, 1e-10, 1e-12) == []
    model.close()
