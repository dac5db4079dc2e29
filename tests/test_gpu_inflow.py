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
