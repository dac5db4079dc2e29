"""Routing tiles of other heights (mhmhip_set_tile_height): the partition a
context binds changes the number and length of the tiled routing launches,
not the results.  At heights 2 (a node, its upstream nodes or proxies), 32,
48 and 64
the per-step path (mhm_step + mrm_step) and the captured fused time loop give
the default height's gauges, cell states and routing states bit for bit, and
the oracle's within 1e-10; two basins packed, two parameter sets, a ring
shorter than the run."""
import sys

import numpy as np
import pytest

from mhm_amd import synth
from oracle_sim import STATE_NAMES, rel_close
from pairs import compare_routing, make_pair

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-10, 1e-12


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _doms():
    return [synth.make_domain(36, 30, seed=31), synth.make_domain(11, 14, seed=32)]


@pytest.mark.parametrize("height", [2, 32, 48, 64])
def test_tile_height_gives_the_default_run(height):
    _log(f"case: tile height {height}, 2 basins, 2 sets, 24 per-step + 160 fused hourly steps")
    n = 184
    ref, sim, steps, recs = make_pair(_doms(), n_sets=2, n_steps=n, seed=5)
    alt, _, _, _ = make_pair(_doms(), n_sets=2, n_steps=n, seed=5, tile_height=height)
    for k in range(24):
        go = sim.step(recs[k])
        for m in (ref, alt):
            m.mhm_step(steps[k])
        g1, g2 = ref.mrm_step(steps[k]), alt.mrm_step(steps[k])
        assert np.array_equal(g1, g2), k
        assert rel_close(g2, go, RTOL, ATOL)[0], k
    kw = dict(mode=1, steps_per_launch=24, route_mode=3, route_window=16, use_graph=True)
    g1 = ref.run(steps[24:], **kw)
    g2 = alt.run(steps[24:], **kw)
    go = np.stack([sim.step(r) for r in recs[24:]])
    assert np.array_equal(g1, g2)
    assert rel_close(g2, go, RTOL, ATOL)[0]
    for s in range(2):
        for name in STATE_NAMES:
            assert np.array_equal(ref.get(name, s), alt.get(name, s)), (name, s)
        r1, r2 = ref.routing_state(s), alt.routing_state(s)
        for name in r1:
            assert np.array_equal(r1[name], r2[name]), (name, s)
    assert compare_routing(alt, sim, RTOL, ATOL) == []
    ref.close()
    alt.close()


def test_tile_height_argument_checks():
    _log("case: tile height argument checks")
    from mhm_amd import engine
    dom = synth.make_domain(5, 6, seed=33)
    m = engine.Model([dom], n_sets=1, timestep=1, route=True, n_forcing_days=2)
    for bad in (-1, 1, 65):
        with pytest.raises(RuntimeError, match="height must be"):
            m.set_tile_height(bad)
    m.set_tile_height(0)
    m.set_tile_height(40)
    m.bind_network()
    with pytest.raises(RuntimeError, match="before bind_network"):
        m.set_tile_height(32)
    m.close()
