"""Adaptive-timestep routing with a constant celerity (processCase(8) = 2).

mrm_update_param (src/mRM/mo_mrm_mpr.f90:268-329) sets K = length /
celerity per link, picks the routing step TSrout from given_TS
(mo_mrm_constants.F90:42-46) with FORCES locate, and C1/C2 with the space
weight 0; mRM_routing then routes each step's qOUT rout_loop = max(1,
nint(1 / tsRoutFactor)) times and averages Qmod (mo_mrm_routing.f90:
224-284).

CPU tests pin the oracle: C1/C2 against Chow's Muskingum coefficients for
x = 0 (c0 = C2, c1 = C1 - C2, c2 = 1 - C1 with dt = TSrout), the routing
step against the largest given_TS not above min K (equality included), the
sub-step count from the factor, the averaging against repeated
one-loop calls, and, for routing steps longer than the model step, each
routing step's qOUT against the period's runoff the oracle reports step by
step.  GPU tests call mhmhip_set_routing_celerity and the per-step
entry points through the C-ABI against the oracle.
"""
import numpy as np
import pytest

import oracle_lib as O

GIVEN_TS = [60.0, 120.0, 180.0, 240.0, 300.0, 360.0, 600.0, 720.0, 900.0, 1200.0, 1800.0,
            3600.0, 7200.0, 10800.0, 14400.0, 21600.0, 28800.0, 43200.0, 86400.0]


@pytest.mark.parametrize("celerity", [0.3, 1.0, 2.5])
def test_coefficients_are_chows_for_zero_space_weight(celerity):
    length = np.array([1000.0, 1414.2135623730951, 2500.0, 1000.0])
    c1, c2, ts, _, _ = O.mrm_update_param_celerity(length, celerity, 1)
    K = length / celerity
    dt = ts
    chow_c0 = dt / (2 * K + dt)
    chow_c1 = dt / (2 * K + dt)
    chow_c2 = (2 * K - dt) / (2 * K + dt)
    assert np.allclose(c2, chow_c0, rtol=1e-14, atol=0)
    assert np.allclose(c1 - c2, chow_c1, rtol=1e-13, atol=1e-16)
    assert np.allclose(1.0 - c1, chow_c2, rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("kmin,want", [(1000.0, 900.0), (2000.0, 1800.0), (600.0, 600.0),
                                       (599.9999, 360.0), (30.0, 60.0), (1e6, 86400.0),
                                       (3600.0, 3600.0)])
def test_routing_step_is_the_largest_given_step_not_above_min_k(kmin, want):
    length = np.array([kmin * 2.0, kmin * 3.0, kmin])  # celerity 1: K = length
    _, _, ts, _, _ = O.mrm_update_param_celerity(length, 1.0, 1)
    assert ts == want
    below = [t for t in GIVEN_TS if t <= kmin]
    assert ts == (max(below) if below else GIVEN_TS[0])


@pytest.mark.parametrize("ts_len,timestep,loops,factor", [(900.0, 1, 4, 0.25), (1800.0, 1, 2, 0.5),
                                                          (3600.0, 1, 1, 1.0), (360.0, 1, 10, 0.1),
                                                          (3600.0, 24, 24, 1.0 / 24)])
def test_sub_steps_from_the_factor(ts_len, timestep, loops, factor):
    _, _, ts, f, n = O.mrm_update_param_celerity(np.array([ts_len]), 1.0, timestep)
    assert ts == ts_len and f == pytest.approx(factor, rel=1e-15) and n == loops


def _chain(n_loops):
    """Routing configuration of a 3-node chain 3 -> 2 -> 1 (outlet 1)."""
    rc = O.RouteCfg()
    rc.n_nodes, rc.n_outlets, rc.timestep, rc.process_case = 3, 1, 1, 2
    rc.map_flag, rc.n_inflow, rc.n_gauges, rc.rout_loop = 1, 0, 1, n_loops
    return rc


def test_averaging_equals_repeated_single_loops():
    """One call with rout_loop = 4 = the mean of four one-loop calls' Qmod on
    the same qOUT, and the same final states."""
    import ctypes as C
    from_n = np.array([3, 2], np.int32)
    to_n = np.array([2, 1], np.int32)
    perm = np.array([1, 2], np.int32)
    length = np.array([1000.0, 1414.2135623730951])
    c1, c2, _, _, _ = O.mrm_update_param_celerity(length, 1.0, 1)
    runoff = np.array([0.3, 0.7, 1.1])
    area = np.array([1.0, 1.0, 1.0])
    ids = np.array([1, 2, 3], np.int32)

    def call(loops, qtin, qtr):
        rc = _chain(loops)
        qout, qmod, g = np.zeros(3), np.zeros(3), np.zeros(1)
        O.lib().or_mrm_routing(C.byref(rc), O.dptr(np.zeros(5)), 3, O.dptr(runoff),
                               O.dptr(area), O.iptr(ids), O.dptr(area), O.iptr(ids),
                               O.iptr(perm), O.iptr(from_n), O.iptr(to_n), O.dptr(length),
                               O.dptr(np.full(2, 0.01)), O.dptr(np.zeros(2)), None, None, None,
                               None, O.iptr(np.array([1], np.int32)), O.dptr(c1.copy()),
                               O.dptr(c2.copy()), O.dptr(qout), O.dptr(qtin), O.dptr(qtr),
                               O.dptr(qmod), O.dptr(g))
        return qmod, g[0]

    qtin_a, qtr_a = np.zeros(6), np.zeros(6)
    qmod_a, g_a = call(4, qtin_a, qtr_a)
    qtin_b, qtr_b = np.zeros(6), np.zeros(6)
    acc = np.zeros(3)
    for _ in range(4):
        q, _ = call(1, qtin_b, qtr_b)
        acc = acc + q
    assert np.array_equal(qmod_a, acc / 4.0)
    assert g_a == qmod_a[0]
    assert np.array_equal(qtin_a, qtin_b) and np.array_equal(qtr_a, qtr_b)


# ---- GPU: the C-ABI against the oracle ---------------------------------------------

def _pair(n_steps=24):
    from mhm_amd import synth
    from pairs import make_pair
    dom = synth.make_domain(12, 10, seed=51)
    return make_pair([dom], n_sets=2, n_steps=n_steps, seed=13)


@pytest.mark.gpu
@pytest.mark.parametrize("celerity,loops", [(1.0, 4), (0.5, 2), (0.25, 1)])
def test_per_step_adaptive_routing_matches_oracle(celerity, loops):
    from oracle_sim import rel_close
    from pairs import compare_cells, compare_routing
    print(f"case: adaptive routing celerity={celerity}", flush=True)
    model, sim, steps, recs = _pair()
    for s in range(2):
        model.set_routing_celerity(celerity, s)
        sim.set_routing_celerity(celerity, s)
    assert sim.ds[0][0].adaptive[2] == loops
    for k in range(len(recs)):
        model.mhm_step(steps[k])
        g = model.mrm_step(steps[k])
        go = sim.step(recs[k])
        ok, worst = rel_close(g, go, 1e-10, 1e-12)
        assert ok, (k, worst)
    assert compare_cells(model, sim, 1e-10, 1e-12) == []
    assert compare_routing(model, sim, 1e-10, 1e-12) == []
    model.close()


@pytest.mark.gpu
def test_adaptive_routing_limits():
    from mhm_amd import _native as N
    print("case: adaptive routing limits", flush=True)
    model, sim, steps, recs = _pair(n_steps=4)
    # sets that disagree on the routing step
    model.set_routing_celerity(1.0, 0)
    model.set_routing_celerity(0.5, 1)
    model.mhm_step(steps[0])
    with pytest.raises(N.MhmHipError, match="differ"):
        model.mrm_step(steps[0])
    # routing steps longer than the model step: per step only
    model.set_routing_celerity(0.1, 0)
    model.set_routing_celerity(0.1, 1)
    model.mrm_step(steps[0])
    with pytest.raises(N.MhmHipError, match="longer"):
        model.run(steps[1:3], mode=N.RUN_FUSED, steps_per_launch=2)
    model.close()


def test_per_link_celerity_equals_constant_when_uniform():
    """processCase 3 with the same celerity on every link gives processCase
    2's coefficients and routing step bit for bit; per link, C1/C2 are Chow's
    x = 0 coefficients of that link's K = length / celerity."""
    rng = np.random.default_rng(4)
    length = rng.uniform(800.0, 1500.0, 40)
    a = O.mrm_update_param_celerity(length, 0.8, 1)
    b = O.mrm_update_param_links(length, np.full(40, 0.8), 1)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    cel = rng.uniform(0.3, 3.0, 40)
    c1, c2, ts, _, _ = O.mrm_update_param_links(length, cel, 1)
    K = length / cel
    assert ts == max(t for t in GIVEN_TS if t <= K.min())
    assert np.allclose(c2, ts / (2 * K + ts), rtol=1e-14, atol=0)
    assert np.allclose(1.0 - c1, (2 * K - ts) / (2 * K + ts), rtol=1e-13, atol=1e-16)


@pytest.mark.gpu
@pytest.mark.parametrize("route_mode,graph,mode", [(3, False, 1), (3, True, 1), (1, False, 1),
                                                   (2, False, 0)])
def test_time_loop_adaptive_routing_matches_oracle(route_mode, graph, mode):
    """Routing sub-steps inside the time loop (rout_loop = 4): every model
    step's qOUT routed four times, gauges and Qmod the average of the four,
    against the oracle -- with a per-link celerity (processCase 3)."""
    from oracle_sim import rel_close
    from pairs import compare_cells, compare_routing
    print(f"case: adaptive routing in the time loop route_mode={route_mode} graph={graph}",
          flush=True)
    model, sim, steps, recs = _pair(n_steps=30)
    dom = sim.domains[0]
    rng = np.random.default_rng(9)
    for s in range(2):
        cel = rng.uniform(0.8, 1.0, dom.n_links)  # K >= 1000 s on every link
        cel[0] = dom.length[0] / 950.0             # one link with K = 950 s: TSrout 900 s
        model.set_routing_celerity_links(cel, s)
        sim.set_routing_celerity_links(cel, s)
    assert sim.ds[0][0].adaptive[2] == sim.ds[0][1].adaptive[2] == 4
    g = np.concatenate([model.run(steps[:12], mode=mode, steps_per_launch=5,
                                  route_mode=route_mode, use_graph=graph),
                        model.run(steps[12:], mode=mode, steps_per_launch=5,
                                  route_mode=route_mode, use_graph=graph)])
    go = np.stack([sim.step(r) for r in recs])
    ok, worst = rel_close(g, go, 1e-10, 1e-12)
    assert ok, worst
    assert compare_cells(model, sim, 1e-10, 1e-12) == []
    assert compare_routing(model, sim, 1e-10, 1e-12) == []
    model.close()


@pytest.mark.gpu
@pytest.mark.parametrize("celerity,every,n_steps", [(0.1, 2, 24), (0.05, 4, 25)])
def test_long_routing_steps_match_oracle(celerity, every, n_steps):
    """Routing steps longer than the model step (tsRoutFactor = 2 and 4): the
    model steps accumulate RunToRout and the inflow (one inflow gauge), a
    routing step converts the sum with TS_rout and routes once; with 25
    steps and a factor of 4 the last period has one step (tt == nTimeSteps).
    Routed steps against the oracle, accumulating steps report nothing."""
    from mhm_amd import synth
    from oracle_sim import rel_close
    from pairs import compare_cells, compare_routing, make_pair
    print(f"case: long routing steps celerity={celerity}", flush=True)
    dom = synth.make_domain(12, 10, seed=52)
    indeg = np.bincount(dom.toN - 1, minlength=dom.n_cells)
    node = int(np.argsort(-indeg, kind="stable")[2]) + 1
    q = 1.5 + np.random.default_rng(3).gamma(2.0, 1.0, size=(n_steps, 1))
    inflow = {"domain": [0], "node": [node], "headwater": [1], "q": q}
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=n_steps, seed=14, inflow=inflow)
    for s in range(2):
        model.set_routing_celerity(celerity, s)
        sim.set_routing_celerity(celerity, s)
    assert sim.ds[0][0].adaptive[1] == float(every)
    routed = 0
    for k in range(n_steps):
        model.set_inflow(q[k:k + 1])
        model.mhm_step(steps[k])
        g = model.mrm_step(steps[k])
        go = sim.step(recs[k])
        due = (k + 1) % every == 0 or k + 1 == n_steps
        assert (g is not None) == due, k
        assert np.isnan(go).all() != due, k
        if due:
            routed += 1
            ok, worst = rel_close(g, go, 1e-10, 1e-12)
            assert ok, (k, worst)
    assert routed == -(-n_steps // every)
    assert compare_cells(model, sim, 1e-10, 1e-12) == []
    assert compare_routing(model, sim, 1e-10, 1e-12) == []
    model.close()


@pytest.mark.parametrize("celerity,every,n_steps", [(0.1, 2, 24), (0.05, 4, 25)])
def test_long_routing_steps_accumulate_the_period_runoff(celerity, every, n_steps):
    """Routing steps longer than the model step in the oracle
    (mo_mhm_interface_run.f90:494-512, 608-618): on a routing step the qOUT
    of every node is the L1_total_runoff summed over the period's model
    steps, times the cell area, converted with the routing step's length:
    sum(q) [mm] * area [km2] * 1000 / (3600 s * timestep * steps in the
    period) -- the last period of 25 steps by 4 is one step long.  Checked
    from the per-step runoff the oracle reports, on a network without
    inflow gauges (qOUT is then the accumulated runoff alone)."""
    from mhm_amd import calendar, parameters as PM, synth
    from oracle_sim import OracleSim
    from pairs import START
    dom = synth.make_domain(9, 8, seed=52)
    synth.make_forcing(dom, calendar.n_forcing_days_needed(n_steps), 53)
    sim = OracleSim([dom], n_sets=1, n_forcing_days=calendar.n_forcing_days_needed(n_steps))
    g = PM.draw(np.random.default_rng(54), 1)[0]
    sim.set_params(synth.cell_params(dom, g), 0)
    sim.default_states()
    sim.set_domain_forcing()
    sim.set_routing_celerity(celerity)
    assert sim.ds[0][0].adaptive[1] == float(every)
    period = []
    for r in calendar.step_records(START, n_steps, 1):
        go = sim.step(r)
        period.append(sim.ds[0][0].X[16].copy())  # this step's L1_total_runoff [mm]
        tt = int(r["tt"])
        if tt % every == 0 or tt == n_steps:
            assert not np.isnan(go).any(), tt
            want = np.sum(period, axis=0) * dom.area_km2 * 1000.0 / (3600.0 * len(period))
            assert np.allclose(sim.ds[0][0].qOUT, want, rtol=1e-14, atol=0.0), tt
            period = []
        else:
            assert np.isnan(go).all(), tt
    assert period == []
