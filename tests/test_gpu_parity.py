"""HIP path against the oracle on identical seeded inputs (FP64).

Bar: every per-cell state and flux, every routing state and every gauge
discharge within 1e-10 relative (|a-b| <= 1e-10*max(|a|,|b|) + 1e-12).
The per-step entry points (mhmhip_mhm_step + mhmhip_mrm_step, the drop-in
for mHM() + mRM_routing()) and the pipelined time loop (mhmhip_run, plain
and fused, both routing schedules) are checked.
"""
import sys

import numpy as np
import pytest

from mhm_amd import synth
from oracle_sim import rel_close
from pairs import compare_cells, compare_routing, make_pair

pytestmark = pytest.mark.gpu

RTOL = 1e-10
ATOL = 1e-12


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_series(sim, recs):
    return np.stack([sim.step(r) for r in recs])


def test_step_by_step_matches_oracle():
    _log("case: step_by_step 24x24 sets=2 hourly")
    dom = synth.make_domain(24, 24, seed=11)
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=30, seed=3)
    for k in range(len(recs)):
        model.mhm_step(steps[k])
        g = model.mrm_step(steps[k])
        go = sim.step(recs[k])
        ok, worst = rel_close(g, go, RTOL, ATOL)
        assert ok, (k, worst, g, go)
        if k in (0, 1, 7, 29):
            assert compare_cells(model, sim, RTOL, ATOL) == [], k
            assert compare_routing(model, sim, RTOL, ATOL) == [], k
    model.close()


@pytest.mark.parametrize("mode,window,S", [(1, 1, 5), (1, 7, 24), (1, 64, 24), (1, 40, 48),
                                           (0, 64, 0)])
def test_route_windows_and_shared_gauge_nodes(mode, window, S):
    """Tiled routing at windows shorter and longer than the cell launches, and
    gauges that share a node (the kernel writes a node's first gauge, the
    host copies it to the others)."""
    _log(f"case: tile windows mode={mode} window={window} S={S}, shared gauge nodes")
    dom = synth.make_domain(30, 26, seed=21)
    outlet, inner = int(dom.gauge_nodes[0]), int(dom.toN[5])
    dom.gauge_nodes = np.array([outlet, inner, outlet, inner, inner], np.int32)
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=70, seed=8)
    g = model.run(steps, mode=mode, steps_per_launch=S, route_mode=3, route_window=window)
    go = _oracle_series(sim, recs)
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert np.array_equal(g[:, :, 0], g[:, :, 2]) and np.array_equal(g[:, :, 1], g[:, :, 4])
    assert compare_routing(model, sim, RTOL, ATOL) == []
    model.close()


@pytest.mark.parametrize("window", [8, 64])
def test_gauge_at_every_node(window):
    """Every node a gauge: at window 8 the tiled kernel stages hundreds of
    output rows per bin in LDS; at window 64 they would not fit and the
    kernel stores them directly.  Both agree with the oracle."""
    _log(f"case: gauge at every node window={window}")
    dom = synth.make_domain(16, 16, seed=23)
    dom.gauge_nodes = np.arange(1, dom.n_cells + 1, dtype=np.int32)
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=40, seed=10)
    g = model.run(steps, mode=1, steps_per_launch=8, route_mode=3, route_window=window)
    go = _oracle_series(sim, recs)
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_routing(model, sim, RTOL, ATOL) == []
    model.close()


@pytest.mark.parametrize("mode,S,route_mode", [(0, 0, 1), (1, 24, 1), (1, 5, 2), (1, 48, 2),
                                               (0, 0, 3), (1, 24, 3), (1, 5, 3), (1, 48, 3)])
def test_run_matches_oracle(mode, S, route_mode):
    _log(f"case: run mode={mode} S={S} route_mode={route_mode}")
    dom = synth.make_domain(20, 28, seed=5)
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=60, seed=7)
    g = model.run(steps, mode=mode, steps_per_launch=S, route_mode=route_mode)
    go = _oracle_series(sim, recs)
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []
    model.close()


def test_run_in_pieces_equals_one_run():
    """Splitting the time loop over several calls changes nothing (states persist on device)."""
    _log("case: run pieces")
    dom = synth.make_domain(16, 16, seed=8)
    m1, _, steps, _ = make_pair([dom], n_sets=1, n_steps=50, seed=2)
    g1 = m1.run(steps, mode=1, steps_per_launch=24)
    m2, _, _, _ = make_pair([dom], n_sets=1, n_steps=50, seed=2)
    parts = [m2.run(steps[a:b], mode=1, steps_per_launch=7) for a, b in ((0, 13), (13, 14), (14, 50))]
    assert np.array_equal(np.concatenate(parts), g1)
    for name in ("SOILMOIST", "SATSTW", "TOTAL_RUNOFF"):
        assert np.array_equal(m1.get(name), m2.get(name))
    m1.close()
    m2.close()


def test_graph_replays_match_eager_and_oracle():
    """The time loop captured as a hipGraph and replayed on later chunks of the
    run (new step tables, eager reads in between) equals the eager loop and the
    oracle on every chunk."""
    _log("case: graph replays")
    dom = synth.make_domain(18, 22, seed=12)
    m1, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=48, seed=5)
    m2, _, _, _ = make_pair([dom], n_sets=2, n_steps=48, seed=5)
    for a in range(0, 48, 12):
        g1 = m1.run(steps[a:a + 12], mode=1, steps_per_launch=5)
        g2 = m2.run(steps[a:a + 12], mode=1, steps_per_launch=5, use_graph=True)
        go = _oracle_series(sim, recs[a:a + 12])
        assert np.array_equal(g1, g2), a
        assert rel_close(g2, go, RTOL, ATOL)[0], a
        for name in ("SOILMOIST", "UNSATSTW", "TOTAL_RUNOFF"):
            assert np.array_equal(m1.get(name, 1), m2.get(name, 1)), (a, name)
        assert compare_routing(m2, sim, RTOL, ATOL) == [], a
    m1.close()
    m2.close()


@pytest.mark.parametrize("n_steps,S,window", [(200, 24, 8), (150, 7, 16), (96, 24, 64)])
def test_graph_two_streams_match_eager(n_steps, S, window):
    """A captured run routes on a stream of its own; with a ring shorter than
    the run (the first two cases) cell launches also wait for the routing
    that last reads the rows they overwrite.  Graph, eager and oracle agree
    on every replay."""
    _log(f"case: graph two streams n_steps={n_steps} S={S} window={window}")
    dom = synth.make_domain(30, 26, seed=22)
    m1, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=2 * n_steps, seed=9)
    m2, _, _, _ = make_pair([dom], n_sets=2, n_steps=2 * n_steps, seed=9)
    for a in (0, n_steps):
        kw = dict(mode=1, steps_per_launch=S, route_mode=3, route_window=window)
        g1 = m1.run(steps[a:a + n_steps], **kw)
        g2 = m2.run(steps[a:a + n_steps], use_graph=True, **kw)
        go = _oracle_series(sim, recs[a:a + n_steps])
        assert np.array_equal(g1, g2), a
        assert rel_close(g2, go, RTOL, ATOL)[0], a
        assert compare_routing(m2, sim, RTOL, ATOL) == [], a
    m1.close()
    m2.close()


def test_profiling_inside_graph_replays():
    """Kernel timing recorded as graph nodes: every replay adds its launches."""
    _log("case: graph profiling")
    dom = synth.make_domain(40, 40, seed=13)
    m1, _, steps, _ = make_pair([dom], n_sets=1, n_steps=48, seed=6)
    m2, _, _, _ = make_pair([dom], n_sets=1, n_steps=48, seed=6)
    m2.set_profiling(True, 3)
    for a in (0, 24):
        for rm in (1, 3):
            g1 = m1.run(steps[a:a + 12], mode=1, steps_per_launch=6, route_mode=rm)
            g2 = m2.run(steps[a:a + 12], mode=1, steps_per_launch=6, route_mode=rm, use_graph=True)
            assert np.array_equal(g1, g2)
            a += 12
    p = m2.profile()
    # one launch in 3 of each kernel is timed, chosen when a graph is
    # captured; the context keeps one graph, so the alternating routing modes
    # recapture on every call: fused launches 0, 3 and 6 of the eight
    assert p["cell_fused"][1] == 3 and p["cell_fused"][0] > 0.0
    assert p["cell_fused"][2] == pytest.approx(3 * 6 * dom.n_cells)
    assert p["route_diag"][1] > 0 and p["route_diag"][0] > 0.0
    # routing units are a call's node-steps, scaled to the timed launches
    assert p["route_tile"][1] > 0 and p["route_tile"][0] > 0.0
    assert 0.0 < p["route_tile"][2] <= 24 * dom.n_cells
    m1.close()
    m2.close()


@pytest.mark.parametrize("route_mode", [1, 3])
def test_packed_domains_match_oracle(route_mode):
    """Uneven basins with their own parameters packed in one context."""
    _log(f"case: packed domains route_mode={route_mode}")
    doms = [synth.make_domain(12, 20, seed=21), synth.make_domain(30, 30, seed=22),
            synth.make_domain(5, 7, seed=23)]
    model, sim, steps, recs = make_pair(doms, n_sets=1, n_steps=40, seed=9)
    g = model.run(steps, mode=1, steps_per_launch=16, route_mode=route_mode)
    go = _oracle_series(sim, recs)
    assert rel_close(g, go, RTOL, ATOL)[0]
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []
    model.close()


@pytest.mark.parametrize("route_mode", [0, 1, 3])
def test_single_node_and_tiny_domains_match_oracle(route_mode):
    """Ragged pack with a one-cell basin (mRM_routing's single-node shortcut,
    Qmod = qOUT with qTIN untouched, mo_mrm_routing.f90:285-289) and a
    two-cell one (one link) beside a larger basin.  route_mode 0: the
    per-step entry points."""
    _log(f"case: tiny domains route_mode={route_mode}")
    doms = [synth.make_domain(1, 1, seed=41), synth.make_domain(6, 5, seed=42),
            synth.make_domain(1, 2, seed=43), synth.make_domain(1, 1, seed=44)]
    model, sim, steps, recs = make_pair(doms, n_sets=2, n_steps=30, seed=12)
    if route_mode == 0:
        for k in range(len(recs)):
            model.mhm_step(steps[k])
            g = model.mrm_step(steps[k])
            go = sim.step(recs[k])
            ok, worst = rel_close(g, go, RTOL, ATOL)
            assert ok, (k, worst, g, go)
    else:
        g = model.run(steps, mode=1, steps_per_launch=8, route_mode=route_mode)
        go = _oracle_series(sim, recs)
        assert rel_close(g, go, RTOL, ATOL)[0]
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []
    # the lone basins' gauges carry their qOUT, not zero
    assert np.all(np.abs(g[..., 0]) > 0.0) or np.all(np.abs(go[..., 0]) == 0.0)
    model.close()


@pytest.mark.parametrize("soil,pet,timestep", [(2, 0, 1), (3, -1, 24), (4, 1, 24), (1, 0, 3)])
def test_process_variants_match_oracle(soil, pet, timestep):
    _log(f"case: variant soil={soil} pet={pet} timestep={timestep}")
    dom = synth.make_domain(10, 14, seed=31)
    n_steps = 40
    model, sim, steps, recs = make_pair([dom], n_sets=1, timestep=timestep, n_steps=n_steps,
                                        soil=soil, pet=pet, seed=4, pet_forcing=(pet <= 0))
    g = model.run(steps, mode=1, steps_per_launch=8)
    go = _oracle_series(sim, recs)
    assert rel_close(g, go, RTOL, ATOL)[0]
    assert compare_cells(model, sim, RTOL, ATOL) == []
    model.close()


@pytest.mark.parametrize("route_mode", [1, 3])
def test_restart_round_trip(route_mode, tmp_path):
    """Suspend and resume (the restart variable set: cell states and the
    routing's qTIN/qTR): a run continued after its states went through a
    file and into a fresh context equals the run that never stopped, bit
    for bit (mo_restart.f90 / mo_mrm_restart.f90 hold the same set)."""
    _log(f"case: restart round trip route_mode={route_mode}")
    dom = synth.make_domain(24, 20, seed=24)
    m1, _, steps, _ = make_pair([dom], n_sets=2, n_steps=60, seed=11)
    kw = dict(mode=1, steps_per_launch=8, route_mode=route_mode)
    m1.run(steps[:36], **kw)
    snap = m1.snapshot()
    path = tmp_path / "restart.npz"
    flat = {f"cells_{s}_{k}": v for s, c in enumerate(snap["cells"]) for k, v in c.items()}
    flat.update({f"routing_{s}_{k}": v for s, r in enumerate(snap["routing"]) for k, v in r.items()})
    np.savez(path, **flat)
    g1 = m1.run(steps[36:], **kw)
    m2, _, _, _ = make_pair([dom], n_sets=2, n_steps=60, seed=11)
    z = np.load(path)
    back = {"cells": [{k: z[f"cells_{s}_{k}"] for k in m2.STATE_NAMES} for s in range(2)],
            "routing": [{k: z[f"routing_{s}_{k}"] for k in ("qTIN", "qTR")} for s in range(2)]}
    m2.restore(back)
    g2 = m2.run(steps[36:], **kw)
    assert np.array_equal(g1, g2)
    for s in range(2):
        for name in m1.STATE_NAMES:
            assert np.array_equal(m1.get(name, s), m2.get(name, s)), (s, name)
        a, b = m1.routing_state(s), m2.routing_state(s)
        for k in ("qTIN", "qTR", "Qmod", "qOUT"):
            assert np.array_equal(a[k], b[k]), (s, k)
    m1.close()
    m2.close()


@pytest.mark.parametrize("route_mode,graph", [(1, False), (2, False), (3, False), (3, True)])
def test_qmod_sums_match_oracle(route_mode, graph):
    """Output aggregation: Qmod summed per node over the routed steps on the
    device (updateVariable, mo_nc_output.f90:140-148) equals the oracle's
    per-step Qmod summed in step order; the count is the steps routed, and
    a restart zeroes both.  Two basins packed, the second chunk of steps run
    step by step through mhm_step/mrm_step."""
    _log(f"case: qmod sums route_mode={route_mode} graph={graph}")
    doms = [synth.make_domain(14, 18, seed=25), synth.make_domain(9, 11, seed=26)]
    model, sim, steps, recs = make_pair(doms, n_sets=2, n_steps=50, seed=12)
    model.qmod_accumulate(True)
    model.run(steps[:40], mode=1, steps_per_launch=8, route_mode=route_mode, use_graph=graph)
    for k in range(40, 50):
        model.mhm_step(steps[k])
        model.mrm_step(steps[k])
    want = [np.zeros(model.n_cells) for _ in range(2)]
    for r in recs:
        sim.step(r)
        for s in range(2):
            want[s] = want[s] + sim.routing_state(s)["Qmod"]
    for s in range(2):
        got, n = model.qmod_sum(s)
        assert n == 50
        ok, worst = rel_close(got, want[s], RTOL, ATOL)
        assert ok, (s, worst)
    model.qmod_accumulate(True)
    got, n = model.qmod_sum(0)
    assert n == 0 and not got.any()
    model.close()


def _with_read_states(steps, recs, lo, hi, flag):
    """Copies of steps[lo:hi] and recs[lo:hi] with read_states = flag."""
    from mhm_amd import _native as N
    st = (N.Step * (hi - lo))()
    for i in range(hi - lo):
        C_src = steps[lo + i]
        for f, _ in N.Step._fields_:
            setattr(st[i], f, getattr(C_src, f))
        st[i].read_states = flag
    rr = recs[lo:hi].copy()
    rr["read_states"] = flag
    return st, rr


def test_restart_coefficients_follow_read_states():
    """C1/C2 loaded as a restart holds them (mrm_read_restart_states) are
    routed with while steps run with read_states = 1 (the reference skips
    reg_rout, mo_mrm_routing.f90:211) and replaced by reg_rout of the routing
    parameters at the first step with read_states = 0; both the time loop and
    the per-step path, against the oracle given the same coefficients."""
    _log("case: restart coefficients and read_states")
    dom = synth.make_domain(16, 18, seed=31)
    model, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=40, seed=13)
    kw = dict(mode=1, steps_per_launch=8, route_mode=3)
    g = model.run(steps[:20], **kw)
    go = np.stack([sim.step(r) for r in recs[:20]])
    assert rel_close(g, go, RTOL, ATOL)[0]
    nl = dom.n_links
    rng = np.random.default_rng(2)
    for s in range(2):
        # coefficients of another parameter draw (still C1 + C2 <= 1)
        c1 = rng.uniform(0.2, 0.6, nl)
        c2 = rng.uniform(0.0, 1.0 - c1)
        model.set_routing_state(None, None, s, c1, c2)
        sim.ds[0][s].C1[:nl] = c1
        sim.ds[0][s].C2[:nl] = c2
    st, rr = _with_read_states(steps, recs, 20, 30, 1)
    g = model.run(st, **kw)
    go = np.stack([sim.step(r) for r in rr])
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_routing(model, sim, RTOL, ATOL) == []
    # a call may not mix read_states values
    mixed, _ = _with_read_states(steps, recs, 30, 32, 0)
    mixed[1].read_states = 1
    with pytest.raises(RuntimeError, match="read_states"):
        model.run(mixed, **kw)
    # read_states = 0: reg_rout again, on the per-step path
    st, rr = _with_read_states(steps, recs, 30, 40, 0)
    for i in range(10):
        model.mhm_step(st[i])
        gi = model.mrm_step(st[i])
        ok, worst = rel_close(gi, sim.step(rr[i]), RTOL, ATOL)
        assert ok, (i, worst)
    assert compare_routing(model, sim, RTOL, ATOL) == []
    assert compare_cells(model, sim, RTOL, ATOL) == []
    model.close()


@pytest.mark.parametrize("how", ["fused", "graph", "plain", "per_step"])
def test_output_sums_match_oracle(how):
    """Output aggregation of the L1 states and fluxes on the device
    (mHM_updateDataset + updateVariable, mo_write_fluxes_states.f90:283-438,
    mo_nc_output.f90:140-148): every output but the neutrons, summed per cell
    over 30 steps in step order, against the oracle's per-step values of the
    same expressions summed the same way; the count is the steps added and a
    restart zeroes both.  Two basins packed, two parameter sets."""
    from oracle_sim import OUTPUT_PER_HORIZON, output_value
    _log(f"case: output sums {how}")
    doms = [synth.make_domain(12, 15, seed=41), synth.make_domain(8, 9, seed=42)]
    model, sim, steps, recs = make_pair(doms, n_sets=2, n_steps=30, seed=43)
    flags = [1] * 21
    flags[17] = 0  # neutrons
    model.output_accumulate(flags)
    if how == "per_step":
        for k in range(30):
            model.mhm_step(steps[k])
            model.mrm_step(steps[k])
    else:
        kw = {"fused": dict(mode=1, steps_per_launch=8), "plain": dict(mode=0),
              "graph": dict(mode=1, steps_per_launch=8, use_graph=True)}[how]
        model.run(steps[:12], **kw)
        model.run(steps[12:], **kw)
    keys = [(k, h) for k in range(1, 22) if k != 18
            for h in (range(model.n_horizons) if k in OUTPUT_PER_HORIZON else (0,))]
    want = {(s, k, h): np.zeros(model.n_cells) for s in range(2) for k, h in keys}
    for r in recs:
        sim.step(r)
        for s in range(2):
            for k, h in keys:
                want[(s, k, h)] = want[(s, k, h)] + output_value(sim, s, k, h)
    for s in range(2):
        for k, h in keys:
            got, n = model.output_sum(k, h + 1, s)
            assert n == 30
            ok, worst = rel_close(got, want[(s, k, h)], RTOL, ATOL)
            assert ok, (s, k, h, worst)
    model.output_accumulate(flags)
    got, n = model.output_sum(11, 1, 0)
    assert n == 0 and not got.any()
    with pytest.raises(RuntimeError, match="neutrons"):
        model.output_accumulate([1] * 21)
    model.close()


@pytest.mark.parametrize("pet,weights,how", [
    (2, False, "fused"), (2, False, "plain"), (2, False, "per_step"),
    (3, False, "fused"), (3, False, "graph"), (3, False, "per_step"),
    (1, True, "fused"), (1, True, "plain"), (3, True, "fused"), (3, True, "per_step")])
def test_radiation_pet_and_meteo_weights_match_oracle(pet, weights, how):
    """Priestley-Taylor (process_pet 2) and Penman-Monteith (3) on the device
    (get_corrected_pet, mo_meteo_handler.f90:1078-1095, mo_pet.f90:149-272)
    and the disaggregation by meteo weights (read_meteo_weights,
    1102-1107, 1183-1189, 1265-1270): every state, flux and gauge against
    the oracle over 30 daily-forced hourly steps crossing a day, two basins,
    two parameter sets."""
    _log(f"case: pet {pet} weights {weights} {how}")
    doms = [synth.make_domain(12, 15, seed=51), synth.make_domain(8, 9, seed=52)]
    model, sim, steps, recs = make_pair(doms, n_sets=2, n_steps=30, seed=53, pet=pet,
                                        weights=weights)
    if how == "per_step":
        g = []
        for k in range(30):
            model.mhm_step(steps[k])
            g.append(model.mrm_step(steps[k]))
        g = np.stack(g)
    else:
        kw = {"fused": dict(mode=1, steps_per_launch=8), "plain": dict(mode=0),
              "graph": dict(mode=1, steps_per_launch=8, use_graph=True)}[how]
        g = np.concatenate([model.run(steps[:12], **kw), model.run(steps[12:], **kw)])
    go = np.stack([sim.step(r) for r in recs])
    ok, worst = rel_close(g, go, RTOL, ATOL)
    assert ok, worst
    assert compare_cells(model, sim, RTOL, ATOL) == []
    assert compare_routing(model, sim, RTOL, ATOL) == []
    model.close()
