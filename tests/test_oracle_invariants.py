"""Invariants and closed forms that pin the oracle beyond the reference's
KATs (which hold only to 1e-3).

* per-cell water balance of one meteo + mHM() step, closed in FP64:
    dI + dSnow + fS dSeal + (1-fS)(sum dTheta + dU + dG)
      = P - [Ecan + fS Eseal + (1-fS) sum aet] - q_total - (1-fS) perc (1-karst)
  (canopy mo_canopy_interc.f90:100-131, snow mo_snow_accum_melt.f90:120-156,
  soil mo_soil_moisture.f90:179-286, runoff mo_runoff.f90:117-272)
* saturated zone recession G_t = G_0 (1 - k2)^t
* special cases: Imax = 0, T > Tth with no pack, polar night, PET thresholds
* Muskingum (mo_mrm_routing.f90:445-449, reg_rout mo_mrm_mpr.f90:101-117) is
  the textbook scheme: with K, x recovered from C1, C2, the reach storage
  K (x I + (1-x) O) changes by the trapezoid of inflow minus outflow, so
  network storage closes against lateral inflow minus outlet discharge.
"""
import ctypes as C
import datetime as dt
import math

import numpy as np
import pytest

import oracle_lib as O
from mhm_amd import calendar, parameters as PM, synth
from oracle_sim import OracleSim

EPS = np.finfo(np.float64).eps


def _sim(nr=16, nc=20, seed=3, steps=240, karst=None, soil=1, timestep=1):
    dom = synth.make_domain(nr, nc, seed=seed)
    n_days = calendar.n_forcing_days_needed(steps)
    synth.make_forcing(dom, n_days, seed + 10)
    sim = OracleSim([dom], n_sets=1, timestep=timestep, process_soil=soil, n_forcing_days=n_days)
    g = PM.draw(np.random.default_rng(seed), 1)[0]
    p = synth.cell_params(dom, g)
    if karst is not None:
        p["KARSTLOSS"] = np.random.default_rng(seed + 1).uniform(karst, 1.0, dom.n_cells)
    sim.set_params(p, 0)
    sim.set_routing_params(PM.routing_gamma(g))
    sim.default_states()
    sim.set_domain_forcing()
    recs = calendar.step_records(dt.datetime(1990, 1, 1), steps, timestep)
    return dom, sim, recs, p


@pytest.mark.parametrize("soil,karst", [(1, None), (2, 0.5)])
def test_water_balance_closes_every_step(soil, karst):
    dom, sim, recs, p = _sim(soil=soil, karst=karst)
    fS = p["FSEALED"]
    ks = p["KARSTLOSS"]
    o = sim.ds[0][0]
    worst = 0.0
    for k, r in enumerate(recs):
        if k == 0:
            sim.step(r)  # soil moisture is reset to FC/2 at tt = 1
            continue
        S0 = o.S.copy()
        sim.step(r)
        S1, X = o.S, o.X
        d = S1 - S0
        nh = sim.nh
        lhs = d[0] + d[1] + fS * d[2] + (1 - fS) * (d[5:5 + nh].sum(0) + d[3] + d[4])
        P = X[2]
        et = X[4] + fS * X[11] + (1 - fS) * X[17 + nh:17 + 2 * nh].sum(0)
        rhs = P - et - X[16] - (1 - fS) * X[14] * (1 - ks)
        scale = 1.0 + np.abs(S0).sum(0)
        worst = max(worst, float(np.max(np.abs(lhs - rhs) / scale)))
    assert worst < 1e-13, worst


def test_saturated_zone_recession_closed_form():
    k2, G0, n = 0.0123, 137.5, 2000
    G = G0
    for _ in range(n):
        G, _b = O.runoff_sat_zone(k2, G)
    assert abs(G - G0 * (1 - k2) ** n) <= 1e-11 * G


def test_canopy_special_cases():
    # no interception capacity: everything falls through, nothing evaporates
    i, thr, e = O.canopy_interc(0.3, 0.0, 1.7, 0.0)
    assert (i, thr, e) == (0.0, 1.7, 0.0)
    # storage + throughfall + evaporation = old storage + precipitation
    rng = np.random.default_rng(0)
    for _ in range(200):
        pet, imax, p, i0 = rng.uniform(0, 1), rng.uniform(0.01, 3), rng.uniform(0, 5), rng.uniform(0, 1)
        i0 = min(i0, imax)
        i, thr, e = O.canopy_interc(pet, imax, p, i0)
        assert abs((i + thr + e) - (i0 + p)) <= 4 * EPS * (i0 + p + 1)
        assert 0.0 <= i <= imax and e >= 0.0


def test_snow_special_cases():
    # warm, no pack: effective precipitation is the throughfall
    pack, ddf, melt, peff, rain, snow = O.snow_accum_melt(0.5, 4.0, 2.0, 1.2, 5.0, 1.0, 0.8, 0.0)
    assert (pack, melt, peff, rain, snow) == (0.0, 0.0, 0.8, 0.8, 0.0)
    # cold: throughfall becomes pack, nothing leaves
    pack, ddf, melt, peff, rain, snow = O.snow_accum_melt(0.5, 4.0, 2.0, 1.2, -5.0, 1.0, 0.8, 3.0)
    assert (pack, melt, peff, snow) == (3.8, 0.0, 0.0, 0.8)
    # degree-day factor saturates at its maximum
    assert O.snow_accum_melt(0.5, 4.0, 2.0, 100.0, 5.0, 1.0, 0.0, 1.0)[1] == 4.0


def test_pet_special_cases():
    L = O.lib()
    assert L.or_pet_hargreaves(0.0023, 17.8, 10.0, 12.0, 12.0, 45.0, 100) == 0.0  # tmax == tmin
    assert L.or_pet_hargreaves(0.0023, 17.8, -17.8, 20.0, 0.0, 45.0, 100) == 0.0  # tavg at limit
    # polar night: the sunset hour angle clamps to 0, no extraterrestrial radiation
    assert L.or_extraterr_rad_approx(355, math.radians(80.0)) == pytest.approx(0.0, abs=1e-15)
    # equator: omega = pi/2 and Ra = const * dr * cos(delta)
    for doy in (1, 80, 172, 266):
        dr = 1 + 0.033 * math.cos(2 * math.pi * doy / 365)
        delta = 0.409 * math.sin(2 * math.pi * doy / 365 - 1.39)
        want = 86400 / math.pi / 2.45e6 * 1367 * dr * math.cos(delta)
        assert L.or_extraterr_rad_approx(doy, 0.0) == pytest.approx(want, rel=1e-14)


def _muskingum_K_x(C1, C2, TS):
    K = TS * (1.0 - C2) / C1
    x = 1.0 - (TS / C1 - TS / 2.0) / K
    return K, x


def test_reg_rout_is_textbook_muskingum():
    """C1, C2 of reg_rout equal Chow's coefficients for the clamped K, xi."""
    dom = synth.make_domain(12, 12, seed=4)
    nl = dom.n_links
    C1 = np.zeros(nl)
    C2 = np.zeros(nl)
    g = PM.routing_gamma(PM.DEFAULT)
    O.lib().or_reg_rout(O.dptr(g), nl, O.dptr(dom.length), O.dptr(dom.slope),
                        O.dptr(dom.fFPimp), 1.0, O.dptr(C1), O.dptr(C2))
    K, x = _muskingum_K_x(C1, C2, 1.0)
    D = K * (1 - x) + 0.5
    assert np.allclose((-K * x + 0.5) / D, C2, rtol=1e-12, atol=1e-14)
    assert np.allclose(1.0 / D, C1, rtol=1e-12)
    assert np.all((x >= 0.005 - 1e-12) & (x <= 0.5 + 1e-12))
    assert np.all(K <= 0.5 / x * (1 + 1e-12)) and np.all(K >= 0.5 / (1 - x) * (1 - 1e-12))


def test_routing_storage_closes():
    """Network storage change = trapezoid of (lateral inflow - outlet discharge)."""
    dom, sim, recs, _ = _sim(nr=14, nc=18, seed=9, steps=200)
    o = sim.ds[0][0]
    n = dom.n_cells
    fr = dom.fromN - 1
    outlet = dom.gauge_nodes[0] - 1
    TS = 3600.0

    def storage():
        K, x = _muskingum_K_x(o.C1[:dom.n_links], o.C2[:dom.n_links], 1.0)
        I, Ou = o.qTIN[:n][fr], o.qTR[:n][fr]
        return float(np.sum(K * 3600.0 * (x * I + (1 - x) * Ou)))

    sim.step(recs[0])
    S_prev = storage()
    bal_prev = o.qOUT.sum() - o.Qmod[outlet]
    err = 0.0
    for r in recs[1:]:
        sim.step(r)
        S = storage()
        bal = o.qOUT.sum() - o.Qmod[outlet]
        err = max(err, abs((S - S_prev) - TS / 2 * (bal + bal_prev)) / (abs(S) + 1.0))
        S_prev, bal_prev = S, bal
    assert err < 1e-12, err


def test_routing_last_link_quirk_and_sum_order():
    """Tiny network: 1 -> 3, 2 -> 3, 3 -> 5 (outlet), 4 -> 5.  Direct recursion
    over the tree gives the same qTIN/qTR as the reference's link loop, and
    the outlet receives its own qOUT only because it is the last link's tNode."""
    fromN = np.array([1, 2, 3, 4], np.int32)
    toN = np.array([3, 3, 5, 5], np.int32)
    n = 5
    ro, pm = np.zeros(n, np.int32), np.zeros(n, np.int32)
    fr = np.zeros(n, np.int32)
    to = np.zeros(n, np.int32)
    fr[:4], to[:4] = fromN, toN
    O.lib().or_routing_order(n, 4, O.iptr(fr), O.iptr(to), O.iptr(ro), O.iptr(pm))
    C1 = np.array([0.3, 0.5, 0.7, 0.2])
    C2 = np.array([0.1, -0.05, 0.2, 0.15])
    qTIN = np.zeros(2 * n)
    qTR = np.zeros(2 * n)
    Qmod = np.zeros(n)
    rng = np.random.default_rng(1)
    prev_in = np.zeros(n)
    prev_tr = np.zeros(n)
    for step in range(20):
        qout = rng.uniform(0, 3, n)
        O.lib().or_L11_routing(n, 4, O.iptr(pm), O.iptr(fr), O.iptr(to), O.dptr(C1), O.dptr(C2),
                               O.dptr(qout), 0, None, None, O.dptr(qTIN), O.dptr(qTR), O.dptr(Qmod))
        # direct recursion: leaves 1, 2, 4; node 3 gets 1 and 2; outlet 5 gets 3 and 4 + own
        tin = np.zeros(n)
        tr = np.zeros(n)
        link = {1: 0, 2: 1, 3: 2, 4: 3}
        for node, ups in ((1, []), (2, []), (4, []), (3, [1, 2])):
            s = 0.0
            for u in ups:
                s = s + tr[u - 1]
            tin[node - 1] = s + qout[node - 1]
            i = link[node]
            tr[node - 1] = prev_tr[node - 1] + C1[i] * (prev_in[node - 1] - prev_tr[node - 1]) + \
                C2[i] * (tin[node - 1] - prev_in[node - 1])
        order = [u for u in (pm[:4] - 1) if toN[u] == 5]  # upstream of 5 in routing rank
        s = 0.0
        for li in order:
            s = s + tr[fromN[li] - 1]
        tin[4] = s + qout[4]
        assert np.array_equal(Qmod, tin), step
        assert np.array_equal(qTR[:4], tr[:4]), step
        prev_in, prev_tr = tin, tr


def test_routing_order_is_topological():
    dom = synth.make_domain(33, 27, seed=17)
    rank = np.empty(dom.n_links, np.int64)
    rank[dom.netPerm - 1] = np.arange(dom.n_links)
    out_link = np.full(dom.n_cells + 1, -1)
    out_link[dom.fromN] = np.arange(dom.n_links)
    down = out_link[dom.toN]
    has = down >= 0
    assert np.all(rank[has] < rank[down[has]])


def test_step_clock():
    recs = calendar.step_records(dt.datetime(1992, 2, 28, 0), 50, 1)
    assert (recs[0]["month"], recs[0]["hour"], recs[0]["doy"], recs[0]["iMeteoTS"]) == (2, 0, 59, 1)
    assert (recs[24]["month"], recs[24]["doy"], recs[24]["iMeteoTS"]) == (2, 60, 2)  # Feb 29, 1992
    assert (recs[48]["month"], recs[48]["doy"], recs[48]["hour"]) == (3, 61, 0)
    # daily steps: the reference's iMeteoTS = ceiling(tt / nint(24 / nTstepForcingDay))
    d = calendar.step_records(dt.datetime(1990, 1, 1), 30, 24)
    assert list(d["iMeteoTS"][:26]) == [1] * 24 + [2] * 2
    assert d[29]["doy"] == 30
