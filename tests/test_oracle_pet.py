"""Pins of the oracle's radiation-based PET and weight disaggregation
(oracle/mhm_oracle.c: or_sat_vap_pressure, or_slope_satpressure,
or_pet_priestly, or_pet_penman restating src/mHM/mo_pet.f90:149-174,
235-272, 384-441; the getter branches of get_corrected_pet, get_temp and
get_prec, src/meteo/mo_meteo_handler.f90:1078-1107, 1183-1189, 1265-1270).

Anchors that do not restate the oracle's formulas:
  * FAO-56 (Allen et al. 1998) Annex 2, Tables 2.3 and 2.4: saturation
    vapour pressure and its slope at 10, 20 and 25 degC, to the tables'
    three decimals;
  * FAO-56 Example 18 (Brussels, 6 July): the reference-crop Penman-Monteith
    evapotranspiration of 3.9 mm/d from T = 16.9 degC, u2 = 2.078 m/s,
    es - ea = 0.589 kPa, Rn = 13.28 MJ m-2 d-1 with ra = 208 / u2 and
    rs = 70 s m-1 (the reference's form differs from FAO's in its
    psychrometric constant and air density, so within 0.1 mm/d);
  * the energy conversion: 2.45 MJ m-2 d-1 (28.3565 W m-2) evaporates 1 mm;
  * identities between the two cases: Penman-Monteith without its
    aerodynamic term (ra -> infinity) is the equilibrium evaporation,
    Priestley-Taylor with alpha = 1;
  * the getter: negative net radiation gives no PET, vapour pressure in Pa
    (an actual vapour pressure equal to the saturation one, with no
    radiation, gives 0, which a missing /1000 would turn strongly
    negative), the resistance divided by the wind speed;
  * the weights: precipitation and PET weights that sum to 1 over the day
    give back the daily values, temperature weights of 1 give the daily
    temperature.
"""
import ctypes as C
import datetime as dt

import numpy as np
import pytest

import oracle_lib as O
from mhm_amd import calendar, parameters as PM, synth
from oracle_sim import OracleSim

L = O.lib()
for _f, _n in (("or_sat_vap_pressure", 1), ("or_slope_satpressure", 1), ("or_pet_priestly", 3),
               ("or_pet_penman", 7)):
    getattr(L, _f).restype = C.c_double
    getattr(L, _f).argtypes = [C.c_double] * _n

# FAO-56 Annex 2, Table 2.3 (saturation vapour pressure [kPa]) and Table 2.4
# (slope of the saturation vapour pressure curve [kPa / degC])
FAO_ES = {10.0: 1.228, 20.0: 2.338, 25.0: 3.168}
FAO_DELTA = {10.0: 0.082, 20.0: 0.145, 25.0: 0.189}
MM_PER_DAY_WM2 = 2.45e6 / 86400.0  # W m-2 that evaporate 1 mm per day
GAMMA = 0.0646  # FORCES mo_constants Psychro_dp [kPa / degC]


@pytest.mark.parametrize("t", sorted(FAO_ES))
def test_saturation_vapour_pressure_and_slope_match_fao56_tables(t):
    assert abs(L.or_sat_vap_pressure(t) - FAO_ES[t]) <= 0.0006
    assert abs(L.or_slope_satpressure(t) - FAO_DELTA[t]) <= 0.0006


def test_priestley_taylor_energy_conversion_and_scaling():
    t = 20.0
    e1 = L.or_pet_priestly(1.0, MM_PER_DAY_WM2, t)
    want = FAO_DELTA[t] / (FAO_DELTA[t] + GAMMA)  # mm/d for 1 mm/d of net radiation
    assert abs(e1 - want) <= 2e-3 * want
    assert L.or_pet_priestly(1.26, 0.0, t) == 0.0
    for a, rn in ((1.26, 150.0), (0.9, 37.5)):
        base = L.or_pet_priestly(a, rn, t)
        assert L.or_pet_priestly(2.0 * a, rn, t) == 2.0 * base
        assert L.or_pet_priestly(a, 2.0 * rn, t) == 2.0 * base


def test_penman_monteith_fao56_example_18():
    t, u2, deficit, rn = 16.9, 2.078, 0.589, 13.28e6 / 86400.0
    es = L.or_sat_vap_pressure(t)
    et = L.or_pet_penman(rn, t, es - deficit, 208.0 / u2, 70.0, 1.0, 1.0)
    assert abs(et - 3.9) <= 0.1, et


@pytest.mark.parametrize("t,rn", [(5.0, 60.0), (20.0, 150.0), (31.0, 220.0)])
def test_penman_monteith_without_aerodynamic_term_is_equilibrium_evaporation(t, rn):
    es = L.or_sat_vap_pressure(t)
    pm = L.or_pet_penman(rn, t, 0.5 * es, 1e300, 70.0, 1.0, 1.0)
    pt = L.or_pet_priestly(1.0, rn, t)
    assert abs(pm - pt) <= 1e-14 * pt


def test_penman_monteith_no_radiation_no_deficit_is_zero_and_deficit_raises_it():
    t = 18.0
    es = L.or_sat_vap_pressure(t)
    assert L.or_pet_penman(0.0, t, es, 80.0, 70.0, 1.0, 1.0) == 0.0
    lo = L.or_pet_penman(100.0, t, 0.9 * es, 80.0, 70.0, 1.0, 1.0)
    hi = L.or_pet_penman(100.0, t, 0.5 * es, 80.0, 70.0, 1.0, 1.0)
    wide = L.or_pet_penman(100.0, t, 0.5 * es, 80.0, 140.0, 1.0, 1.0)
    assert 0.0 < lo < hi and wide < hi  # drier air raises it, a closed surface lowers it


def _day_sim(pet_case, forcing, params=None, weights=None, n_steps=24, month=7):
    dom = synth.make_domain(4, 5, seed=18)
    synth.make_forcing(dom, 2, 19)
    synth.make_pet_forcing(dom, 20)
    for k, v in forcing.items():
        dom.forcing[k] = np.tile(np.asarray(v, dtype=np.float64), (2, 1))
    sim = OracleSim([dom], n_sets=1, timestep=1, process_pet=pet_case, n_forcing_days=2,
                    route=False)
    p = synth.cell_params(dom, PM.draw(np.random.default_rng(21), 1)[0], pet_case=pet_case)
    p.update(params or {})
    sim.set_params(p, 0)
    sim.default_states()
    sim.set_domain_forcing()
    if weights is not None:
        sim.set_meteo_weights(*weights)
    X = []
    for r in calendar.step_records(dt.datetime(1991, month, 10), n_steps, 1):
        sim.step(r)
        X.append(sim.ds[0][0].X[:3].copy())
    return dom, p, np.array(X)  # (steps, [pet, temp, prec], n)


def test_getter_clamps_negative_net_radiation():
    n = 20
    rn = np.where(np.arange(n) % 2 == 0, -40.0, 90.0)
    dom, p, X = _day_sim(2, {"NETRAD": rn}, n_steps=1)
    assert np.all(X[0, 0][rn < 0] == 0.0)
    assert np.all(X[0, 0][rn > 0] > 0.0)
    dom, p, X = _day_sim(3, {"NETRAD": rn}, n_steps=1)
    # no radiation: only the vapour deficit term, the same as at 0 W m-2
    dom0, p0, X0 = _day_sim(3, {"NETRAD": np.where(rn < 0, 0.0, rn)}, n_steps=1)
    assert np.array_equal(X[0, 0], X0[0, 0])


def test_getter_takes_vapour_pressure_in_pa_and_resistance_per_wind_speed():
    n = 20
    t = np.linspace(2.0, 28.0, n)
    es_pa = 1000.0 * np.array([L.or_sat_vap_pressure(v) for v in t])
    dom, p, X = _day_sim(3, {"TAVG": t, "NETRAD": np.zeros(n), "ABSVAPPRESS": es_pa}, n_steps=1)
    assert np.all(np.abs(X[0, 0]) <= 1e-15)
    ws = np.linspace(0.5, 6.0, n)
    f = {"TAVG": t, "NETRAD": np.full(n, 120.0), "ABSVAPPRESS": 0.6 * es_pa, "WINDSPEED": ws}
    dom, p, X = _day_sim(3, f, n_steps=1)
    f2 = dict(f, WINDSPEED=2.0 * ws)
    dom2, p2, X2 = _day_sim(3, f2, params={"AERORESIST": 2.0 * p["AERORESIST"]}, n_steps=1)
    assert np.array_equal(X[0, 0], X2[0, 0])


@pytest.mark.parametrize("pet_case", [1, 2, 3])
def test_weights_give_back_the_daily_values(pet_case):
    n = 20
    pre_w, temp_w, pet_w = synth.make_meteo_weights(n, 22)
    temp_one = np.ones_like(temp_w)
    dom, p, X = _day_sim(pet_case, {}, weights=(pre_w, temp_one, pet_w))
    pre = dom.forcing["PRE"][0]
    assert np.allclose(X[:, 2].sum(0), pre, rtol=1e-14, atol=1e-300)
    assert np.allclose(X[:, 1], dom.forcing["TAVG"][0][None, :], rtol=0.0, atol=1e-12)
    # the daily PET: the same cells with hourly forcing give it undivided
    sim_h = OracleSim([dom], n_sets=1, timestep=1, process_pet=pet_case, n_forcing_days=2,
                      hourly_forcing=True, route=False)
    sim_h.set_params(p, 0)
    sim_h.default_states()
    sim_h.set_domain_forcing()
    sim_h.step(calendar.step_records(dt.datetime(1991, 7, 10), 1, 1)[0])
    pet_day = sim_h.ds[0][0].X[0].copy()
    assert np.allclose(X[:, 0].sum(0), pet_day, rtol=1e-13, atol=1e-300)
    # weights of the step's month and hour: a planted month shift fails
    w_shift = np.roll(pre_w, 1, axis=1)
    dom_s, p_s, X_s = _day_sim(pet_case, {}, weights=(w_shift, temp_one, pet_w))
    assert not np.allclose(X_s[:, 2], X[:, 2], rtol=1e-6)
