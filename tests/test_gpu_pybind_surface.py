"""The Python surface (model/run/get/set, get_variable, get_runoff,
set_meteo) driven with the call sequence of the reference's own binding
tests (pybind/tests/test_pybind.py:31-96), checked against the oracle on
the same synthetic domain instead of the network-downloaded test domain."""
import datetime as dt
import sys

import numpy as np
import pytest

import mhm_amd
from mhm_amd import calendar, parameters as PM, synth
from oracle_sim import OracleSim, rel_close

pytestmark = pytest.mark.gpu

N_DAYS = 4
START = dt.datetime(1991, 3, 1)


def _domain():
    dom = synth.make_domain(14, 17, seed=61)
    synth.make_forcing(dom, N_DAYS, 62, doy0=START.timetuple().tm_yday)
    return dom, PM.draw(np.random.default_rng(63), 1)[0]


def _oracle(dom, gam):
    sim = OracleSim([dom], n_sets=1, timestep=1, n_forcing_days=N_DAYS)
    sim.set_params(synth.cell_params(dom, gam), 0)
    sim.set_routing_params(PM.routing_gamma(gam))
    sim.default_states()
    sim.set_domain_forcing()
    recs = calendar.step_records(START, N_DAYS * 24, 1)
    q, qmod = [], []
    for r in recs:
        q.append(sim.step(r)[0])
        qmod.append(sim.routing_state(0)["Qmod"].copy())
    return np.array(q), np.array(qmod)


def test_coupling():
    print("case: pybind coupling", flush=True)
    dom, gam = _domain()
    mhm_amd.model.config_coupling(couple_case=1, meteo_expect_pre=True, meteo_expect_temp=True,
                                  meteo_expect_tmin=True, meteo_expect_tmax=True,
                                  meteo_timestep=24)
    mhm_amd.model.set_verbosity(level=1)
    mhm_amd.model.init(mhm_amd.Setup([dom], gamma=gam, start=START, n_days=N_DAYS))
    mhm_amd.model.disable_output()
    mhm_amd.run.prepare()
    mhm_amd.run.prepare_domain()
    while not mhm_amd.run.finished():
        time = dt.datetime(*mhm_amd.run.current_time())
        if time.hour % 24 == 0:
            day = (time - START).days
            f = dom.forcing
            mhm_amd.set_meteo(time=time, pre=f["PRE"][day], temp=f["TAVG"][day],
                              tmin=f["TMIN"][day], tmax=f["TMAX"][day])
        mhm_amd.run.do_time_step()
        mhm_amd.run.write_output()
    mhm_amd.run.finalize_domain()
    mhm_amd.run.finalize()
    runoff = mhm_amd.get_runoff()
    mhm_amd.model.finalize()
    q, _ = _oracle(dom, gam)
    assert runoff.shape == (N_DAYS * 24, 1)
    ok, worst = rel_close(runoff, q, 1e-10, 1e-12)
    assert ok, worst


def test_get_variable():
    print("case: pybind get_variable", flush=True)
    dom, gam = _domain()
    mhm_amd.model.set_verbosity(level=1)
    mhm_amd.model.init(mhm_amd.Setup([dom], gamma=gam, start=START, n_days=N_DAYS))
    mhm_amd.model.disable_output()
    mhm_amd.run.prepare()
    mhm_amd.run.prepare_domain()
    stream, n = 0.0, 0
    while not mhm_amd.run.finished():
        mhm_amd.run.do_time_step()
        stream = (stream * n + mhm_amd.get_variable("L11_QMOD")) / (n + 1)
        n += 1
    sm = mhm_amd.get_variable("L1_SOILMOIST", index=2, compressed=True)
    mhm_amd.run.finalize_domain()
    mhm_amd.run.finalize()
    mhm_amd.model.finalize()
    _, qmod = _oracle(dom, gam)
    mean = 0.0
    for k in range(qmod.shape[0]):
        mean = (mean * k + qmod[k]) / (k + 1)
    # reshape the oracle's packed node values like get_variable does
    want = mean.reshape((dom.ncols, dom.nrows), order="C")
    assert stream.shape == (dom.ncols, dom.nrows)
    ok, worst = rel_close(stream.filled(np.nan), want, 1e-10, 1e-12)
    assert ok, worst
    assert sm.shape == (dom.n_cells,)


def test_wrong_timestamp_is_an_error():
    print("case: pybind wrong timestamp", flush=True)
    dom, gam = _domain()
    mhm_amd.model.config_coupling(couple_case=1, meteo_expect_pre=True, meteo_expect_temp=True,
                                  meteo_expect_tmin=True, meteo_expect_tmax=True,
                                  meteo_timestep=24)
    mhm_amd.model.init(mhm_amd.Setup([dom], gamma=gam, start=START, n_days=N_DAYS))
    mhm_amd.run.prepare()
    wrong = START + dt.timedelta(days=1)
    f = dom.forcing
    mhm_amd.set_meteo(time=wrong, pre=f["PRE"][1], temp=f["TAVG"][1], tmin=f["TMIN"][1],
                      tmax=f["TMAX"][1])
    with pytest.raises(RuntimeError, match="wrong time-stamp"):
        mhm_amd.run.do_time_step()
    mhm_amd.model.finalize()
