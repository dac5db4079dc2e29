"""Module-level model state behind the `model`, `run`, `get` and `set`
modules -- the counterpart of the reference's Fortran module globals that its
pybind wrapper drives (pybind/src/wrapper.f90:13-949).

A `Setup` replaces the namelists and input files the reference reads in
`model.init` (config and readers are outside the accelerated path): it
carries the domain(s), the parameter vector, the simulation period and the
process switches.  Everything from `run.prepare()` on runs through the HIP
library.
"""
import datetime as _dt
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import calendar, engine, parameters as PM, synth


@dataclass
class Setup:
    domains: list
    gamma: np.ndarray = None              # one parameter vector (parameters.GAMMA order)
    start: _dt.datetime = _dt.datetime(1990, 1, 1)
    n_days: int = 365
    timestep: int = 1                     # [h]
    process_soil: int = 1
    process_pet: int = 1
    route: bool = True
    forcing: bool = True                  # upload each domain's synthetic forcing at prepare()
    device: int = 0
    params: dict = None                   # explicit L1 fields (overrides MPR-lite)


@dataclass
class _State:
    setup: Setup = None
    coupling: dict = field(default_factory=dict)
    verbosity: int = 1
    output: bool = True
    model: engine.Model = None
    prepared: bool = False
    domain: int = 1
    tt: int = 0
    n_steps: int = 0
    steps: object = None
    runoff: np.ndarray = None             # mRM_runoff(time, gauge)
    last_gauges: np.ndarray = None
    coupled: dict = field(default_factory=dict)  # coupled variable -> (year, month, day)


S = _State()


def require(cond, msg):
    if not cond:
        raise RuntimeError(msg)


def current_datetime():
    """Date and hour of the next step (run_cfg%domainDateTime after increment)."""
    return S.setup.start + _dt.timedelta(hours=S.tt * S.setup.timestep)


def prepare():
    """mhm_interface_run_prepare + prepare_domain: allocation, MPR, default
    states (src/mHM/mo_mhm_interface_run.f90:160-331)."""
    st = S.setup
    require(st is not None, "model.init was not called")
    n_steps = st.n_days * 24 // st.timestep
    couple = bool(S.coupling)
    n_forcing_days = 1 if couple else calendar.n_forcing_days_needed(n_steps)
    m = engine.Model(st.domains, n_sets=1, timestep=st.timestep, process_soil=st.process_soil,
                     process_pet=st.process_pet, route=st.route, n_forcing_days=n_forcing_days,
                     device=st.device)
    if st.route:
        m.bind_network()
    gam = PM.DEFAULT.copy() if st.gamma is None else np.asarray(st.gamma, dtype=np.float64)
    if st.params is not None:
        m.set_params(st.params, 0)
    else:
        parts = [synth.cell_params(d, gam) for d in st.domains]
        m.set_params({k: np.concatenate([p[k] for p in parts], axis=0) for k in parts[0]}, 0)
    if st.route:
        m.set_routing_params(PM.routing_gamma(gam))
    m.default_states()
    if not couple and st.forcing:
        for d in st.domains:
            if not d.forcing or d.n_days < n_forcing_days:
                synth.make_forcing(d, n_forcing_days, d.seed + 1000,
                                   doy0=st.start.timetuple().tm_yday)
        m.set_domain_forcing()
    S.model = m
    S.n_steps = n_steps
    S.tt = 0
    n_g = sum(d.gauge_nodes.shape[0] for d in st.domains) if st.route else 0
    S.runoff = np.zeros((n_steps, max(n_g, 1)))[:, :n_g]
    S.steps = calendar.step_table(st.start, n_steps, st.timestep, n_tstep_forcing_day=1)
    if couple:
        for k in range(n_steps):  # coupled forcing lives in column 1 (mo_meteo_handler.f90:605)
            S.steps[k].iMeteoTS = 1
    S.prepared = True


def do_time_step():
    """mhm_interface_run_do_time_step: meteo getters + mHM() + mRM_routing()
    on the device (src/mHM/mo_mhm_interface_run.f90:341-638)."""
    require(S.prepared, "run.prepare was not called")
    require(S.tt < S.n_steps, "time loop already finished")
    if S.coupling:
        # mo_meteo_handler.f90:970-1030: every expected variable must carry
        # the time stamp of the current day
        t = current_datetime()
        key = (t.year, t.month, t.day)
        for var, expected in S.coupling.items():
            require(not expected or S.coupled.get(var) == key,
                    f"meteo_handler: {var} was expected from coupler, but has a wrong time-stamp.")
    st = S.steps[S.tt]
    S.model.mhm_step(st)
    if S.setup.route:
        g = S.model.mrm_step(st)
        if g is None:
            # a routing step longer than the model step: runoff accumulates
            S.pending = getattr(S, "pending", 0) + 1
        else:
            # the routed discharge also stands for the period's earlier steps
            # (mo_mhm_interface_run.f90:614-618)
            k = getattr(S, "pending", 0)
            S.runoff[S.tt - k:S.tt + 1, :] = g[0]
            S.pending = 0
            S.last_gauges = g[0]
    S.tt += 1


def finished():
    return S.tt >= S.n_steps


def finalize_run():
    if S.model is not None:
        S.model.close()
    S.model = None
    S.prepared = False
