"""Step clock of the time loop.

The reference advances a datetime by `timestep` hours per step
(src/common/mo_common_datetime_type.f90:106-133) and hands mHM() the
decimal Julian time newTime - 0.5, from which dec2date yields month, day and
hour (src/mHM/mo_mhm_interface_run.f90:360,396; mo_mhm.f90:440).  Here the
same month/hour/day-of-year come from an integer proleptic Gregorian
calendar.  iMeteoTS follows meteo_handler%update_timestep
(src/meteo/mo_meteo_handler.f90:607): ceiling(tt / nint(24 / nTstepForcingDay)).
"""
import datetime as _dt
import math

import numpy as np

from . import _native as N


def step_table(start, n_steps, timestep, n_tstep_forcing_day=1, read_states=False, tt0=1):
    """ctypes array of mhmhip_step for steps tt0 .. tt0+n_steps-1."""
    arr = (N.Step * n_steps)()
    meteo_div = float(round(24.0 / float(n_tstep_forcing_day)))
    for k in range(n_steps):
        tt = tt0 + k
        t = start + _dt.timedelta(hours=(tt - 1) * timestep)
        s = arr[k]
        s.tt = tt
        s.month = t.month
        s.hour = t.hour
        s.doy = t.timetuple().tm_yday
        s.iMeteoTS = int(math.ceil(float(tt) / meteo_div))
        s.read_states = 1 if read_states else 0
        s.n_steps = tt0 + n_steps - 1  # the table's last step ends the run
    return arr


def step_records(start, n_steps, timestep, n_tstep_forcing_day=1, read_states=False, tt0=1):
    """The same table as a numpy structured array (for tests and the oracle)."""
    tab = step_table(start, n_steps, timestep, n_tstep_forcing_day, read_states, tt0)
    return np.array([(s.tt, s.month, s.hour, s.doy, s.iMeteoTS, s.read_states, s.n_steps)
                     for s in tab],
                    dtype=[("tt", "i4"), ("month", "i4"), ("hour", "i4"), ("doy", "i4"),
                           ("iMeteoTS", "i4"), ("read_states", "i4"), ("n_steps", "i4")])


def n_forcing_days_needed(n_steps, n_tstep_forcing_day=1, tt0=1):
    meteo_div = float(round(24.0 / float(n_tstep_forcing_day)))
    return int(math.ceil(float(tt0 + n_steps - 1) / meteo_div))
