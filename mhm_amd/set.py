"""Write access, as the reference's `mhm.set` module: meteo coupling
(pybind/src/wrapper.f90:906-948 -> meteo_handler%set_meteo,
src/meteo/mo_meteo_handler.f90:1462-1605)."""
import numpy as np

from . import session as _s

_VARS = {"PRE": "PRE", "TEMP": "TAVG", "PET": "PET", "TMIN": "TMIN", "TMAX": "TMAX"}


def meteo(data, var, year, month, day, hour):
    """Push one daily field (packed L1 cells) for the given time stamp."""
    S = _s.S
    if not S.coupling:
        raise RuntimeError("meteo_handler: coupling is not configured")
    var = var.upper()
    if var not in _VARS:
        raise ValueError(f"meteo_handler: unknown coupled variable {var}")
    if hour != 0:
        raise ValueError("meteo_handler: daily coupled forcing must be stamped at hour 0")
    data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(1, -1))
    S.model.set_forcing(_VARS[var], data, day0=1)
    S.coupled[_VARS[var]] = (year, month, day)
