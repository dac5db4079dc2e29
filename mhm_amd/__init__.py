"""mhm_amd -- the per-timestep hot path of mHM (meteo getters, per-cell
process chain, mRM Muskingum routing) on AMD MI355X (gfx950), FP64.

Compute runs only in the HIP library mhm_amd/libmhmhip.so (C-ABI:
include/mhmhip.h); there is no CPU fallback.  The modules `model`, `run`,
`get` and `set` and the helpers below mirror the reference's Python
surface (pybind/src/mhm/__init__.py, tools.py).
"""
import numpy as np

from . import get, model, run, set  # noqa: F401  (module names follow the reference)
from .session import Setup

__all__ = ["Setup", "model", "run", "get", "set", "get_variable", "get_runoff", "get_mask",
           "set_meteo"]


def get_runoff():
    """Runoff of all gauges, dims (time, gauge) (pybind/src/mhm/tools.py:30-38)."""
    return get.runoff()


def get_mask(level="L1", indexing="ij", selection=False):
    """Mask of a level as in tools.py:51-72 (True outside the domain unless selection)."""
    mask = np.array(get.domain_mask(level), dtype=bool)
    if not selection:
        mask = ~mask
    return mask.T if indexing == "ij" else mask


def get_variable(name, index=1, indexing="ij", compressed=False):
    """A variable of the current step, as tools.py:75-104: compressed values, or
    a masked 2-D array filled with the no-data value outside the domain."""
    name = name.upper()
    grid = name.split("_")[0].lower()
    if grid not in ("l0", "l1", "l11", "l2"):
        raise ValueError(f"Unknown variable: {name}")
    if grid == "l1":
        var = get.L1_variable(name, index)
    elif grid == "l11":
        var = get.L11_variable(name, index)
    else:
        raise ValueError(f"{grid.upper()} variables are not held by the accelerated path")
    if compressed:
        return np.asarray(var, dtype=float)
    ncols, nrows, _, _, _, _, nodata = get.domain_info(grid)
    sel = get_mask(grid, indexing="xy", selection=True).ravel(order="F")
    out = np.ma.empty_like(sel, dtype=float)
    out.fill_value = nodata
    out.mask = ~sel
    out[sel] = var[: int(sel.sum())]
    out = out.reshape((ncols, nrows), order="C")
    return out.T if indexing == "xy" else out


def set_meteo(time, pre=None, temp=None, pet=None, tmin=None, tmax=None, compressed=True,
              indexing="ij", **_other):
    """Push forcing with a time stamp (tools.py:107-169); fields are compressed
    L1 values unless compressed=False (then 2-D, masked by the L1 mask)."""
    sel = slice(None) if compressed else get_mask("L1", indexing, selection=True)
    for name, v in (("PRE", pre), ("TEMP", temp), ("PET", pet), ("TMIN", tmin), ("TMAX", tmax)):
        if v is not None:
            set.meteo(np.asarray(v)[sel], name, time.year, time.month, time.day, time.hour)
