"""Read access, as the reference's `mhm.get` module (pybind/src/wrapper.f90:246-860).

L1 and L11 variables are synchronised from the device on demand."""
import numpy as np

from . import session as _s

_L1_STATES = {"L1_INTER": "INTER", "L1_SNOWPACK": "SNOWPACK", "L1_SEALSTW": "SEALSTW",
              "L1_UNSATSTW": "UNSATSTW", "L1_SATSTW": "SATSTW", "L1_SOILMOIST": "SOILMOIST"}
_L1_FLUXES = {"L1_PET_CALC": "PET_CALC", "L1_TEMP_CALC": "TEMP_CALC",
              "L1_PREC_CALC": "PREC_CALC", "L1_AETCANOPY": "AETCANOPY",
              "L1_AETSEALED": "AETSEALED", "L1_TOTAL_RUNOFF": "TOTAL_RUNOFF",
              "L1_RUNOFFSEAL": "RUNOFFSEAL", "L1_FASTRUNOFF": "FASTRUNOFF",
              "L1_SLOWRUNOFF": "SLOWRUNOFF", "L1_BASEFLOW": "BASEFLOW", "L1_PERCOL": "PERCOL",
              "L1_PREEFFECT": "PREEFFECT", "L1_AETSOIL": "AETSOIL", "L1_INFILSOIL": "INFILSOIL",
              "L1_THROUGHFALL": "THROUGHFALL", "L1_MELT": "MELT", "L1_RAIN": "RAIN",
              "L1_SNOW": "SNOW", "L1_DEGDAY": "DEGDAY"}


def number_of_horizons():
    return _s.S.model.n_horizons


def _params():
    st = _s.S.setup
    from . import parameters as PM, synth
    if st.params is not None:
        return st.params
    gam = PM.DEFAULT if st.gamma is None else st.gamma
    parts = [synth.cell_params(d, gam) for d in st.domains]
    return {k: np.concatenate([p[k] for p in parts], axis=0) for k in parts[0]}


def L1_variable(name, idx=1):
    """get.L1_variable (wrapper.f90:569-661): values of the L1 cells, packed."""
    name = name.upper()
    m = _s.S.model
    if name in _L1_STATES or name in _L1_FLUXES:
        key = _L1_STATES.get(name) or _L1_FLUXES[name]
        v = m.get(key)
        return v[:, idx - 1].copy() if v.ndim == 2 else v
    if name == "L1_SOILMOIST_VOL":
        return m.get("SOILMOIST")[:, idx - 1] / _params()["SOILMOIST_SAT"][:, idx - 1]
    if name == "L1_SOILMOIST_VOL_ALL":
        return m.get("SOILMOIST").sum(axis=1) / _params()["SOILMOIST_SAT"].sum(axis=1)
    if name == "L1_SOILMOISTSAT":
        return _params()["SOILMOIST_SAT"][:, idx - 1].copy()
    if name == "L1_FSEALED":
        return _params()["FSEALED"].copy()
    if name == "L1_FNOTSEALED":
        return 1.0 - _params()["FSEALED"]
    raise ValueError(f"mhm.get.L1_variable: unkown variable name: {name}")


def L11_variable(name, idx=1):
    """get.L11_variable (wrapper.f90:751-777): L11_QMOD, L11_QOUT, L11_QTIN, L11_QTR."""
    name = name.upper()
    rs = _s.S.model.routing_state(0)
    table = {"L11_QMOD": "Qmod", "L11_QOUT": "qOUT", "L11_QTIN": "qTIN", "L11_QTR": "qTR"}
    if name not in table:
        raise ValueError(f"mhm.get.L11_variable: unkown variable name: {name}")
    return rs[table[name]].copy()


def runoff():
    """mRM_runoff(time, gauge) of the steps run so far."""
    return _s.S.runoff.copy()


def runoff_shape():
    return _s.S.runoff.shape


def domain_info(level="L1"):
    """ncols, nrows, ncells, xll, yll, cellsize, nodata of the first domain."""
    d = _s.S.setup.domains[0]
    return d.ncols, d.nrows, d.n_cells, 0.0, 0.0, d.cellsize, -9999.0


def domain_mask(level="L1"):
    """Fortran (nrows, ncols) mask of the first domain (all cells active)."""
    d = _s.S.setup.domains[0]
    return np.ones((d.nrows, d.ncols), dtype=bool, order="F")
