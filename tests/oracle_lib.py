"""ctypes view of the oracle (oracle/liboracle.so) for the test-suite.

The oracle is the CPU restatement of the reference path and is test
infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it.  This module builds it on first use when the
shared object is missing (gcc is all it needs).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

N_P_SCALAR = 17
N_S_SCALAR = 5
N_X_SCALAR = 17
P_ROWS = ["FSEALED", "DEGDAY_INCR", "DEGDAY_MAX", "DEGDAY_NOPRE", "TEMP_THRESH",
          "SEALED_THRESH", "JARVIS_C1", "K0", "K1", "K2", "KP", "ALPHA", "KARST",
          "UNSAT_THRESH", "FASP", "HARSAM", "LAT"]
S_ROWS = ["INTERC", "SNOWPACK", "SEALSTW", "UNSATSTW", "SATSTW"]
X_ROWS = ["PET", "TEMP", "PREC", "THROUGHFALL", "AET_CANOPY", "DEG_DAY", "MELT",
          "PREC_EFFECT", "RAIN", "SNOW", "RUNOFF_SEALED", "AET_SEALED", "FAST", "SLOW",
          "PERC", "BASEFLOW", "TOTAL_RUNOFF"]


class CellsCfg(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("ld", C.c_int64), ("n_h", C.c_int32),
                ("process_soil", C.c_int32), ("process_pet", C.c_int32),
                ("hourly_forcing", C.c_int32), ("c2TSTu", C.c_double),
                ("nTstepDay", C.c_double), ("forcing_ld", C.c_int64),
                ("forcing_days", C.c_int32), ("pad", C.c_int32),
                ("evap_coeff", C.c_double * 12), ("fnight_prec", C.c_double * 12),
                ("fnight_pet", C.c_double * 12), ("fnight_temp", C.c_double * 12),
                ("meteo_weights", C.c_void_p)]


class StepInfo(C.Structure):
    _fields_ = [("tt", C.c_int32), ("month", C.c_int32), ("hour", C.c_int32),
                ("doy", C.c_int32), ("iMeteoTS", C.c_int32), ("read_states", C.c_int32)]


class RouteCfg(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_outlets", C.c_int32), ("timestep", C.c_int32),
                ("process_case", C.c_int32), ("read_states", C.c_int32),
                ("map_flag", C.c_int32), ("n_inflow", C.c_int32), ("n_gauges", C.c_int32),
                ("rout_loop", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        d, i, v = C.c_double, C.c_int, None
        sig = {
            "or_canopy_interc": (v, [d, d, d, _dp, _dp, _dp]),
            "or_snow_accum_melt": (v, [d, d, d, d, d, d, d] + [_dp] * 6),
            "or_feddes_et_reduction": (d, [d, d, d, d]),
            "or_jarvis_et_reduction": (d, [d, d, d, d, d]),
            "or_soil_moisture": (v, [i, i, d, d, d, d, _dp, _dp, _dp, _dp, _dp, d, d, d,
                                     _dp, _dp, _dp, _dp, _dp, _dp]),
            "or_runoff_unsat_zone": (v, [d] * 7 + [_dp] * 5),
            "or_runoff_sat_zone": (v, [d, _dp, _dp]),
            "or_L1_total_runoff": (d, [d] * 5),
            "or_LE": (i, [d, d]),
            "or_extraterr_rad_approx": (d, [i, d]),
            "or_pet_hargreaves": (d, [d, d, d, d, d, d, i]),
            "or_pet_priestly": (d, [d, d, d]),
            "or_pet_penman": (d, [d] * 7),
            "or_slope_satpressure": (d, [d]),
            "or_sat_vap_pressure": (d, [d]),
            "or_disagg_meteo_weights": (d, [d, d, d]),
            "or_disagg_flux_daynight": (d, [i, d, d, d, d]),
            "or_disagg_state_daynight": (d, [i, d, d, d, d, i]),
            "or_n_param_rows": (i, [i]),
            "or_n_state_rows": (i, [i]),
            "or_n_flux_rows": (i, [i]),
            "or_mhm_step": (v, [C.POINTER(CellsCfg), C.POINTER(StepInfo), _dp, _dp, _dp, _dp, i]),
            "or_default_states": (v, [C.c_int64, C.c_int64, i, _dp, _dp, _dp]),
            "or_reg_rout": (v, [_dp, C.c_int64, _dp, _dp, _dp, d, _dp, _dp]),
            "or_L11_runoff_acc": (v, [C.c_int64, _dp, _dp, _ip, C.c_int64, _dp, _ip,
                                      C.c_int32, C.c_int32, _dp]),
            "or_add_inflow": (v, [C.c_int32, _ip, _ip, _ip, _dp, _dp]),
            "or_L11_routing": (v, [C.c_int64, C.c_int64, _ip, _ip, _ip, _dp, _dp, _dp,
                                   C.c_int32, _ip, _ip, _dp, _dp, _dp]),
            "or_mrm_routing": (v, [C.POINTER(RouteCfg), _dp, C.c_int64, _dp, _dp, _ip, _dp,
                                   _ip, _ip, _ip, _ip, _dp, _dp, _dp, _ip, _ip, _ip, _dp,
                                   _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "or_routing_order": (v, [C.c_int64, C.c_int64, _ip, _ip, _ip, _ip]),
            "or_mrm_update_param_celerity": (v, [C.c_int64, _dp, d, C.c_int32, _dp, _dp, _dp,
                                                 _dp, _ip]),
            "or_mrm_update_param_links": (v, [C.c_int64, _dp, _dp, C.c_int32, C.c_int32, _dp,
                                              _dp, _dp, _dp, _ip]),
            "or_meteo_l2_to_l1": (i, [C.c_int32, C.c_int32, d, C.c_void_p, C.c_int32,
                                      C.c_int32, d, C.c_void_p, C.c_int32, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    """Pointer to a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def ref(x=0.0):
    return C.c_double(x)


# ---- convenience wrappers for scalar leaf functions ------------------------

def canopy_interc(pet, imax, precip, interc):
    a, b, c = ref(interc), ref(), ref()
    lib().or_canopy_interc(pet, imax, precip, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def snow_accum_melt(incr, dmax, dnp, prec, temp, tth, thr, pack):
    p = ref(pack)
    out = [ref() for _ in range(5)]
    lib().or_snow_accum_melt(incr, dmax, dnp, prec, temp, tth, thr, C.byref(p),
                             *[C.byref(o) for o in out])
    return (p.value,) + tuple(o.value for o in out)


def soil_moisture(case, fS, thrS, pet, evap, sat, froots, fc, wp, bexp, c1, ecan, peff,
                  runoff_sealed, storage_sealed, sm):
    nH = len(sat)
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (sat, froots, fc, wp, bexp)]
    smv = np.array(sm, dtype=np.float64)
    infil = np.zeros(nH)
    aet = np.zeros(nH)
    rs, ss, aes = ref(runoff_sealed), ref(storage_sealed), ref()
    lib().or_soil_moisture(case, nH, fS, thrS, pet, evap, *[dptr(x) for x in arr], c1, ecan,
                           peff, C.byref(rs), C.byref(ss), dptr(infil), dptr(smv), dptr(aet),
                           C.byref(aes))
    return rs.value, ss.value, infil, smv, aet, aes.value


def runoff_unsat_zone(k1, kp, k0, alpha, karst, pefec, thr, sat, unsat):
    s, u = ref(sat), ref(unsat)
    out = [ref() for _ in range(3)]
    lib().or_runoff_unsat_zone(k1, kp, k0, alpha, karst, pefec, thr, C.byref(s), C.byref(u),
                               *[C.byref(o) for o in out])
    return (s.value, u.value) + tuple(o.value for o in out)


def runoff_sat_zone(k2, sat):
    s, b = ref(sat), ref()
    lib().or_runoff_sat_zone(k2, C.byref(s), C.byref(b))
    return s.value, b.value


def meteo_l2_to_l1(data2, mask2, cellsize2, mask1, cellsize1):
    """The oracle's meteo_forcings_wrapper mapping.  data2: (n_days, nrows2,
    ncols2); masks (nrows, ncols) booleans.  Returns (n_days, count(mask1))
    in pack order (column-major over mask1), or None where the oracle reports
    a valid cell mapping outside the other grid."""
    data2 = np.asarray(data2, dtype=np.float64)
    nd, nr2, nc2 = data2.shape
    d2 = np.ascontiguousarray(np.transpose(data2, (0, 2, 1)))
    m2 = np.ascontiguousarray(np.asarray(mask2, dtype=bool).T.astype(np.uint8))
    m1b = np.asarray(mask1, dtype=bool)
    nr1, nc1 = m1b.shape
    m1 = np.ascontiguousarray(m1b.T.astype(np.uint8))
    out = np.zeros((nd, int(m1b.sum())))
    rc = lib().or_meteo_l2_to_l1(nr2, nc2, float(cellsize2), m2.ctypes.data, nr1, nc1,
                                 float(cellsize1), m1.ctypes.data, nd, dptr(d2), dptr(out))
    return None if rc != 0 else out


def mrm_update_param_celerity(length, celerity, timestep):
    """The oracle's adaptive routing parameters: (C1, C2, ts_rout, factor, rout_loop)."""
    length = np.ascontiguousarray(length, dtype=np.float64)
    n = length.shape[0]
    c1, c2 = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    ts, f = ref(), ref()
    loops = C.c_int32()
    lib().or_mrm_update_param_celerity(n, dptr(length), float(celerity), int(timestep),
                                       dptr(c1), dptr(c2), C.byref(ts), C.byref(f),
                                       C.byref(loops))
    return c1[:n], c2[:n], ts.value, f.value, loops.value


def mrm_update_param_links(length, celerity, timestep):
    """processCase(8) = 3: (C1, C2, ts_rout, factor, rout_loop) of a per-link celerity."""
    length = np.ascontiguousarray(length, dtype=np.float64)
    cel = np.ascontiguousarray(celerity, dtype=np.float64)
    n = length.shape[0]
    c1, c2 = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    ts, f = ref(), ref()
    loops = C.c_int32()
    lib().or_mrm_update_param_links(n, dptr(length), dptr(cel), 1, int(timestep), dptr(c1),
                                    dptr(c2), C.byref(ts), C.byref(f), C.byref(loops))
    return c1[:n], c2[:n], ts.value, f.value, loops.value
