"""ctypes binding of the C-ABI in include/mhmhip.h (mhm_amd/libmhmhip.so).

The HIP library is the only compute path: when it is missing or cannot be
loaded, importing anything that needs it raises -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.environ.get("MHMHIP_LIB", os.path.join(PKG, "libmhmhip.so"))
SYN_LIB = os.path.join(PKG, "libmhmsyn.so")

OK, EINVAL, EHIP, ESTATE, ENOMEM = 0, -1, -2, -3, -4
PENDING = 1  # mhmhip_mrm_step: runoff accumulated, no routing at this step
MAX_HORIZONS = 6

# enum mhmhip_var
P_NAMES = ["FSEALED", "DEGDAY_INC", "DEGDAY_MAX", "DEGDAY_NOPRE", "TEMP_THRESH", "SEALED_THRESH",
           "JARVIS_C1", "KFASTFLOW", "KSLOWFLOW", "KBASEFLOW", "KPERCO", "ALPHA", "KARSTLOSS",
           "UNSAT_THRESH", "FASP", "HARSAMCOEFF", "LATITUDE", "SOILMOIST_SAT", "FROOTS",
           "SOILMOIST_FC", "WILTINGPOINT", "SOILMOIST_EXP", "MAXINTER", "PETLAICOR",
           "PRIETAYALPHA", "AERORESIST", "SURFRESIST"]
P = {n: i for i, n in enumerate(P_NAMES)}
S_NAMES = ["INTER", "SNOWPACK", "SEALSTW", "UNSATSTW", "SATSTW", "SOILMOIST"]
S = {n: 32 + i for i, n in enumerate(S_NAMES)}
X_NAMES = ["PET_CALC", "TEMP_CALC", "PREC_CALC", "THROUGHFALL", "AETCANOPY", "DEGDAY", "MELT",
           "PREEFFECT", "RAIN", "SNOW", "RUNOFFSEAL", "AETSEALED", "FASTRUNOFF", "SLOWRUNOFF",
           "PERCOL", "BASEFLOW", "TOTAL_RUNOFF", "INFILSOIL", "AETSOIL"]
X = {n: 64 + i for i, n in enumerate(X_NAMES)}
HORIZON_VARS = {"SOILMOIST_SAT", "FROOTS", "SOILMOIST_FC", "WILTINGPOINT", "SOILMOIST_EXP",
                "SOILMOIST", "INFILSOIL", "AETSOIL"}
MONTHLY_VARS = {"MAXINTER", "PETLAICOR", "PRIETAYALPHA", "AERORESIST", "SURFRESIST"}
F_NAMES = ["PRE", "TAVG", "TMIN", "TMAX", "PET", "NETRAD", "ABSVAPPRESS", "WINDSPEED"]
F = {n: i for i, n in enumerate(F_NAMES)}

RUN_PLAIN, RUN_FUSED = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class Config(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("n_domains", C.c_int32), ("n_horizons", C.c_int32),
                ("n_sets", C.c_int32), ("timestep", C.c_int32), ("process_soil", C.c_int32),
                ("process_pet", C.c_int32), ("process_route", C.c_int32),
                ("hourly_forcing", C.c_int32), ("n_forcing_days", C.c_int32),
                ("device", C.c_int32), ("domain_cell_start", _lp),
                ("evap_coeff", C.c_double * 12), ("fnight_prec", C.c_double * 12),
                ("fnight_pet", C.c_double * 12), ("fnight_temp", C.c_double * 12)]


class Step(C.Structure):
    _fields_ = [("tt", C.c_int32), ("month", C.c_int32), ("hour", C.c_int32), ("doy", C.c_int32),
                ("iMeteoTS", C.c_int32), ("read_states", C.c_int32), ("n_steps", C.c_int32)]


class Network(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("domain_node_start", _lp), ("n_outlets", _ip),
                ("netPerm", _ip), ("fromN", _ip), ("toN", _ip), ("length", _dp),
                ("slope", _dp), ("fFPimp", _dp), ("L1_area", _dp), ("L1_L11_Id", _ip),
                ("n_gauges", C.c_int32), ("n_inflow", C.c_int32), ("gauge_domain", _ip),
                ("gauge_node", _ip), ("inflow_domain", _ip), ("inflow_node", _ip),
                ("inflow_headwater", _ip)]


class RunOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("steps_per_launch", C.c_int32), ("block", C.c_int32),
                ("use_graph", C.c_int32), ("route_mode", C.c_int32),
                ("route_window", C.c_int32), ("reserved", C.c_int32 * 2)]


# every entry point declared in include/mhmhip.h
SIGNATURES = {
    "mhmhip_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "mhmhip_destroy": (C.c_int, [C.c_void_p]),
    "mhmhip_last_error": (C.c_char_p, [C.c_void_p]),
    "mhmhip_device_bytes": (C.c_int64, [C.c_void_p]),
    "mhmhip_sync": (C.c_int, [C.c_void_p]),
    "mhmhip_set_cell_var": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, C.c_int64, C.c_int32]),
    "mhmhip_get_cell_var": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, C.c_int64, C.c_int32]),
    "mhmhip_default_states": (C.c_int, [C.c_void_p, _dp]),
    "mhmhip_set_forcing": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _dp, C.c_int64]),
    "mhmhip_stream_forcing": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32,
                                        C.POINTER(_dp), C.c_int64]),
    "mhmhip_set_forcing_l2": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                        _dp, C.c_int32, C.c_int32, C.c_double, C.c_void_p,
                                        C.c_int32, C.c_int32, C.c_double, C.c_void_p]),
    "mhmhip_mhm_step": (C.c_int, [C.c_void_p, C.POINTER(Step)]),
    "mhmhip_routing_order": (C.c_int, [C.c_int64, C.c_int64, _ip, _ip, _ip, _ip]),
    "mhmhip_plan_network": (C.c_int, [C.c_int64, C.c_int32, _ip, _ip, _ip, C.c_void_p, _lp, _ip,
                                      C.c_int64]),
    "mhmhip_bind_network": (C.c_int, [C.c_void_p, C.POINTER(Network)]),
    "mhmhip_set_tile_height": (C.c_int, [C.c_void_p, C.c_int32]),
    "mhmhip_set_routing_params": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp]),
    "mhmhip_set_routing_celerity": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "mhmhip_set_routing_celerity_links": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp]),
    "mhmhip_set_inflow": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mhmhip_get_routing_state": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mhmhip_set_routing_state": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp]),
    "mhmhip_qmod_accumulate": (C.c_int, [C.c_void_p, C.c_int32]),
    "mhmhip_get_qmod_sum": (C.c_int, [C.c_void_p, C.c_int32, _dp, _lp]),
    "mhmhip_mrm_step": (C.c_int, [C.c_void_p, C.POINTER(Step), _dp]),
    "mhmhip_output_accumulate": (C.c_int, [C.c_void_p, _ip]),
    "mhmhip_get_output_sum": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _dp,
                                        C.POINTER(C.c_int64)]),
    "mhmhip_set_meteo_weights": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int64]),
    "mhmhip_mprlite_inputs": (C.c_int, [C.c_void_p, _dp, C.c_int64, _dp, _dp]),
    "mhmhip_mprlite_params": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mhmhip_run": (C.c_int, [C.c_void_p, C.POINTER(Step), C.c_int32, C.POINTER(RunOpts), _dp]),
    "mhmhip_run_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Step), C.c_int32, C.POINTER(RunOpts),
                                     _dp]),
    "mhmhip_set_profiling": (C.c_int, [C.c_void_p, C.c_int32]),
    "mhmhip_get_profile": (C.c_int, [C.c_void_p, C.c_int32, _dp, _lp, _dp]),
}
K_NAMES = ["cell_step", "cell_fused", "route_diag", "route_wg", "route_tile"]
MPRLITE_NGAMMA = 26  # gamma entries mhmhip_mprlite_params reads (the routing ones follow)


class Plan(C.Structure):
    _fields_ = [("dmax", C.c_int32), ("tau_max", C.c_int32), ("n_tiles", C.c_int64),
                ("n_bins", C.c_int64), ("n_bnd", C.c_int64), ("n_slots", C.c_int64)]


class MhmHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mhmhip error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded HIP library; raises when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB):
            raise ImportError(f"{HIP_LIB} is missing: run `python -m mhm_amd.build` first "
                              "(there is no CPU fallback)")
        L = C.CDLL(HIP_LIB)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_syn = None


def syn():
    global _syn
    if _syn is None:
        if not os.path.exists(SYN_LIB):
            raise ImportError(f"{SYN_LIB} is missing: run `python -m mhm_amd.build` first")
        L = C.CDLL(SYN_LIB)
        L.syn_priority_flood.restype = C.c_int
        L.syn_priority_flood.argtypes = [C.c_int64, C.c_int64, _dp, C.c_int64, C.c_double, _dp,
                                         _lp, _ip]
        L.syn_tree_depth.restype = C.c_int
        L.syn_tree_depth.argtypes = [C.c_int64, _lp, _ip]
        _syn = L
    return _syn


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def lptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_lp)


def check(rc, ctx=None):
    if rc != OK:
        msg = lib().mhmhip_last_error(ctx)
        raise MhmHipError(rc, msg.decode() if msg else "")
    return rc


def routing_order(from_n, to_n, n_nodes):
    """netPerm and rOrder of L11_routing_order for one domain (1-based arrays)."""
    n_links = len(from_n)
    fr = np.zeros(n_nodes, np.int32)
    to = np.zeros(n_nodes, np.int32)
    fr[:n_links] = from_n
    to[:n_links] = to_n
    r_order = np.zeros(n_nodes, np.int32)
    net_perm = np.zeros(n_nodes, np.int32)
    rc = lib().mhmhip_routing_order(n_nodes, n_links, iptr(fr), iptr(to), iptr(r_order),
                                    iptr(net_perm))
    if rc != OK:
        raise MhmHipError(rc, "routing_order: invalid network")
    return r_order, net_perm


def plan_network(domain_arrays, slots=False):
    """Device layout of one domain (mhmhip_plan_network): (plan, pos, slots)
    with slots an (n_slots, 4) int32 array when asked for.  domain_arrays:
    object with n_cells, n_outlets, netPerm, fromN, toN (1-based)."""
    d = domain_arrays
    n = int(d.n_cells)
    perm = np.ascontiguousarray(d.netPerm, np.int32)
    fr = np.ascontiguousarray(d.fromN, np.int32)
    to = np.ascontiguousarray(d.toN, np.int32)
    plan = Plan()
    pos = np.zeros(n, np.int64)
    rc = lib().mhmhip_plan_network(n, int(d.n_outlets), iptr(perm), iptr(fr), iptr(to),
                                   C.byref(plan), lptr(pos), None, 0)
    if rc != OK:
        raise MhmHipError(rc, "plan_network: invalid network")
    table = None
    if slots:
        table = np.zeros((plan.n_slots, 4), np.int32)
        rc = lib().mhmhip_plan_network(n, int(d.n_outlets), iptr(perm), iptr(fr), iptr(to),
                                       C.byref(plan), None, iptr(table), plan.n_slots)
        if rc != OK:
            raise MhmHipError(rc, "plan_network: slot table")
    return plan, pos, table
