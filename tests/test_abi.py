"""The C-ABI library loads and exports every entry point include/mhmhip.h
declares; host-only entry points behave as documented without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mhm_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mhmhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|int64_t|const char \*)\s*\**\s*(mhmhip_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    assert "mhmhip_create" in names and "mhmhip_run" in names
    assert len(names) == len(N.SIGNATURES), sorted(set(names) ^ set(N.SIGNATURES))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(N.HIP_LIB)
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) == set(N.SIGNATURES)


def test_create_rejects_bad_config_without_touching_a_device():
    L = N.lib()
    cfg = N.Config()
    cfg.n_cells = 0  # invalid
    ctx = C.c_void_p()
    assert L.mhmhip_create(C.byref(cfg), C.byref(ctx)) == N.EINVAL
    assert b"out of range" in L.mhmhip_last_error(None)
    cfg.n_cells, cfg.n_horizons, cfg.n_sets, cfg.timestep = 10, 3, 1, 5  # 24 % 5 != 0
    cfg.process_soil, cfg.process_pet, cfg.n_forcing_days, cfg.n_domains = 1, 1, 1, 1
    assert L.mhmhip_create(C.byref(cfg), C.byref(ctx)) == N.EINVAL
    cfg.timestep = 1
    for m in range(12):
        cfg.evap_coeff[m] = 1.0
    cfg.evap_coeff[4] = 0.0  # a zero evaporation coefficient is rejected before any device call
    assert L.mhmhip_create(C.byref(cfg), C.byref(ctx)) == N.EINVAL
    assert b"evap_coeff" in L.mhmhip_last_error(None)


def test_routing_order_rejects_cycles_and_double_outlinks():
    fr = np.array([1, 2, 3, 0], np.int32)
    to = np.array([2, 3, 1, 0], np.int32)  # 1->2->3->1
    ro = np.zeros(4, np.int32)
    pm = np.zeros(4, np.int32)
    assert N.lib().mhmhip_routing_order(4, 3, N.iptr(fr), N.iptr(to), N.iptr(ro), N.iptr(pm)) == N.EINVAL
    fr = np.array([1, 1, 0], np.int32)
    to = np.array([2, 3, 0], np.int32)
    assert N.lib().mhmhip_routing_order(3, 2, N.iptr(fr), N.iptr(to), N.iptr(ro[:3]), N.iptr(pm[:3])) == N.EINVAL


def test_create_rejects_rows_beyond_32bit_offsets():
    """The kernels address (cell, set) slots by 32-bit byte offsets: a row of
    all sets must stay below 4 GiB, and create says so before any device call."""
    L = N.lib()
    cfg = N.Config()
    cfg.n_horizons, cfg.timestep, cfg.process_soil, cfg.process_pet = 3, 1, 1, 1
    cfg.n_forcing_days, cfg.n_domains = 1, 1
    for m in range(12):
        cfg.evap_coeff[m] = 1.0
    ctx = C.c_void_p()
    cfg.n_cells, cfg.n_sets = 1 << 26, 8  # 2^29 slots x 8 B = 4 GiB
    assert L.mhmhip_create(C.byref(cfg), C.byref(ctx)) == N.EINVAL
    assert b"4 GiB" in L.mhmhip_last_error(None)
    cfg.n_cells, cfg.n_sets = (1 << 26) - 64, 8  # one padded set below: passes the bound
    cfg.evap_coeff[4] = 0.0  # and fails at the next check, before any device call
    assert L.mhmhip_create(C.byref(cfg), C.byref(ctx)) == N.EINVAL
    assert b"evap_coeff" in L.mhmhip_last_error(None)


def test_mprlite_gamma_indices_follow_the_parameter_table():
    """The header's MHMHIP_MPRLITE_* gamma indices name the entries of
    parameters.GAMMA at those positions, and the count stops where the
    routing group begins."""
    from mhm_amd import parameters as PM
    txt = open(os.path.join(ROOT, "include", "mhmhip.h")).read()
    found = re.findall(r"#define MHMHIP_MPRLITE_(\w+) (\d+)\s+/\*\s*(\w+)", txt)
    named = [(int(i), name) for key, i, name in found if key != "NGAMMA"]
    assert len(named) == 26 and sorted(i for i, _ in named) == list(range(26))
    for i, name in named:
        assert PM.NAMES[i] == name, (i, name)
    n = int(re.search(r"#define MHMHIP_MPRLITE_NGAMMA (\d+)", txt).group(1))
    assert n == N.MPRLITE_NGAMMA == PM.NAMES.index(PM.ROUTING[0])
