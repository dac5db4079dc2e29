"""MPR-lite on the device (mhmhip_mprlite_params): every parameter row of a
set built by one kernel from the set's gamma vector gives the run the host
path gives -- the numpy statement (parameters.mpr_lite) uploaded field by
field through mhmhip_set_cell_var -- bit for bit in every state and flux,
for the soil-moisture and PET cases whose rows differ, with inputs given
before and after a network binding moves the cells; and the run matches the
oracle.  The host's root fractions take c**(z/10) from the chain's own power
(the host build of chain.h in libmhmcheck.so), the device kernel's power,
instead of numpy's: every other field is the numpy expression itself."""
import ctypes as C
import os

import numpy as np
import pytest

from mhm_amd import calendar, engine, parameters as PM, synth
from oracle_sim import OracleSim, STATE_NAMES, FLUX_NAMES, rel_close
from pairs import START, compare_cells

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _chain_pow():
    lib = C.CDLL(os.path.join(ROOT, "tests", "native", "libmhmcheck.so"))
    f = lib.mhmhip_check_eval_host
    f.restype = C.c_int
    f.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 3 + [C.c_int64]

    def power(x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(x)
        p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        assert f(2, p(x), p(y), p(out), x.shape[0]) == 0  # exp_ylogx(y, x)
        return out
    return power


def _host_params(dom, g, pet, power):
    """parameters.mpr_lite's fields with the root fractions' power from the
    chain (the rest exactly as cell_params gives them)."""
    p = synth.cell_params(dom, g, pet_case=pet)
    fF, fI, fP = dom.fractions
    G = lambda k: g[PM.INDEX[k]]
    c = (fF * G("rootFractionCoefficient_forest") + fI * G("rootFractionCoefficient_impervious")
         + fP * G("rootFractionCoefficient_pervious"))
    c = np.clip(c, 1e-3, 0.999)
    hd = np.asarray(dom.horizon_depth, dtype=np.float64)
    n, nh = c.shape[0], hd.shape[0]
    cum = 1.0 - power(np.repeat(c, nh), np.tile(hd / 10.0, n)).reshape(n, nh)
    top = np.concatenate([np.zeros((n, 1)), cum[:, :-1]], axis=1)
    fr = np.maximum(cum - top, 0.0)
    p["FROOTS"] = fr / np.maximum(fr.sum(axis=1, keepdims=True), 1e-12)
    return p


def _run_pair(soil, pet, inputs_first, n_steps=30, seed=61):
    doms = [synth.make_domain(9, 8, seed=seed), synth.make_domain(6, 11, seed=seed + 1)]
    n_days = calendar.n_forcing_days_needed(n_steps)
    for i, d in enumerate(doms):
        synth.make_forcing(d, n_days, seed + 10 + i)
        if pet >= 2:
            synth.make_pet_forcing(d, seed + 20 + i)
        if pet <= 0:
            d.forcing["PET"] = np.random.default_rng(seed + 40 + i).gamma(2.0, 1.2,
                                                                        (d.n_days, d.n_cells))
    power = _chain_pow()
    gam = PM.draw(np.random.default_rng(seed + 30), 2)
    kw = dict(n_sets=2, timestep=1, process_soil=soil, process_pet=pet, route=True,
              n_forcing_days=n_days)
    host, dev = engine.Model(doms, **kw), engine.Model(doms, **kw)
    sim = OracleSim(doms, n_sets=2, timestep=1, process_soil=soil, process_pet=pet,
                    n_forcing_days=n_days, route=True)
    if inputs_first:
        # inputs and rows in creation order, then the binding moves the cells
        dev.set_mprlite_inputs()
        for s in range(2):
            dev.mprlite_params(gam[s], s)
    for m in (host, dev):
        m.bind_network()
    if not inputs_first:
        dev.set_mprlite_inputs()
    for s in range(2):
        parts = [_host_params(d, gam[s], pet, power) for d in doms]
        params = {k: np.concatenate([p[k] for p in parts], axis=0) for k in parts[0]}
        host.set_params(params, s)
        sim.set_params(params, s)
        if not inputs_first:
            dev.mprlite_params(gam[s], s)
        for i in range(len(doms)):
            for m in (host, dev, sim):
                m.set_routing_params(PM.routing_gamma(gam[s]), s, i)
    for m in (host, dev, sim):
        m.default_states()
        m.set_domain_forcing()
    if pet <= 0:
        for m in (host, dev):
            m.set_forcing("PET", np.concatenate([d.forcing["PET"] for d in doms], axis=1))
    steps = calendar.step_table(START, n_steps, 1)
    gh = host.run(steps, mode=1, steps_per_launch=10)
    gd = dev.run(steps, mode=1, steps_per_launch=10)
    go = np.stack([sim.step(r) for r in calendar.step_records(START, n_steps, 1)])
    return host, dev, sim, gh, gd, go


@pytest.mark.parametrize("soil,pet,inputs_first", [(1, 1, False), (2, 2, True), (1, 3, False),
                                                   (3, -1, True), (4, 0, False)])
def test_device_mprlite_gives_the_host_fields_run(soil, pet, inputs_first):
    print(f"case: MPR-lite soil {soil} pet {pet} inputs_first {inputs_first}", flush=True)
    host, dev, sim, gh, gd, go = _run_pair(soil, pet, inputs_first)
    assert np.array_equal(gd, gh)
    for s in range(2):
        for name in STATE_NAMES + FLUX_NAMES:
            assert np.array_equal(dev.get(name, s), host.get(name, s)), (name, s)
    ok, worst = rel_close(gd, go, 1e-10, 1e-12)
    assert ok, worst
    assert compare_cells(dev, sim) == []
    host.close()
    dev.close()


def test_device_mprlite_rejects_missing_inputs_and_bad_sets():
    print("case: MPR-lite argument checks", flush=True)
    dom = synth.make_domain(4, 5, seed=67)
    m = engine.Model([dom], n_sets=2, timestep=24, route=False, n_forcing_days=2)
    g = PM.draw(np.random.default_rng(68), 1)[0]
    with pytest.raises(RuntimeError, match="no inputs"):
        m.mprlite_params(g, 0)
    m.set_mprlite_inputs()
    with pytest.raises(RuntimeError, match="bad pointer or set"):
        m.mprlite_params(g, 2)
    with pytest.raises(ValueError):
        m.mprlite_params(g[:10], 0)
    m.mprlite_params(g, 1)
    m.close()


def test_ensemble_builder_with_device_parameter_sets():
    """workloads.plain_domain(device_params=True), as the ensemble benchmark
    builds its sets, gives the run of the host-built rows set by set."""
    print("case: ensemble builder, 8 sets x 12x10 cells, daily x 40, device parameter sets",
          flush=True)
    from mhm_amd import workloads
    power = _chain_pow()
    kw = dict(timestep=24, seed=5, n_sets=8)
    dev, steps, dom, gam = workloads.plain_domain(12, 10, 40, device_params=True, **kw)
    host, _, _, _ = workloads.plain_domain(12, 10, 40, **kw)
    for s in range(8):
        host.set_params(_host_params(dom, gam[s], 1, power), s)
    gd = dev.run(steps, mode=1, steps_per_launch=20)
    gh = host.run(steps, mode=1, steps_per_launch=20)
    assert np.array_equal(gd, gh)
    for s in (0, 7):
        for name in STATE_NAMES:
            assert np.array_equal(dev.get(name, s), host.get(name, s)), (name, s)
    dev.close()
    host.close()
