"""Streamed forcing (mhmhip_stream_forcing): the meteo handler's chunked
reads (mo_meteo_helper.f90:283-327) as a ring of device forcing columns fed
on a copy stream while the time loop runs.  A run whose forcing arrives in
2-day chunks into 4 columns (double buffering: the next chunk uploads while
the current one runs) equals the run with all forcing resident, bit for
bit -- time loop eager and captured, and the per-step path -- and a step
whose day is not on the device is refused."""
import sys

import numpy as np
import pytest

from mhm_amd import synth
from oracle_sim import rel_close
from pairs import make_pair

pytestmark = pytest.mark.gpu

STATES = ("INTER", "SNOWPACK", "SEALSTW", "UNSATSTW", "SATSTW", "SOILMOIST")
VARS = ("PRE", "TAVG", "TMIN", "TMAX")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _push(model, dom, k, chunk):
    # every variable the domain has; the ones the PET case does not read are
    # accepted and skipped
    model.stream_forcing({v: a[k * chunk:(k + 1) * chunk] for v, a in dom.forcing.items()},
                         day0=k * chunk + 1)


@pytest.mark.parametrize("how,pet", [("fused", 1), ("graph", 1), ("plain", 1), ("per_step", 1),
                                     ("fused", 3), ("per_step", 3)])
def test_streamed_forcing_equals_resident(how, pet):
    _log(f"case: streamed forcing {how} pet {pet}")
    dom = synth.make_domain(18, 22, seed=61)
    n_days, chunk = 10, 2
    n_steps = 24 * n_days
    ref, sim, steps, recs = make_pair([dom], n_sets=2, n_steps=n_steps, seed=62, pet=pet)
    sut, _, _, _ = make_pair([dom], n_sets=2, n_steps=n_steps, seed=62, pet=pet,
                             forcing_columns=2 * chunk)
    kw = {"fused": dict(mode=1, steps_per_launch=12, route_mode=3),
          "graph": dict(mode=1, steps_per_launch=12, route_mode=3, use_graph=True),
          "plain": dict(mode=0, route_mode=3),
          "per_step": None}[how]
    n_chunks = n_days // chunk
    _push(sut, dom, 0, chunk)
    _push(sut, dom, 1, chunk)
    per = 24 * chunk
    for k in range(n_chunks):
        lo, hi = k * per, (k + 1) * per
        if kw is None:
            g = np.stack([_step(sut, steps[t]) for t in range(lo, hi)])
            g_ref = np.stack([_step(ref, steps[t]) for t in range(lo, hi)])
        else:
            g = sut.run(steps[lo:hi], **kw)
            g_ref = ref.run(steps[lo:hi], **kw)
        assert np.array_equal(g, g_ref), k
        if k + 2 < n_chunks:
            _push(sut, dom, k + 2, chunk)  # overwrites chunk k's columns
    for s in range(2):
        for name in STATES:
            assert np.array_equal(sut.get(name, s), ref.get(name, s)), (s, name)
    go = np.stack([sim.step(r) for r in recs])
    ok, worst = rel_close(g_ref, go[-per:], 1e-10, 1e-12)  # the last chunk, against the oracle
    assert ok, worst
    ref.close()
    sut.close()


def _step(model, st):
    model.mhm_step(st)
    return model.mrm_step(st)


def test_day_not_on_device_is_refused():
    _log("case: streamed forcing, missing day")
    dom = synth.make_domain(10, 12, seed=63)
    m, _, steps, _ = make_pair([dom], n_sets=1, n_steps=96, seed=64, forcing_columns=2)
    m.stream_forcing({v: dom.forcing[v][0:2] for v in VARS}, day0=1)
    m.run(steps[:48], mode=1, steps_per_launch=24)
    with pytest.raises(RuntimeError, match="not on the device"):
        m.run(steps[48:72], mode=1, steps_per_launch=24)
    m.stream_forcing({v: dom.forcing[v][2:4] for v in VARS}, day0=3)
    m.run(steps[48:96], mode=1, steps_per_launch=24)
    with pytest.raises(RuntimeError, match="fit"):
        m.stream_forcing({v: dom.forcing[v][0:3] for v in VARS}, day0=1)
    m.close()
